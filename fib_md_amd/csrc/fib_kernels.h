// fib_kernels.h -- host-side launchers of the gfx950 kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fibgpu.h"
#include "fib_device.h"

// Workgroup size of the per-instance reductions (k_finalize).
#define FIB_FIN_T 1024
#define FIB_LEAVE_CAP 256  // atoms that may leave the simulation in one MD step

// Per-step arguments of the integrate kernels (onestep, timeint.f90:158-242).
struct StepArgs {
  double mass[3];            // kg by type 1..3
  double ts, tau, ttar;      // timestep, Berendsen time constant, target T
  double temp;               // equilibration: the reference's `temp` variable
  double kB, ec;
  double sidewidth;          // nm
  double outz[4], atomz[2], ionz;
  int32_t mode;              // 0 equilibrate, 1 impact (side shells), 2 free
  int32_t frz;               // apply frzsptrdatms in this step
  // MD-step mode
  double dt_max, dx_max;     // dt = min(dt_max, dx_max / |v|max) per instance
  double remove_dz;          // atoms higher than surf_z + remove_dz are removed
  double shell_w;            // Berendsen shells: x/y sides and the bottom (0 = off)
  double disp_r;             // displaced-atom threshold
  double skin;               // rebuild when the two largest displacements exceed it
};

// Per-kernel timing keys (index into fib_timing.ms).
enum KKey {
  KT_BIN = 0,   // k_bin
  KT_SORT,      // scan, scatter, fixup, reorder
  KT_NBUILD,    // k_nbuild
  KT_LISTAUX,   // k_clear_counts, k_maxcnt
  KT_LEGS,      // k_legs
  KT_TRI,       // k_tri
  KT_GATHER,    // k_force3
  KT_EXTRA,     // k_extra
  KT_POS,       // k_pos
  KT_REDUCE,    // k_finalize, k_sums
  KT_IO,        // import / export / set-atom / cell-id kernels
  KT_STEP,      // one whole MD step (a graph launch or the step's launches)
  KT_N
};

// Optional timing hook: the runtime brackets launches with events.
struct KTimer {
  void *ctx;
  int (*begin)(void *ctx, int key);  // -1: not timed
  void (*end)(void *ctx, int id, int key);
};

// All device buffers of a handle (sorted order, see fib_device.h).  Neighbour
// lists and their stencil bookkeeping are atom-major (row i = [i * W, i * W + W));
// G and RA are slot-major by compact leg index ([n * NA + i]).
struct DevBufs {
  int NA;                    // total padded atoms
  int ninst;
  int nbucket;               // total buckets
  int total_cells;           // total real cells
  int Wsw, Wion;             // ELL widths
  int cur;                   // current buffer of the double-buffered arrays
  // launch tiles (0: the defaults): legs per k_tri stage run and legs per
  // lane in flight; cells per k_nbuild wavefront.  Results do not depend on
  // them, only the time.
  int tri_ns = 0, tri_tb = 0, nb_cells = 0;
  DInst *I;                  // [ninst]
  DCtl *ctl;                 // [ninst]
  double4 *X[2];
  double *V[2], *F[2], *Xi[2];
  int *inv;                  // [NA] atom_off + orig -> sorted position
  int *ibase;                // [ninst] scan offset of each instance's first bucket
  int *binst;                // [NA / FIB_BLOCK] instance of each block of FIB_BLOCK atoms
  int *bkey, *brank, *bcnt, *cstart, *perm, *bsum;
  int *cnt_sw, *nbr_sw, *cnt_ion, *nbr_ion;
  uint16_t *rk_sw;           // [NA * Wsw]: (stencil index << 8) | rank of each row entry
  uint8_t *off_sw;           // [NA * 32]: row start of each stencil cell, [27] = has a row
  void *G;                   // centre-term force on each leg, T4 [NA * Wsw]
  uint32_t *RA;              // per leg [NA * Wsw]: index into G of the partner's force
                             // on the atom (~0: none); needs NA * Wsw < 2^32
  uint64_t *amask;           // legs of the last evaluation, by list slot [NA]
  int *cnt_leg;              // legs per atom of the last evaluation [NA]
  double *Fpart, *upart;     // own-term force and energy [3*NA], [NA]
  double4 *Xb;               // positions at the last list build [NA]
  float *disp;               // |X - Xb| after the last position update [NA]
  void *EX;                  // stale-graph terms: force for list slot k, T4 [NA * Wsw]
  void *EXself;              // stale-graph terms: force on the bridge atom, T4 [NA]
  uint64_t *exmask;          // slots of EX written in the last evaluation [NA]
  double *Fout, *uout;       // force-only results [3*NA], [NA]
  double *partials;          // [2 * FIB_NRED * pstride], layout [region][slot][block]
  int max_pad;               // largest padded atom count of an instance
  double *fin_tmp;           // chunk results of big instances' reductions
                             // [ninst * (max_pad / 64 / FIB_FIN_T + 1) * FIB_NRED]
  int pstride;               // NA/64 + 1
  TabEntry *tab;             // [ntable + 3]
  double *xfer;              // staging [3*NA] for host transfers
  const KTimer *kt;          // timing hook (may be null)
  const int *gate;           // non-null: [2 + 2*ninst] build gates and the cell prefix of
                             // the gated instances, see k_md_decide
  int *halt;                 // MD-step mode: set when a gated build overflowed its
                             // width; the step's later kernels then return at once
};

namespace fibk {
void launch_bin(const DevBufs &d, hipStream_t s);
void launch_scan(const DevBufs &d, hipStream_t s);
// scatter + fixup + reorder: flips cur, or (inplace) copies back and keeps cur
void launch_sort(const DevBufs &d, hipStream_t s, bool inplace = false);
// list build + longest-list reduction (DCtl::maxcnt_*)
void launch_nbuild(const DevBufs &d, const FConst &fc, hipStream_t s);
void launch_cellid(const DevBufs &d, int inst, int32_t *dout, hipStream_t s);
// positions of the new build (Xb), displacement bookkeeping reset; count =
// false repeats a build after the lists grew (the build count stays)
void launch_mark_build(const DevBufs &d, hipStream_t s, bool count = true);
// ELL rows into a wider layout: dst[p * wnew + k] = src[p * wold + k]
void launch_copy_rows(int NA, int wold, const int *src, int wnew, int *dst, hipStream_t s);
void launch_copy_rows(int NA, int wold, const uint16_t *src, int wnew, uint16_t *dst,
                      hipStream_t s);
// handle-free evaluations: instance record, zero fill (bytes a multiple of
// 4), energy total of instance 0 into a device double
void launch_put_inst(const DevBufs &d, const DInst &in, hipStream_t s);
void launch_fill_zero(void *p, size_t bytes, hipStream_t s);
void launch_copy_utot(const DevBufs &d, double *out, hipStream_t s);
// clears the leaving mark of every atom (after a removal-list overflow)
void launch_clear_leaving(const DevBufs &d, hipStream_t s);
// lists = true: clears the list-overflow flags and the halt flag; false: the
// removal-list overflow flags
void launch_clear_overflow(const DevBufs &d, hipStream_t s, bool lists = true);
// passes 1-2 (k_legs, k_tri): own pair terms and legs, centre-term shares
// into G, energy and virial reductions
// extra = false leaves out the stale-list pass when no term can be missing
// (the MD step, whose rebuild gate keeps every pair in range listed)
void launch_force_a(const DevBufs &d, const FConst &fc, bool mixed, bool energy, bool virial,
                    hipStream_t s, bool extra = true);
// pass 3 (k_force3): gather of G; mode 0: store forces into Fout; mode 1:
// velocity update with Fout the persistent force array F (timeint.f90:231-236);
// mode 2: MD-step mode second half kick
void launch_force_b(const DevBufs &d, const FConst &fc, bool mixed, bool energy, int mode,
                    const StepArgs &sa, double *Fout, hipStream_t s);
void launch_pos(const DevBufs &d, const StepArgs &sa, hipStream_t s);
// MD-step mode: head of the step (kick, drift, pins, removal) + rebuild gate
// gate: [2 + 2*ninst] rebuild gates and cell prefix, then the removal list
// [1 + FIB_LEAVE_CAP] (count, atoms removed in the step)
void launch_md_pos(const DevBufs &d, const FConst &fc, const StepArgs &sa, int *gate,
                   hipStream_t s);
// adds the interaction energy of the atom with unsorted index q to its
// instance's energy brought in (fired ions)
void launch_fire_energy(const DevBufs &d, const FConst &fc, int q, hipStream_t s);
// MD-step mode: time, next dt, shell factor (after the MODE 2 force pass)
void launch_md_ctl(const DevBufs &d, const StepArgs &sa, hipStream_t s);
// MD-step mode: dt from the current |v|max (reset = also time, steps, factor)
void launch_md_init(const DevBufs &d, const StepArgs &sa, bool reset, hipStream_t s);
// MD-step mode: removed / displaced counts and ion height per instance
void launch_md_report(const DevBufs &d, const StepArgs &sa, hipStream_t s);
void launch_sums(const DevBufs &d, const StepArgs &sa, hipStream_t s);
void launch_import(const DevBufs &d, int inst, const DInst &in, const double *X,
                   const double *V, const int32_t *atype, const int32_t *xsides,
                   const double *Xi, hipStream_t s);
void launch_export_x(const DevBufs &d, int inst, const DInst &in, double *out, hipStream_t s);
void launch_export_vec(const DevBufs &d, int inst, const DInst &in, const double *src,
                       double *out, hipStream_t s);
void launch_export_scalar(const DevBufs &d, int inst, const DInst &in, const double *src,
                          double *out, hipStream_t s);
// sa supplies the masses for the ledger's e_in
void launch_set_atom(const DevBufs &d, int p, const double x[3], const double v[3],
                     const StepArgs &sa, hipStream_t s);
}  // namespace fibk

// Handle-free force evaluation with every array inside `arena` (device
// memory of arena_bytes); allocates nothing (fib_runtime.cpp).
int32_t fib_arena_force(const fib_params *p, int32_t precision, int32_t natm, int32_t nsg,
                        int32_t nlj, const double lb[3], int32_t impact, const double *X,
                        const int32_t *atype, double *F, double *utot, void *arena,
                        size_t arena_bytes, int32_t max_atoms, hipStream_t s);
