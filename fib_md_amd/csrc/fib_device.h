// fib_device.h -- data layout shared by the HIP kernels and the host runtime.
//
// All per-atom device arrays are indexed by the *sorted* position p (atoms
// grouped by cell, instance by instance).  Each instance owns the contiguous
// range [atom_off, atom_off + natm_pad); natm_pad is a multiple of BLOCK so
// that every thread block belongs to exactly one instance.  Slots
// [atom_off + natm, atom_off + natm_pad) hold padding atoms that never move.
//
// Positions are double4: (x, y, z, w) where w carries the 64-bit word
// (orig << 32) | meta, so one 32-byte gather yields position, species, flags
// and original index of a neighbour.
#pragma once
#include <stddef.h>
#include <stdint.h>

#define FIB_BLOCK 256
#define FIB_WAVE 64

// meta bits
#define META_TYPE_MASK 0x3u   // 1 = Si, 2 = Ge, 3 = Ga
#define META_FROZEN 0x4u      // xsides > 0: pinned to Xi every step
#define META_DEAD 0x8u        // removed from the simulation
#define META_PAD 0x10u        // padding slot
#define META_LEAVING 0x20u    // removed in the current MD step (with META_DEAD)
#define META_INST_SHIFT 8

// Per-instance geometry and bookkeeping (device copy in an array).
struct DInst {
  double lb[3], ilb[3], h[3];  // box, 1/box (reference iLb), cell size Lb/N
  int32_t nc[3];               // reference cell counts N (si.f90:57)
  int32_t next[3];             // extended grid: N+1 when N <= 2 (phantom layer), else N
  int32_t bucket_off;          // first global bucket of this instance
  int32_t nbucket;             // next0*next1*next2 + 2 (dropped, dead)
  int32_t atom_off, natm, natm_pad;
  int32_t nsg, nlj, impact;    // ions with orig >= nsg + impact are unfired
  int32_t pad_before;          // atom_off - sum of natm of earlier instances
  int32_t cell_real_off;       // prefix of N0*N1*N2 over earlier instances
  int32_t block_off;           // first thread block of this instance
  int32_t nblocks;             // natm_pad / FIB_BLOCK
  double surf_z, bot_z;        // MD-step mode: slab top and bottom heights (m)
};

// Force-field constants common to all instances (kernel argument by value).
struct FConst {
  double isig[2][2], epssig[2][2], eps[2][2], leps[2][2][2];
  double A, B, p, q, al, gamma;
  double cut_sw;      // 0.999*al: reduced-distance gate of SW terms
  double cut_sw2;     // cut_sw^2 (1 + 1e-6): FP32 pre-gate on r^2
  double r2gate;      // (1.001 u_izero)^2: pairs beyond it have no table force
                      // and no SW leg (pre-gate on r^2 / sigma^2)
  double rv2, rv2lj;  // squared list cutoffs rc^2, rclj^2 (si.f90:115-116)
  double rlj2;        // squared Moliere force cutoff (2.5 sigma)^2
  double molc[3], molb[3];
  double aamol[3][3], zmol[3];
  double ie0ec2;      // ie0*ec*ec is NOT precomputed in the reference; kept
                      // separate factors below for the FP64 path
  double ie0, ec;
  double umax;        // table abscissa range (si.f90:583)
  double tabn1;       // REAL(Ntable-1)
  int32_t ntable;
  int32_t izero;      // first table index (1-based) whose T and successors are 0
  double mixed_r64;   // mixed precision: Moliere pairs below this use FP64
  double skin;        // rc - (three-body range): pairs farther apart than rc at
                      // the build need a relative move above this to form a leg
  double mixed_xleg;  // mixed precision: three-body legs need r/sigma below
                      // this, i.e. E = exp(gamma/(r/sigma - a)) > e^-30
  float rho[2][2];    // mixed precision: (float)(isig[a][b] / isig[0][0]), the
                      // pair's length unit over sigma_Si (1 for Si-Si)
  float fsw[2][2];    // mixed precision: (float)(epssig[a][b] / (eps[0][0] isig[0][0])),
                      // the pair's eps/sigma in units of eps_Si/sigma_Si
};

// Pair-table entry for the reference's 3-point Lagrange interpolation
// (si.f90:554-573): abscissas u_i, u_{i+1}, u_{i+2} and values T_i..T_{i+2}.
struct TabEntry {
  double u0, u1, u2, t0, t1, t2;
};

// Per-instance reduction outputs / control scalars (device).  All reduced
// fields are doubles so the finalize kernel can address them by index.
struct DCtl {
  double utot;          // potential energy of the last evaluation
  double virial[9];
  double side_sum;      // sum m v^2 over side-thermostat atoms (after V update)
  double side_cnt;      // their count
  double glob_sum;      // sum m v^2 over atype < 3 (after V update)
  double temp_eq;       // equilibration: T of the rescaled V (temp.f90:183-205)
  double kemax;         // max 0.5 m v^2 / ec after the position update (eV)
  double vmax;          // max |v| (m/s) after the V update
  double ke_sum;        // sum 0.5 m v^2 (J) after the V update
  double maxcnt_sw;     // longest SW / ion neighbour list of the last build
  double maxcnt_ion;
  double eq_sum;        // sum m v^2 of the rescaled V (equilibration)
  double maxdisp;       // largest displacement since the last list build (m)
  double maxdisp2;      // second largest (MD-step mode; equal to maxdisp otherwise):
                        // a pair unlisted at the build can be within the cutoff
                        // only when maxdisp + maxdisp2 exceeds the skin
  double nmaxdisp;      // atoms at maxdisp (MD-step mode)
  double builds;        // list builds performed (count)
  // MD-step mode (per-instance adaptive dt)
  double cfg_dt;        // timestep of the next step (s)
  double cfg_time;      // simulated time (s)
  double cfg_lambda;    // Berendsen factor of the thermostat shells, next step
  double cfg_steps;     // steps taken
  double cfg_dt_min;    // smallest timestep taken since fib_md_begin (s)
  double n_removed;     // SW atoms removed above the surface (sputtered)
  double n_ga_removed;  // ions removed (backscattered)
  double n_displaced;   // SW atoms farther than the threshold from their site
  double ion_z;         // height of the instance's last ion (m)
  double ke_shell;      // sum m v^2 over thermostat-shell atoms
  double n_shell;       // their count
  // energy ledger of MD-step mode (J, cumulative since fib_md_begin): the
  // total energy plus what left minus what came in stays constant
  double e_bath;        // taken out by the Berendsen shells, sum 1/2 m v^2 (1 - lambda^2)
  double e_out;         // carried out by removed atoms, their 1/2 m v^2 at removal
  double e_in;          // brought in by fired ions, 1/2 m v^2 at firing
  double e_bath_step;   // the bath and removal terms of the last step (k_md_ctl
  double e_out_step;    // adds them in)
  int32_t nonfinite;    // set by k_bin on a NaN/Inf position
  int32_t extra_any;    // k_extra found vertex terms outside a centre's list
  int32_t overflow;     // a neighbour list exceeded its width (sticky)
  int32_t leave_overflow;  // more atoms left in one MD step than the removal
                           // list holds (sticky)
};

// Double index of a DCtl field (for the reduction finalize kernel).
#define DCTL_IDX(field) ((int)(offsetof(DCtl, field) / sizeof(double)))

// Number of doubles in a reduction partial record.
#define FIB_NRED 16

// Reduction finalize description: slot k of the producer's partials is
// combined with op[k] (0 sum, 1 max) and stored at DCtl double index dst[k].
struct RedSpec {
  int32_t nv;
  int32_t bs;            // block size of the producing kernel
  int32_t nreg;          // 1, or 2: two producers wrote partials of the same
                         // blocks at partials and partials + FIB_NRED * pstride
  int32_t op[FIB_NRED];
  int32_t dst[FIB_NRED];
};
