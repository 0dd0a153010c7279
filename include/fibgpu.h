/*
 * fibgpu.h -- C-ABI of the MI355X (gfx950) force-and-integrate library.
 *
 * This is the drop-in boundary for FIB_MD's per-step hot path.  The
 * reference has no FFI: its boundary is four Fortran module procedures with
 * implicit module state.  Each entry point below replaces one of them (cited
 * as reference src/file:line) and is bound from Fortran through
 * iso_c_binding (fortran/fibgpu_iface.f90) or from Python through ctypes.
 *
 * Conventions
 *  - Every function returns an int32 status (FIB_OK == 0); fib_last_error()
 *    gives a message.  No function prints or aborts.
 *  - Host arrays are owned by the caller, passed by pointer, never retained.
 *    Per-atom vectors are Fortran (natm,3) column-major arrays: x[0..natm),
 *    then y, then z -- the reference's X(Natm,3).  Atom indices in exported
 *    lists are 1-based, as in the reference.
 *  - Device memory is owned by the library.  Calls on one handle must come
 *    from one host thread; each handle owns one HIP stream.  Handles are
 *    independent (no global state), so several may coexist.
 *  - There is no CPU fallback: without a usable gfx950 device fib_create()
 *    returns FIB_ERR_NODEVICE.
 *  - A handle holds one or more independent instances (substrates).  The
 *    reference corresponds to one instance; batched ensembles put many
 *    instances of different sizes into the same launches.
 */
#ifndef FIBGPU_H
#define FIBGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FIBGPU_ABI_VERSION 1

enum fib_status {
  FIB_OK = 0,
  FIB_ERR_ARG = 1,       /* bad argument (size, index, null pointer)       */
  FIB_ERR_HIP = 2,       /* a HIP runtime call failed                       */
  FIB_ERR_OVERFLOW = 3,  /* a neighbor list exceeded capacity even after growth */
  FIB_ERR_TABLE = 4,     /* pair-table range miss (reference si.f90:562-564) */
  FIB_ERR_NONFINITE = 5, /* NaN/Inf detected in state                       */
  FIB_ERR_STATE = 6,     /* call order violated (e.g. force before lists)   */
  FIB_ERR_NODEVICE = 7   /* no gfx950 GPU visible                           */
};

enum fib_precision {
  FIB_PREC_FP64 = 0,  /* all force arithmetic in FP64 (reference arithmetic) */
  FIB_PREC_MIXED = 1  /* FP32 pair/triplet terms from FP64 pair and leg
                         geometry (cos + 1/3 compensated), compensated FP32
                         force sums, FP64 positions, velocities,
                         integration, energies and close Moliere pairs */
};

/* Physical constants and numerical knobs.  fib_default_params() fills the
 * reference's values (prms.f90:151-216). SI units throughout. */
typedef struct fib_params {
  /* Stillinger-Weber (prms.f90:189-201) */
  double sw_A, sw_B, sw_p, sw_q, sw_a, sw_lambda_si, sw_lambda_ge, sw_gamma;
  double eps_si, eps_ge, sigma_si, sigma_ge;
  /* masses, kg (prms.f90:178-180) */
  double mass_si, mass_ge, mass_ga;
  /* Moliere (prms.f90:175-187, setupMol :433-449) */
  double z_si, z_ge, z_ga;
  double mol_c[3], mol_b[3];
  double bohr;      /* ao = 0.053e-9 m (prms.f90:154)            */
  double kB, eps0, ec;
  /* numerics */
  double listfac;   /* neighbor-list padding factor (prms.f90:212)   */
  int32_t ntable;   /* pair-derivative table size (prms.f90:216)     */
  /* mixed precision: Moliere pairs closer than this use FP64 (m) */
  double mixed_fp64_moliere_r;
} fib_params;

void fib_default_params(fib_params *p);

/* One substrate/instance.  Atoms 1..nsg are SW (Si/Ge) atoms, nsg+1..natm are
 * Ga ions in firing order (prms.f90:541).  lb is the periodic box (m). */
typedef struct fib_instance_desc {
  int32_t natm, nsg, nlj;
  int32_t reserved;
  double lb[3];
} fib_instance_desc;

typedef struct fib_handle fib_handle;

/* Create a handle on the current HIP device.  Computes the reference's
 * derived quantities per instance: rc, Nc = MAX(1,INT(Lb/rc)), Moliere
 * cutoffs (si.f90:48-62) and the SW pair table (si.f90:575-597). */
int32_t fib_create(const fib_params *p, int32_t ninst,
                   const fib_instance_desc *inst, int32_t precision,
                   fib_handle **out);
int32_t fib_destroy(fib_handle *h);
const char *fib_last_error(const fib_handle *h);
int32_t fib_get_cells(const fib_handle *h, int32_t inst, int32_t nc[3]);

/* Upload state (reference ic / data.f90:48-112 results).  X,V (natm,3) m and
 * m/s; atype (natm) 1=Si 2=Ge 3=Ga; xsides (natm, nullable) >0 marks side
 * atoms pinned to Xi every step (tools.f90:38-51); Xi (natm,3, nullable)
 * reference positions (defaults to X).  Forces are reset to zero. */
int32_t fib_set_state(fib_handle *h, int32_t inst, const double *X,
                      const double *V, const int32_t *atype,
                      const int32_t *xsides, const double *Xi);
/* Download state in the caller's original atom order (any pointer may be
 * null).  F is the force of the last completed evaluation. */
int32_t fib_get_state(fib_handle *h, int32_t inst, double *X, double *V,
                      double *F);
/* Number of fired ions; atoms with index > nsg+impact are skipped in the
 * force (reference si.f90:415,428,457). */
int32_t fib_set_impact(fib_handle *h, int32_t inst, int32_t impact);

/* Cell assignment of every atom with the reference's arithmetic
 * (replaces atomcell, nlist.f90:20-32 as called by chainlist :44-45):
 *   Icell = INT((X - FLOOR(X*iLb)*Lb)/(Lb/N)) + 1,  1-based, (natm,3). */
int32_t fib_bin_cells(fib_handle *h, int32_t inst, int32_t *icell);

/* Rebuild the neighbor lists of every instance from current positions
 * (replaces si_nlist, si.f90:87-276): counting-sort binning, spatial
 * reordering, LDS-staged 27-cell scan.  SW pairs need both atoms non-Ga and
 * r^2 < rc^2; ion pairs need a Ga atom and r^2 < rclj^2 (si.f90:156-170). */
int32_t fib_build_nlist(fib_handle *h);

/* Export the half lists in the reference's CSR layout (si.f90:234-260):
 * nss (natm,2) 1-based [first,last] into nlist2 (a first > last range is
 * empty); nlist2 holds 1-based j > i.  Same for the ion list.  Order inside
 * one atom's range is unspecified (sets are what the reference defines).
 * Call with nlist2 == NULL first to obtain the lengths. */
int32_t fib_get_nlist(fib_handle *h, int32_t inst, int32_t *nss,
                      int32_t *nlist2, int32_t *len2, int32_t *nsslj,
                      int32_t *nlistlj, int32_t *lenlj);

/* Forces on all instances at the current positions with the current lists
 * (replaces si_force, si.f90:384-552; energies replace si_potential
 * si.f90:278-368, restricted to the pairs and triplets the force includes;
 * the virial W[a][b] = sum r_a f_b has no reference counterpart). */
enum { FIB_EVAL_FORCE = 1, FIB_EVAL_ENERGY = 2, FIB_EVAL_VIRIAL = 4 };
int32_t fib_compute(fib_handle *h, int32_t flags);
/* Results of the last fib_compute for one instance (pointers nullable). */
int32_t fib_get_forces(fib_handle *h, int32_t inst, double *F, double *u,
                       double *utot, double virial[9]);
/* Convenience: fib_set_impact + fib_compute + fib_get_forces. */
int32_t fib_force(fib_handle *h, int32_t inst, int32_t impact, double *F,
                  double *u, double *utot, double virial[9]);

/* Force (and, with utot non-null, the potential energy) of one
 * configuration without allocating device memory -- for callers that
 * cannot call hipMalloc (code under HIP graph capture, tight serving
 * loops).  It has no counterpart in the reference; it is fib_force of a
 * one-instance handle without the handle: every intermediate array lives in
 * a fixed device array of FIB_NOALLOC_BYTES that each GPU holds for the life
 * of the process.  X (natm,3) and atype (natm) are device arrays in the
 * reference's layout, F (natm,3) a device array, utot a device double or
 * NULL; stream is a hipStream_t or NULL.  At most FIB_NOALLOC_MAX_ATOMS
 * atoms (and as many cells as the array holds after them); list widths are
 * fixed (SW 32, ion 64): a longer list returns FIB_ERR_OVERFLOW.  The call
 * returns once the lists are checked; the forces follow on the stream.
 * Calls that overlap on one GPU race for the array.  fib_last_error(NULL)
 * gives the message of the calling thread's last failure. */
#define FIB_NOALLOC_BYTES (64u << 20)
#define FIB_NOALLOC_MAX_ATOMS 16384
int32_t fib_force_noalloc(const fib_params *p, int32_t precision, int32_t natm,
                          int32_t nsg, int32_t nlj, const double lb[3],
                          int32_t impact, const double *X, const int32_t *atype,
                          double *F, double *utot, void *stream);

/* ---- time integration: the reference's onestep/TI loop ----------------
 * Step control mirrors prms/timeint module state: Ts and the cadences switch
 * from the *_i to the *_r set when the largest kinetic energy drops below
 * ke_limit_1 (calc_atm, timeint.f90:277-298, ionrun prms.f90:563-584). */
typedef struct fib_run_params {
  double ts_i, ts_r;                 /* timesteps (s)                       */
  double tau_i, tau_r;               /* Berendsen time constants (s)        */
  double ttar;                       /* target temperature (K)              */
  double teps;                       /* equilibration tolerance (K)         */
  double ke_limit_1;                 /* eV                                  */
  double dti;                        /* time between impacts (s)            */
  double sidewidth;                  /* thermostat shell width (nm)         */
  double outz[4], atomz[2], ionz;    /* storage windows, units of Lb(3)/10  */
  int32_t ntlist_i, ntlist_r, atlist_i, atlist_r, dslist_i, dslist_r;
  int32_t atm_out_i, atm_out_r, res_out;
  int32_t reserved;
} fib_run_params;

int32_t fib_set_run(fib_handle *h, int32_t inst, const fib_run_params *rp);

/* Phases of the reference's TI (timeint.f90:123-155). */
enum { FIB_TI_EQUILIBRATE = 0, FIB_TI_IMPACT = 1, FIB_TI_FREE = 2 };

typedef struct fib_step_ctl {
  int32_t whichctrl;      /* FIB_TI_*                                     */
  int32_t max_steps;      /* return after at most this many onestep calls  */
  int32_t stop_on_output; /* return after a step on the atm_out/res_out cadence */
  int32_t reserved;
} fib_step_ctl;

typedef struct fib_step_report {
  int64_t lt;             /* reference step counter                        */
  double time;            /* simulated time (s)                            */
  double temp;            /* the reference's `temp` variable (K)           */
  double ts;              /* current timestep (s)                          */
  int32_t steps_done;     /* onestep calls in this fib_step                */
  int32_t ended;          /* TI loop condition reached                     */
  int32_t output_due;     /* a step on the atm_out/res_out cadence ended the
                             call (stop_on_output) or occurred in it      */
  int32_t regime;         /* 1 = fast (Ts_i), 2 = slow (Ts_r)              */
  int32_t nlist_builds;   /* rebuilds performed in this call               */
  int32_t atoms_due;      /* that step is on the atm_out cadence (writeatoms) */
} fib_step_report;

/* Begin the reference's run on instance 0: lt = lt0, time = time0, initial
 * temperature and the initial force (tint, timeint.f90:47-69). */
int32_t fib_tint_init(fib_handle *h, int64_t lt0, double time0);
/* Start a TI phase: ionrun(1) for impacts, Tau = 1e-14 for equilibration
 * (timeint.f90:85,101-103). `impact` = number of the ion being fired. */
int32_t fib_ti_begin(fib_handle *h, int32_t whichctrl, int32_t impact);
/* Place and launch ion `knc` (knock.f90:21-55 with the ranindex already
 * drawn by the host): X, V of atom nsg+knc. */
int32_t fib_knock(fib_handle *h, int32_t knc, const double xion[3],
                  const double vion[3]);
/* Run TI loop iterations (each: lt += 1, end test, onestep). */
int32_t fib_step(fib_handle *h, const fib_step_ctl *ctl, fib_step_report *rep);
/* Recompute the reference's temperature (temp.f90:183-205). */
int32_t fib_temperature(fib_handle *h, int32_t inst, double *temp);

/* ---- MD-step mode: impacts as the benchmark configurations define them ---
 * Velocity Verlet with a per-instance timestep dt = min(dt_max, dx_max /
 * |v|max), lists rebuilt when the two largest displacements since the
 * build add up to more than the list skin, atoms
 * pinned through xsides (e.g. the bottom layers), atoms leaving the slab top
 * by more than remove_height taken out (sputtered), optional Berendsen
 * shells on the x/y sides and the bottom.  Every instance of the handle
 * steps together (a batched ensemble) with its own dt; a step runs entirely
 * on the device and can be replayed as a captured HIP graph.  This mode has
 * no reference counterpart: the reference's own loop is fib_step. */
typedef struct fib_md_params {
  double dt_max;          /* s, timestep cap                                    */
  double dx_max;          /* m, largest displacement per step                   */
  double remove_height;   /* m above surface_z beyond which atoms are removed   */
  double shell_width;     /* m, Berendsen shells (0: off)                       */
  double ttar, tau;       /* shell target temperature (K), time constant (s)    */
  double displaced_r;     /* m, displacement that counts an atom as displaced   */
  int32_t use_graph;      /* replay the step as a captured HIP graph            */
  int32_t reserved;
} fib_md_params;

/* Enter MD-step mode on all instances: surface_z / bottom_z (per instance,
 * m) are the slab's top and bottom; forces of the current state, first dt. */
int32_t fib_md_begin(fib_handle *h, const fib_md_params *p, const double *surface_z,
                     const double *bottom_z);
/* Fire ion `ion` (1-based among the instance's nlj ions) from x with
 * velocity v: the ion counts as fired from now on, lists and forces are
 * refreshed and the instance's dt follows the ion's speed. */
int32_t fib_md_fire(fib_handle *h, int32_t inst, int32_t ion, const double x[3],
                    const double v[3]);
/* Replay the step as a captured HIP graph (on) or as ordinary launches
 * (off; per-kernel timers only see ordinary launches).  The trajectory is
 * the same either way, bit for bit. */
int32_t fib_md_use_graph(fib_handle *h, int32_t on);
/* Fire one ion in each of `count` distinct instances at once: inst[q],
 * ion[q], x[3q..3q+2], v[3q..3q+2].  Only the fired instances' lists are
 * rebuilt; the others keep their state bit for bit. */
int32_t fib_md_fire_many(fib_handle *h, int32_t count, const int32_t *inst,
                         const int32_t *ion, const double *x, const double *v);
/* Advance every instance by nsteps steps of its own dt.  A list build that
 * outgrows the list width halts the step on the device; the lists are then
 * grown, the build repeated and the step finished, so the trajectory is the
 * one wider lists give, bit for bit.  More than 256 atoms leaving in one
 * step returns FIB_ERR_OVERFLOW: the step is complete, but the ledger
 * misses those atoms' interaction energy. */
int32_t fib_md_run(fib_handle *h, int32_t nsteps);
/* fib_md_run in two halves, so that the host can work (write a frame) while
 * the device steps: _begin enqueues the steps and returns at once, _end
 * waits for them, recovers from a list overflow as fib_md_run does, and
 * returns its status.  Between the two, only fib_snapshot_wait may be
 * called on the handle (other calls return FIB_ERR_STATE).  No reference
 * counterpart (onestep runs on the host, timeint.f90:158-242). */
int32_t fib_md_run_begin(fib_handle *h, int32_t nsteps);
int32_t fib_md_run_end(fib_handle *h);

/* Asynchronous snapshots for the frame writers (the reference writes X and
 * V of every atom from the host, out.f90:63-100).  _begin orders, on the
 * handle's stream, a copy of instance inst's X and V (original atom order)
 * into a device staging buffer of `slot`, then a device-to-host copy of it
 * into pinned memory on a second stream, and returns at once: later steps
 * run while the copy is in flight.  _wait blocks until the slot's copy has
 * arrived and copies X, V ((natm,3) each, either may be null) out of it.
 * A slot holds one snapshot; beginning a slot whose snapshot was not
 * waited for first waits for its copy.  Not while a run is in flight. */
#define FIB_SNAPSHOT_SLOTS 2
int32_t fib_snapshot_begin(fib_handle *h, int32_t inst, int32_t slot);
int32_t fib_snapshot_wait(fib_handle *h, int32_t slot, double *X, double *V);

typedef struct fib_md_report {
  double time;            /* s, simulated time of the instance                  */
  double dt;              /* s, timestep of its next step                       */
  double ion_z;           /* m, height of its last fired ion (1e300: removed)   */
  double ke;              /* J, kinetic energy after the last step              */
  int64_t steps;          /* steps taken since fib_md_begin                     */
  int32_t removed;        /* SW atoms removed above the surface (sputter yield) */
  int32_t ions_removed;   /* ions removed (backscattered)                       */
  int32_t displaced;      /* SW atoms farther than displaced_r from their site  */
  int32_t builds;         /* list builds so far                                 */
  /* energy ledger since fib_md_begin (J): ke + potential + e_bath + e_removed
   * - e_fired stays at its starting value up to the integration error */
  double e_bath;          /* taken out by the Berendsen shells                  */
  double e_removed;       /* carried out by removed atoms: kinetic energy plus
                             interaction energy with the atoms that stay        */
  double e_fired;         /* brought in by fired ions: kinetic energy plus
                             interaction energy at the start point              */
  double dt_min;          /* s, smallest timestep taken (0 before the first)    */
} fib_md_report;
int32_t fib_md_get_report(fib_handle *h, int32_t inst, fib_md_report *r);
/* Potential energy (J) of every instance's current state, FP64 sums of the
 * handle's precision: epot[k] for instance k.  The step's forces are kept. */
int32_t fib_md_energy(fib_handle *h, double *epot);

/* ELL list widths (entries per atom): SW list (a multiple of 4, at most 64)
 * and ion list (a multiple of 4).  Lists grow by themselves; setting a width
 * reserves room up front (or shrinks it), and drops the current lists.  Not
 * in MD-step mode.  No reference counterpart: the reference sizes its lists
 * once (init_nlist, si.f90:63-74) and only prints on overflow
 * (si.f90:167-169,264-272). */
int32_t fib_set_list_width(fib_handle *h, int32_t wsw, int32_t wion);
int32_t fib_get_list_width(const fib_handle *h, int32_t *wsw, int32_t *wion);

/* Launch tiles, for autotuning: tri_ns legs per stage run of the vertex pass
 * (128 or 256), tri_batch legs per lane whose neighbour loads are in flight
 * together (2 or 4), nb_cells cells per wavefront of the list build (4, 8 or
 * 16); 0 keeps a tile at its default (256, 4, 8).  Results are the same bit
 * for bit in every tile, only the time changes.  The reference's only
 * tuning knob of this kind is its cell count, rc and Nc (si.f90:55-56). */
int32_t fib_set_tiles(fib_handle *h, int32_t tri_ns, int32_t tri_batch, int32_t nb_cells);
int32_t fib_get_tiles(const fib_handle *h, int32_t tiles[3]);

/* Device-side kernel timers (HIP events on the handle's stream around each
 * launch), accumulated since fib_timing_enable(h, mask); for benchmarks.
 * Bit k of mask turns on timer k, which covers the kernels fib_timer_name(k)
 * names; mask 0 turns timing off. */
#define FIB_NTIMERS 16
typedef struct fib_timing {
  double ms[FIB_NTIMERS];      /* accumulated device time */
  int32_t calls[FIB_NTIMERS];  /* timed launches */
} fib_timing;
/* Name of timer k, or NULL when k is not a timer. */
const char *fib_timer_name(int32_t k);
int32_t fib_timing_enable(fib_handle *h, int32_t on);
int32_t fib_get_timing(fib_handle *h, fib_timing *t);
int32_t fib_synchronize(fib_handle *h);

#ifdef __cplusplus
}
#endif
#endif /* FIBGPU_H */
