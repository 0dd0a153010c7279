/*
 * fib_oracle.h -- CPU restatement of the FIB_MD force-and-integrate path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (fib_md_amd/, include/,
 * fortran/) may include, link or call this code.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as
 * the checker (or as the timed CPU baseline), never as the thing measured.
 *
 * The state struct mirrors the reference's Fortran module variables
 * (prms.f90, data.f90, si.f90 module stillweb).  Every routine cites the
 * reference subroutine it restates.  All arithmetic is FP64 and follows the
 * reference's left-to-right evaluation order; the recipe compiles with
 * -ffp-contract=off so no FMA is formed (the reference's x86-64 gfortran build
 * has none either).
 *
 * Parity unpinned: the order of the steps inside orc_onestep / orc_ti_iter
 * (timeint.f90:36-242) as a sequence.  Each step is pinned on its own by the
 * tests; their integrated order is a restatement that only the compiled
 * reference could pin, and that cannot be built without an MPI module.
 *
 * Array conventions: positions/velocities/forces are Fortran (Natm,3)
 * column-major arrays, i.e. x[0..N), y[0..N), z[0..N).  Atom indices stored in
 * neighbor / triplet lists are 1-based, exactly as in the reference.
 */
#ifndef FIB_ORACLE_H
#define FIB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The 42 siga.in values in read order (prms.f90:223-275). */
typedef struct orc_siga {
  int32_t amorcrys, whichic, startimpact;
  double eV;
  int32_t nc[3];
  int32_t npc[3];
  int32_t nlj;
  int32_t nt;
  double ts_i, ts_r;
  double ttar1, ke_limit_1, ke_limit_2;
  double tau_i, tau_r, teps;
  double uo;
  int32_t ntlist_i, ntlist_r, atlist_i, atlist_r, dslist_i, dslist_r;
  int32_t atm_out_i, atm_out_r, res_out, eng_out, tmp_out;
  int32_t nrand;
  double outz[4];
  double atomz[2];
  double ionz, knockz, sidewidth, dti;
  double fwhm;
  double phiz, phixy;
} orc_siga;

typedef struct orc orc;

orc *orc_new(void);
void orc_free(orc *o);

/* --- parameters ------------------------------------------------------- */
/* prms.f90:172-216 constants + setupSiGe (:452-480) + setupMol (:433-449). */
void orc_constants(orc *o);
/* readprms after the reads (prms.f90:279-357): pureepi when amorcrys==0,
 * setupSiGe, setupMol, ionrun(1), iLb = 1/Lb.  ranseed is taken from the
 * caller (the reference draws it from the wall clock, prms.f90:220,296). */
int orc_readprms(orc *o, const orc_siga *s, int32_t ranseed);
/* Free geometry for bulk / synthetic tests: Natm atoms (Nsg SW atoms then Nlj
 * ions), box Lb.  Sets the same derived constants as readprms. */
int orc_set_geometry(orc *o, int32_t natm, int32_t nsg, int32_t nlj,
                     const double lb[3]);
void orc_ionrun(orc *o, int32_t tic);
/* Copy the run parameters of a siga record without building the lattice. */
void orc_set_run(orc *o, const orc_siga *s);

/* --- state I/O -------------------------------------------------------- */
void orc_set_state(orc *o, const double *X, const double *V,
                   const int32_t *atype, const double *mass);
void orc_get_state(const orc *o, double *X, double *V, double *F);
void orc_set_force(orc *o, const double *F);
void orc_set_xsides(orc *o, const int32_t *xsides);
void orc_set_xi(orc *o, const double *Xi);
void orc_get_xi(const orc *o, double *Xi);
void orc_get_xsides(const orc *o, int32_t *xsides);
void orc_get_types(const orc *o, int32_t *atype, double *mass);
void orc_get_xrand(const orc *o, double *xrand);
int32_t orc_get_int(const orc *o, const char *name);
double orc_get_dbl(const orc *o, const char *name);
void orc_set_int(orc *o, const char *name, int32_t v);
void orc_set_dbl(orc *o, const char *name, double v);

/* --- RNG / initial conditions ---------------------------------------- */
double orc_ran1(orc *o);                   /* ran1.f:10-36 on o->ranseed */
void orc_initranions(orc *o);              /* data.f90:263-285 (+E12.6 text) */
int orc_initdata(orc *o);                  /* data.f90:32-112, whichic=4 */
void orc_knock(orc *o, int32_t knc);       /* knock.f90:21-55 */

/* --- neighbor lists --------------------------------------------------- */
void orc_init_nlist(orc *o);               /* si.f90:48-80 */
void orc_mktables(orc *o);                 /* si.f90:575-597 */
/* atomcell via chainlist's spacing (nlist.f90:20-32,44): icell (Natm,3),
 * 1-based. */
void orc_atomcell(const orc *o, const double *X, int32_t *icell);
int orc_si_nlist(orc *o, const double *X); /* si.f90:87-276 */
/* Export: half lists after compaction.  nss (Natm,2) 1-based CSR pointers
 * into nlist2 (reference layout). Returns counts. */
int32_t orc_nlist2_len(const orc *o);
int32_t orc_nlistlj_len(const orc *o);
int32_t orc_nbrs(const orc *o);
void orc_get_nlists(const orc *o, int32_t *nss, int32_t *nlist2,
                    int32_t *nsslj, int32_t *nlistlj, int32_t *nlist3);
/* Full (pre-compaction) lists of the last build, CSR over atoms. */
void orc_get_full_counts(const orc *o, int32_t *cnt2, int32_t *cntlj);
void orc_get_full_lists(const orc *o, int32_t *full2, int32_t *fulllj);

/* --- forces / energies ------------------------------------------------ */
double orc_itable2bd(const orc *o, double ue);  /* si.f90:554-573 */
double orc_table_value(const orc *o, int32_t i1);  /* table2bd(i1) */
/* si_force (si.f90:384-552); virial (9, row-major W[a][b] = sum r_a F_b) is
 * an addition for pinning the GPU virial (nullable). */
void orc_si_force(orc *o, const double *X, double *F, double *virial);
void orc_si_potential(orc *o, const double *X, double *utot, double *u);

/* --- time integration ------------------------------------------------- */
void orc_temperature(const orc *o, double *temp);          /* temp.f90:183 */
void orc_adjusttemp(orc *o, double temp, double ttar);     /* temp.f90:160 */
void orc_controlsidetemp(orc *o, double ttar);             /* temp.f90:32 */
void orc_freezsides(orc *o);                               /* tools.f90:38 */
void orc_rebox(orc *o);                                    /* tools.f90:26 */
void orc_frzsptrdatms(orc *o);                             /* temp.f90:99 */
void orc_calc_atm(orc *o);                                 /* timeint.f90:277 */
/* onestep (timeint.f90:158-242).  Uses o->F as F (kept between calls).
 * Returns 1 if this step is an output step (iostuff cadence). */
int orc_onestep(orc *o, int32_t whichctrl, double ttar, int32_t lt);
/* tint prologue: temperature, initial si_force (timeint.f90:59-69). */
void orc_tint_init(orc *o);
void orc_ti_begin(orc *o, int32_t whichctrl);
/* One iteration of TI's loop (timeint.f90:123-155).  whichctrl 0: returns 1
 * when |temp-Ttar| <= Teps (loop ended), whichctrl 1: returns 1 when
 * time > impact*dti (lt was incremented, as in the reference). */
int orc_ti_iter(orc *o, int32_t whichctrl, double ttar);

/* --- config mode: the integrator of the benchmark configurations --------
 * No reference counterpart (the reference's own loop is onestep above).
 * Velocity Verlet with dt = min(dt_max, dx_max / |v|max) chosen after each
 * step, atoms with xsides > 0 pinned to Xi, atoms more than remove_height
 * above surface_z removed (out of every list and term from then on),
 * optional Berendsen shells on the x/y sides and the bottom, lists rebuilt
 * when the two largest displacements since the build exceed the list skin
 * rc - 0.999 a sigma, and an FP64 energy ledger.  One instance; the GPU's
 * batches are checked against single instances on the GPU itself. */
typedef struct orc_md_params {
  double dt_max, dx_max, remove_height, shell_width, ttar, tau, displaced_r;
} orc_md_params;
/* Enter config mode: lists of the current state, forces, first dt; clock,
 * step count and ledger reset. */
int orc_md_begin(orc *o, const orc_md_params *p, double surface_z, double bottom_z);
/* Fire ion `ion` (1-based among the Nlj ions) from x with velocity v. */
int orc_md_fire(orc *o, int32_t ion, const double x[3], const double v[3]);
/* One step. */
int orc_md_step(orc *o);
/* Report: out[0..15] = time, dt, lambda, ke, steps, builds, removed SW
 * atoms, removed ions, displaced SW atoms, e_bath, e_removed, e_fired,
 * height of the last fired ion (1e300 once removed), vmax, potential
 * energy of the current state, number of live atoms. */
void orc_md_report(orc *o, double *out);
void orc_md_dead(const orc *o, int32_t *dead);

#ifdef __cplusplus
}
#endif
#endif
