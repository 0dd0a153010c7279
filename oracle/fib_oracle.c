/*
 * fib_oracle.c -- CPU restatement of the FIB_MD force-and-integrate path.
 *
 * TEST INFRASTRUCTURE ONLY (see fib_oracle.h).  This is the checker against
 * which the HIP library is compared; it is never linked into the product.
 *
 * Every function names the reference subroutine (file:line under the
 * reference's src/) it restates.  Evaluation order follows the Fortran
 * source left to right; the recipe builds with -O2 -ffp-contract=off.
 * Fortran intrinsics map as: NINT -> round() (half away from zero),
 * INT -> truncation, FLOOR -> floor(), MOD(real) -> fmod(), x**2 -> x*x,
 * x**(real) -> pow().
 */
#include "fib_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NTAB 32

struct orc {
  /* ---- prms (prms.f90:20-159) ---- */
  int32_t natm, nlj, nsg, nrand, impact, ions, nt, ntable;
  int32_t npc[3];
  double ts, ts_i, ts_r;
  double lb[3], ilb[3];
  double outz[4], atomz[2], ionz, knockz, sidewidth;
  double ke_limit_1, ke_limit_2;
  int32_t atom_is_slow, atom_is_fast;
  double restart_time;
  int32_t *xsides;
  double A, B, psi, qsi, al, lambdaSi, lambdaGe, gamma_, epsSi, epsGe, sigmaSi,
      sigmaGe, massSi, massGe, massGa;
  double ZSi, ZGe, ZGa, Amol, Bmol, Cmol, alpha_mol, beta_mol, gamma_mol;
  double aamol[3][3], Zmol[3], ie0;
  double isigma[2][2], eps[2][2], leps[2][2][2], epssig[2][2];
  double Ttar1, Tau, Tau_i, Tau_r, Teps, Uo, listfac, eV, dti, fwhm, phiz,
      phixy;
  int32_t atm_out, atm_out_i, atm_out_r, res_out, eng_out, tmp_out;
  int32_t ntlist, ntlist_i, ntlist_r, atlist, atlist_i, atlist_r, dslist,
      dslist_i, dslist_r;
  int32_t ranseed, whichic, amorcrys, startimpact;
  double Pi, thrd;
  int32_t nc_lat[3];
  /* ---- data (data.f90:19-25) ---- */
  double *X, *Xi, *V, *mass, *Xrand, *F, *Fp;
  int32_t *atype;
  double *lat;  /* si.dat contents after the E20.10 round trip */
  int32_t nt0;
  /* ---- stillweb (si.f90:20-41) ---- */
  int32_t nc[3];
  double rc, rlj2, rclj, umax;
  double *table2bd, *u; /* 1-based: [1..Ntable] used, [0] unused */
  int32_t *nss, *nsslj; /* (Natm,2) column-major */
  int32_t *nlist2, *nlistlj, *nlist3;
  int64_t cap2, caplj, cap3;
  int32_t len2, lenlj, nbrs;
  int32_t *full2_cnt, *fulllj_cnt, *full2, *fulllj; /* pre-compaction copy */
  int32_t ml2, mllj, ml3;
  /* ---- ran1 SAVE state (ran1.f:16-17) ---- */
  int32_t iv[NTAB], iy;
  /* ---- timeint locals kept between TI iterations ---- */
  int32_t lt;
  double time, temp;
  /* ---- config mode (no reference counterpart): removed atoms, positions at
   * the last list build, step control and energy ledger ---- */
  int32_t *dead;  /* NULL until config mode starts; 1 = removed */
  double *Xb;
  orc_md_params md;
  double surf_z, bot_z, skin;
  double md_dt, md_lambda, md_time, md_vmax, md_ke;
  double e_bath, e_out, e_in;
  int64_t md_steps, builds;
};

static const double kB = 1.381E-23;        /* prms.f90:151 */
static const double eps0 = 8.854187E-12;   /* prms.f90:152 */
static const double ec = 1.60217646E-19;   /* prms.f90:153 */
static const double ao = 0.053E-09;        /* prms.f90:154 */

#define XA(a, i, l) (a)[(size_t)(l) * (size_t)o->natm + (size_t)(i)]

static int fnint(double x) { return (int)round(x); }

orc *orc_new(void) {
  orc *o = (orc *)calloc(1, sizeof(orc));
  return o;
}

static void free_arrays(orc *o) {
  free(o->X); free(o->Xi); free(o->V); free(o->mass); free(o->F); free(o->Fp);
  free(o->atype); free(o->xsides);
  o->X = o->Xi = o->V = o->mass = o->F = o->Fp = NULL;
  o->atype = o->xsides = NULL;
}

void orc_free(orc *o) {
  if (!o) return;
  free_arrays(o);
  free(o->Xrand); free(o->lat);
  free(o->table2bd); free(o->u);
  free(o->nss); free(o->nsslj); free(o->nlist2); free(o->nlistlj);
  free(o->nlist3);
  free(o->full2_cnt); free(o->fulllj_cnt); free(o->full2); free(o->fulllj);
  free(o->dead); free(o->Xb);
  free(o);
}

/* prms.f90:172-216 */
void orc_constants(orc *o) {
  o->Pi = 4. * atan(1.0);
  o->thrd = 1. / 3.;
  o->ZSi = 14.0; o->ZGe = 32.0; o->ZGa = 31.0;
  o->massSi = 46.637063E-27;
  o->massGe = 120.6049350E-27;
  o->massGa = 1.157777417E-25;
  o->Amol = 0.35; o->Bmol = 0.55; o->Cmol = 0.10;
  o->alpha_mol = 0.3; o->beta_mol = 1.2; o->gamma_mol = 6.0;
  o->A = 7.049556277; o->B = 0.6022245584;
  o->psi = 4.0; o->qsi = 0.0; o->al = 1.8;
  o->lambdaSi = 21.0; o->lambdaGe = 31.0; o->gamma_ = 1.2;
  o->epsSi = 3.473928E-19; o->epsGe = 3.085E-19;
  o->sigmaSi = 2.0951E-10; o->sigmaGe = 2.0951E-10;
  o->listfac = 1.1;
  o->ntable = 12000;
}

/* setupSiGe, prms.f90:452-480 (tables indexed [type-1]) */
static void setup_sige(orc *o) {
  double eSi = o->epsSi, eGe = o->epsGe, sSi = o->sigmaSi, sGe = o->sigmaGe;
  double lSi = o->lambdaSi, lGe = o->lambdaGe;
  o->eps[0][0] = eSi;
  o->eps[0][1] = 0.5 * (eSi + eGe);
  o->eps[1][1] = eGe;
  o->eps[1][0] = 0.5 * (eSi + eGe);
  o->isigma[0][0] = 1. / sSi;
  o->isigma[0][1] = 2. / (sSi + sGe);
  o->isigma[1][1] = 1. / sGe;
  o->isigma[1][0] = 2. / (sSi + sGe);
  o->leps[0][0][0] = lSi * eSi;
  double m1 = 2. * lSi * eSi / 3. + lGe * eGe / 3.;
  double m2 = lSi * eSi / 3. + 2. * lGe * eGe / 3.;
  o->leps[0][0][1] = m1; o->leps[0][1][0] = m1; o->leps[1][0][0] = m1;
  o->leps[0][1][1] = m2; o->leps[1][1][0] = m2; o->leps[1][0][1] = m2;
  o->leps[1][1][1] = lGe * eGe;
  o->epssig[0][0] = eSi / sSi;
  o->epssig[0][1] = (eSi + eGe) / (sSi + sGe);
  o->epssig[1][0] = (eSi + eGe) / (sSi + sGe);
  o->epssig[1][1] = eGe / sGe;
}

/* setupMol, prms.f90:433-449 */
static void setup_mol(orc *o) {
  o->Zmol[0] = o->ZSi; o->Zmol[1] = o->ZGe; o->Zmol[2] = o->ZGa;
  memset(o->aamol, 0, sizeof(o->aamol));
  o->aamol[0][2] = 0.885 * ao * pow(sqrt(o->ZSi) + sqrt(o->ZGa), -2. / 3.);
  o->aamol[1][2] = 0.885 * ao * pow(sqrt(o->ZGe) + sqrt(o->ZGa), -2. / 3.);
  o->aamol[2][0] = o->aamol[0][2];
  o->aamol[2][1] = o->aamol[1][2];
  o->aamol[2][2] = 0.885 * ao * pow(sqrt(o->ZGa) + sqrt(o->ZGa), -2. / 3.);
  o->ie0 = 1 / (4 * o->Pi * eps0);
}

/* ionrun, prms.f90:563-584 */
void orc_ionrun(orc *o, int32_t tic) {
  if (tic == 1) {
    o->ts = o->ts_i; o->atm_out = o->atm_out_i; o->ntlist = o->ntlist_i;
    o->atlist = o->atlist_i; o->Tau = o->Tau_i; o->dslist = o->dslist_i;
  } else if (tic == 2) {
    o->ts = o->ts_r; o->atm_out = o->atm_out_r; o->ntlist = o->ntlist_r;
    o->atlist = o->atlist_r; o->Tau = o->Tau_r; o->dslist = o->dslist_r;
  }
}

static void alloc_atoms(orc *o) {
  size_t n = (size_t)o->natm;
  free_arrays(o);
  free(o->nss); free(o->nsslj);
  o->nss = o->nsslj = NULL;
  o->X = (double *)calloc(3 * n, sizeof(double));
  o->Xi = (double *)calloc(3 * n, sizeof(double));
  o->V = (double *)calloc(3 * n, sizeof(double));
  o->F = (double *)calloc(3 * n, sizeof(double));
  o->Fp = (double *)calloc(3 * n, sizeof(double));
  o->mass = (double *)calloc(n, sizeof(double));
  o->atype = (int32_t *)calloc(n, sizeof(int32_t));
  o->xsides = (int32_t *)calloc(n, sizeof(int32_t));
  free(o->dead); free(o->Xb);
  o->dead = NULL; o->Xb = NULL;
}

/* Text round trips of the reference's intermediate files. */
static double rt_e20_10(double x) { /* (E20.10): 10 significant digits */
  char buf[64];
  snprintf(buf, sizeof buf, "%.9E", x);
  return strtod(buf, NULL);
}
static double rt_e12_6(double x) { /* (E12.6): 6 significant digits */
  char buf[64];
  snprintf(buf, sizeof buf, "%.5E", x);
  return strtod(buf, NULL);
}

/* pureepi, prms.f90:484-557.  Writes the lattice to o->lat through the
 * E20.10 text round trip of si.dat (prms.f90:523-528, data.f90:135-138). */
static void pureepi(orc *o, int nc1, int nc2, int nc3) {
  double uc[8][3];
  int nsg = 8 * nc1 * nc2 * nc3;
  double *XX = (double *)malloc(sizeof(double) * 3 * (size_t)nsg);
  double a[3];
  a[0] = a[1] = a[2] = 5.431073E-10;
  uc[0][0] = 0.00;        uc[0][1] = 0.00;        uc[0][2] = 0.00;
  uc[1][0] = 0.00;        uc[1][1] = 0.50 * a[1]; uc[1][2] = 0.50 * a[2];
  uc[2][0] = 0.250 * a[0]; uc[2][1] = 0.250 * a[1]; uc[2][2] = 0.750 * a[2];
  uc[3][0] = 0.250 * a[0]; uc[3][1] = 0.750 * a[1]; uc[3][2] = 0.250 * a[2];
  uc[4][0] = 0.50 * a[0]; uc[4][1] = 0.0 * a[1];  uc[4][2] = 0.50 * a[2];
  uc[5][0] = 0.50 * a[0]; uc[5][1] = 0.50 * a[1]; uc[5][2] = 0.0 * a[2];
  uc[6][0] = 0.750 * a[0]; uc[6][1] = 0.250 * a[1]; uc[6][2] = 0.250 * a[2];
  uc[7][0] = 0.750 * a[0]; uc[7][1] = 0.750 * a[1]; uc[7][2] = 0.750 * a[2];
  int m = 0;
  for (int i = 1; i <= nc1; ++i)
    for (int j = 1; j <= nc2; ++j)
      for (int k = 1; k <= nc3; ++k)
        for (int l = 0; l < 8; ++l) {
          XX[m] = uc[l][0] + a[0] * (double)(i - 1);
          XX[(size_t)nsg + m] = uc[l][1] + a[1] * (double)(j - 1);
          XX[2 * (size_t)nsg + m] = uc[l][2] + a[2] * (double)(k - 1);
          ++m;
        }
  o->lb[0] = a[0] * nc1;
  o->lb[1] = a[1] * nc2;
  o->lb[2] = a[2] * nc3 * 10;
  o->nsg = nsg;
  o->natm = nsg + o->nlj;
  double maxx = XX[0], maxy = XX[nsg];
  for (int i = 0; i < nsg; ++i) {
    if (XX[i] > maxx) maxx = XX[i];
    if (XX[(size_t)nsg + i] > maxy) maxy = XX[(size_t)nsg + i];
  }
  alloc_atoms(o);
  for (int i = 0; i < nsg; ++i) {
    double x = XX[i], y = XX[(size_t)nsg + i];
    if (x < 1.0e-10 || y < 1.0e-10 || x > maxx - 1.0e-10 || y > maxy - 1.0e-10)
      o->xsides[i] = i + 1;
  }
  free(o->lat);
  o->lat = (double *)malloc(sizeof(double) * 3 * (size_t)nsg);
  for (size_t q = 0; q < 3 * (size_t)nsg; ++q) o->lat[q] = rt_e20_10(XX[q]);
  free(XX);
}

static void derived(orc *o) {
  setup_sige(o);
  setup_mol(o);
  orc_ionrun(o, 1);
  for (int l = 0; l < 3; ++l) o->ilb[l] = 1. / o->lb[l];
}

static void copy_siga(orc *o, const orc_siga *s);

/* readprms after the reads, prms.f90:279-357 */
int orc_readprms(orc *o, const orc_siga *s, int32_t ranseed) {
  copy_siga(o, s);
  if (o->amorcrys != 0) return -1; /* amorphous input file: out of scope */
  pureepi(o, s->nc[0], s->nc[1], s->nc[2]);
  o->ranseed = ranseed;
  derived(o);
  return 0;
}

/* Run parameters only (no lattice): for synthetic geometries. */
void orc_set_run(orc *o, const orc_siga *s) {
  copy_siga(o, s);
  orc_ionrun(o, 1);
}

static void copy_siga(orc *o, const orc_siga *s) {
  o->amorcrys = s->amorcrys; o->whichic = s->whichic;
  o->startimpact = s->startimpact; o->eV = s->eV;
  for (int l = 0; l < 3; ++l) { o->npc[l] = s->npc[l]; o->nc_lat[l] = s->nc[l]; }
  o->nlj = s->nlj; o->nt = s->nt; o->ts_i = s->ts_i; o->ts_r = s->ts_r;
  o->Ttar1 = s->ttar1; o->ke_limit_1 = s->ke_limit_1;
  o->ke_limit_2 = s->ke_limit_2; o->Tau_i = s->tau_i; o->Tau_r = s->tau_r;
  o->Teps = s->teps; o->Uo = s->uo;
  o->ntlist_i = s->ntlist_i; o->ntlist_r = s->ntlist_r;
  o->atlist_i = s->atlist_i; o->atlist_r = s->atlist_r;
  o->dslist_i = s->dslist_i; o->dslist_r = s->dslist_r;
  o->atm_out_i = s->atm_out_i; o->atm_out_r = s->atm_out_r;
  o->res_out = s->res_out; o->eng_out = s->eng_out; o->tmp_out = s->tmp_out;
  o->nrand = s->nrand;
  for (int l = 0; l < 4; ++l) o->outz[l] = s->outz[l];
  o->atomz[0] = s->atomz[0]; o->atomz[1] = s->atomz[1];
  o->ionz = s->ionz; o->knockz = s->knockz; o->sidewidth = s->sidewidth;
  o->dti = s->dti; o->fwhm = s->fwhm; o->phiz = s->phiz; o->phixy = s->phixy;
}

int orc_set_geometry(orc *o, int32_t natm, int32_t nsg, int32_t nlj,
                     const double lb[3]) {
  o->natm = natm; o->nsg = nsg; o->nlj = nlj;
  for (int l = 0; l < 3; ++l) o->lb[l] = lb[l];
  alloc_atoms(o);
  derived(o);
  return 0;
}

/* ---------------------------------------------------------------------- */
void orc_set_state(orc *o, const double *X, const double *V,
                   const int32_t *atype, const double *mass) {
  size_t n = (size_t)o->natm;
  if (X) memcpy(o->X, X, 3 * n * sizeof(double));
  if (V) memcpy(o->V, V, 3 * n * sizeof(double));
  if (atype) memcpy(o->atype, atype, n * sizeof(int32_t));
  if (mass) memcpy(o->mass, mass, n * sizeof(double));
}
void orc_get_state(const orc *o, double *X, double *V, double *F) {
  size_t n = (size_t)o->natm;
  if (X) memcpy(X, o->X, 3 * n * sizeof(double));
  if (V) memcpy(V, o->V, 3 * n * sizeof(double));
  if (F) memcpy(F, o->F, 3 * n * sizeof(double));
}
void orc_set_force(orc *o, const double *F) {
  memcpy(o->F, F, 3 * (size_t)o->natm * sizeof(double));
}
void orc_set_xsides(orc *o, const int32_t *xs) {
  memcpy(o->xsides, xs, (size_t)o->natm * sizeof(int32_t));
}
void orc_set_xi(orc *o, const double *Xi) {
  memcpy(o->Xi, Xi, 3 * (size_t)o->natm * sizeof(double));
}
void orc_get_xi(const orc *o, double *Xi) {
  memcpy(Xi, o->Xi, 3 * (size_t)o->natm * sizeof(double));
}
void orc_get_xsides(const orc *o, int32_t *xs) {
  memcpy(xs, o->xsides, (size_t)o->natm * sizeof(int32_t));
}
void orc_get_types(const orc *o, int32_t *atype, double *mass) {
  if (atype) memcpy(atype, o->atype, (size_t)o->natm * sizeof(int32_t));
  if (mass) memcpy(mass, o->mass, (size_t)o->natm * sizeof(double));
}
void orc_get_xrand(const orc *o, double *xr) {
  if (o->Xrand) memcpy(xr, o->Xrand, 2 * (size_t)o->nrand * sizeof(double));
}

#define INTFIELD(nm) if (!strcmp(name, #nm)) return o->nm;
int32_t orc_get_int(const orc *o, const char *name) {
  INTFIELD(natm) INTFIELD(nsg) INTFIELD(nlj) INTFIELD(impact) INTFIELD(lt)
  INTFIELD(ntlist) INTFIELD(atlist) INTFIELD(dslist) INTFIELD(atm_out)
  INTFIELD(ranseed) INTFIELD(atom_is_fast) INTFIELD(atom_is_slow)
  INTFIELD(nrand) INTFIELD(ntable) INTFIELD(nbrs) INTFIELD(len2)
  INTFIELD(lenlj)
  if (!strcmp(name, "nc1")) return o->nc[0];
  if (!strcmp(name, "nc2")) return o->nc[1];
  if (!strcmp(name, "nc3")) return o->nc[2];
  return -999999;
}
#define DBLFIELD(nm) if (!strcmp(name, #nm)) return o->nm;
double orc_get_dbl(const orc *o, const char *name) {
  DBLFIELD(ts) DBLFIELD(time) DBLFIELD(temp) DBLFIELD(Tau) DBLFIELD(rc)
  DBLFIELD(rclj) DBLFIELD(rlj2) DBLFIELD(umax) DBLFIELD(ie0) DBLFIELD(Pi)
  DBLFIELD(restart_time) DBLFIELD(dti) DBLFIELD(eV)
  if (!strcmp(name, "lb1")) return o->lb[0];
  if (!strcmp(name, "lb2")) return o->lb[1];
  if (!strcmp(name, "lb3")) return o->lb[2];
  if (!strcmp(name, "aamol13")) return o->aamol[0][2];
  if (!strcmp(name, "aamol33")) return o->aamol[2][2];
  return NAN;
}
void orc_set_int(orc *o, const char *name, int32_t v) {
  if (!strcmp(name, "impact")) o->impact = v;
  else if (!strcmp(name, "lt")) o->lt = v;
  else if (!strcmp(name, "ranseed")) o->ranseed = v;
  else if (!strcmp(name, "atom_is_fast")) o->atom_is_fast = v;
  else if (!strcmp(name, "atom_is_slow")) o->atom_is_slow = v;
}
void orc_set_dbl(orc *o, const char *name, double v) {
  if (!strcmp(name, "ts")) o->ts = v;
  else if (!strcmp(name, "time")) o->time = v;
  else if (!strcmp(name, "temp")) o->temp = v;
  else if (!strcmp(name, "Tau")) o->Tau = v;
  else if (!strcmp(name, "sidewidth")) o->sidewidth = v;
}

/* ---------------------------------------------------------------------- */
/* ran1, ran1.f:10-36 (Press et al., Numerical Recipes in Fortran 77, 2nd
 * ed., sec. 7.1: Park-Miller minimal standard with Bays-Durham shuffle). */
double orc_ran1(orc *o) {
  const int32_t IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
  const int32_t NDIV = 1 + (IM - 1) / NTAB;
  const double AM = 1. / IM, EPS = 1.2e-7, RNMX = 1. - EPS;
  int32_t *idum = &o->ranseed;
  int32_t j, k;
  if (*idum <= 0 || o->iy == 0) {
    *idum = (-*idum > 1) ? -*idum : 1;
    for (j = NTAB + 8; j >= 1; --j) {
      k = *idum / IQ;
      *idum = IA * (*idum - k * IQ) - IR * k;
      if (*idum < 0) *idum += IM;
      if (j <= NTAB) o->iv[j - 1] = *idum;
    }
    o->iy = o->iv[0];
  }
  k = *idum / IQ;
  *idum = IA * (*idum - k * IQ) - IR * k;
  if (*idum < 0) *idum += IM;
  j = 1 + o->iy / NDIV;
  o->iy = o->iv[j - 1];
  o->iv[j - 1] = *idum;
  double r = AM * o->iy;
  return r < RNMX ? r : RNMX;
}

/* initranions, data.f90:263-285 (fwhm.dat written as E12.6, re-read by
 * readrandimp data.f90:151-166). */
void orc_initranions(orc *o) {
  free(o->Xrand);
  o->Xrand = (double *)malloc(sizeof(double) * 2 * (size_t)o->nrand);
  for (int i = 0; i < o->nrand; ++i) {
    double r1 = orc_ran1(o);
    double r2 = orc_ran1(o);
    double r3 = sqrt(-2 * log(r1)) * cos(2 * o->Pi * r2) * o->fwhm / 2.35 * 1.E-9;
    double r4 = sqrt(-2 * log(r2)) * sin(2 * o->Pi * r1) * o->fwhm / 2.35 * 1.E-9;
    o->Xrand[i] = rt_e12_6(r3);
    o->Xrand[(size_t)o->nrand + i] = rt_e12_6(r4);
  }
}

/* initIons, data.f90:223-260 */
static void init_ions(orc *o) {
  int nsg = o->nsg, natm = o->natm;
  for (int i = nsg; i < natm; ++i) {
    XA(o->X, i, 0) = o->lb[0] * orc_ran1(o);
    XA(o->X, i, 1) = o->lb[1] * orc_ran1(o);
    XA(o->X, i, 2) = o->ionz * o->lb[2] / 10.0;
    o->atype[i] = 3;
  }
  for (int i = nsg; i < natm; ++i)
    for (int j = nsg; j < natm; ++j) {
      if (i > j) {
        for (;;) { /* label 79 */
          double dx = XA(o->X, i, 0) - XA(o->X, j, 0);
          double dy = XA(o->X, i, 1) - XA(o->X, j, 1);
          double rij = sqrt(dx * dx + dy * dy);
          if (rij < o->sigmaSi) {
            XA(o->X, i, 0) = o->lb[0] * orc_ran1(o);
            XA(o->X, i, 1) = o->lb[1] * orc_ran1(o);
            continue;
          }
          break;
        }
      }
    }
}

/* assignmass, data.f90:207-220 */
static void assignmass(orc *o) {
  for (int i = 0; i < o->natm; ++i) {
    if (o->atype[i] == 1) o->mass[i] = o->massSi;
    else if (o->atype[i] == 2) o->mass[i] = o->massGe;
    else if (o->atype[i] == 3) o->mass[i] = o->massGa;
  }
}

/* randvel, data.f90:169-204 (amorcrys == 0 branch) */
static void randvel(orc *o) {
  double mm[3] = {0, 0, 0};
  int nsg = o->nsg;
  for (int i = 0; i < nsg; ++i) {
    for (int l = 0; l < 3; ++l) XA(o->V, i, l) = o->Uo * (orc_ran1(o) - 0.5);
    for (int l = 0; l < 3; ++l) mm[l] = mm[l] + o->mass[i] * XA(o->V, i, l);
  }
  for (int l = 0; l < 3; ++l) mm[l] = mm[l] / nsg;
  for (int i = 0; i < nsg; ++i)
    for (int l = 0; l < 3; ++l)
      XA(o->V, i, l) = XA(o->V, i, l) - mm[l] / o->mass[i];
  for (int i = nsg; i < o->natm; ++i)
    for (int l = 0; l < 3; ++l) XA(o->V, i, l) = 0.0;
}

/* initdata + ic(whichic=4), data.f90:32-68 and initlat :115-123 */
int orc_initdata(orc *o) {
  if (o->whichic <= 0) return -1; /* restart paths not restated */
  o->nt0 = 0; o->ions = 0; o->restart_time = 0.0;
  for (int i = 0; i < o->nsg; ++i) {
    for (int l = 0; l < 3; ++l)
      XA(o->X, i, l) = o->lat[(size_t)l * o->nsg + i];
    o->atype[i] = 1;
  }
  init_ions(o);
  memcpy(o->Xi, o->X, 3 * (size_t)o->natm * sizeof(double));
  assignmass(o);
  if (o->amorcrys == 0) randvel(o);
  return 0;
}

/* knock, knock.f90:21-55.  ranindex = NINT(ran1*Nrand) can be 0, an
 * out-of-bounds read of Xrand(0,:) in the reference; it is clamped to 1. */
void orc_knock(orc *o, int32_t knc) {
  double Pi = o->Pi;
  double kv = 1663.6334 * sqrt(o->eV);
  double dr = tan(o->phiz * Pi / 180) * (o->knockz * o->lb[2] / 10 - o->lb[2] / 10);
  double dx = dr * cos(o->phixy * Pi / 180) * sin(o->phiz * Pi / 180);
  double dy = dr * sin(o->phixy * Pi / 180) * sin(o->phiz * Pi / 180);
  int ri = fnint(orc_ran1(o) * o->nrand);
  if (ri < 1) ri = 1;
  int ion = o->nsg + knc - 1;
  XA(o->V, ion, 0) = kv * cos(o->phixy * Pi / 180) * sin(o->phiz * Pi / 180);
  XA(o->V, ion, 1) = kv * sin(o->phixy * Pi / 180) * sin(o->phiz * Pi / 180);
  XA(o->V, ion, 2) = -1.0 * kv * cos(o->phiz * Pi / 180);
  XA(o->X, ion, 0) = o->Xrand[ri - 1] + o->lb[0] / 2.0 - dx;
  XA(o->X, ion, 1) = o->Xrand[(size_t)o->nrand + ri - 1] + o->lb[1] / 2.0 - dy;
  XA(o->X, ion, 2) = o->knockz * o->lb[2] / 10.0;
}

/* ---------------------------------------------------------------------- */
/* init_nlist, si.f90:48-80 (capacities kept as advisory limits: the
 * restatement grows its arrays instead of writing past them). */
void orc_init_nlist(orc *o) {
  int nl = o->natm;
  o->ml3 = nl * 200;
  o->mllj = o->nlj * 1500;
  o->ml2 = nl * 40;
  double smax = o->sigmaGe > o->sigmaSi ? o->sigmaGe : o->sigmaSi;
  o->rc = o->listfac * o->al * smax;
  for (int l = 0; l < 3; ++l) {
    int n = (int)(o->lb[l] / o->rc);
    o->nc[l] = n > 1 ? n : 1;
  }
  o->rlj2 = 2.5 * smax;
  o->rclj = o->listfac * o->rlj2;
  o->rlj2 = o->rlj2 * o->rlj2;
  /* nss/nsslj are allocated once (si.f90:70-72).  The reference re-allocates
   * nlist2/nlist3/... here, leaving them undefined until the next si_nlist;
   * every atlist step is also an ntlist step in the reference's inputs.  The
   * restatement keeps the previous lists instead (well defined either way). */
  if (!o->nss) {
    o->nss = (int32_t *)calloc(2 * (size_t)o->natm, sizeof(int32_t));
    o->nsslj = (int32_t *)calloc(2 * (size_t)o->natm, sizeof(int32_t));
    for (int i = 0; i < o->natm; ++i) { o->nss[o->natm + i] = -1; o->nsslj[o->natm + i] = -1; }
  }
}

/* mktables, si.f90:575-597 */
void orc_mktables(orc *o) {
  int nt = o->ntable;
  free(o->table2bd); free(o->u);
  o->table2bd = (double *)calloc((size_t)nt + 3, sizeof(double));
  o->u = (double *)calloc((size_t)nt + 3, sizeof(double));
  o->umax = 2.5 * o->listfac * o->al;
  double du = o->umax / (double)(nt - 1);
  double A = o->A, B = o->B, psi = o->psi, qsi = o->qsi, al = o->al;
  for (int i = 2; i <= nt; ++i) {
    double ui = o->umax * (double)(i - 1) / (double)(nt - 1);
    o->u[i] = ui;
    if (ui < .999 * al) {
      double t1 = (B * pow(ui, -psi) - pow(ui, -qsi)) / ((ui - al) * (ui - al));
      double t2 = psi * B * pow(ui, -psi - 1) - qsi * pow(ui, -qsi - 1);
      o->table2bd[i] = -(A * exp(1. / (ui - al)) * (t1 + t2) / (du * du));
    } else {
      o->table2bd[i] = 0.;
    }
  }
}

double orc_table_value(const orc *o, int32_t i1) { return o->table2bd[i1]; }

/* itable2bd, si.f90:554-573 */
double orc_itable2bd(const orc *o, double ue) {
  int i = (int)(ue / o->umax * (double)(o->ntable - 1)) + 1;
  if (i > o->ntable) return 0.0;
  double d1 = ue - o->u[i];
  double d2 = ue - o->u[i + 1];
  double d3 = ue - o->u[i + 2];
  return 0.5 * d2 * d3 * o->table2bd[i] - d1 * d3 * o->table2bd[i + 1] +
         0.5 * d1 * d2 * o->table2bd[i + 2];
}

/* atomcell, nlist.f90:20-32, with hspace = Lb/REAL(N) from chainlist :44 */
void orc_atomcell(const orc *o, const double *X, int32_t *icell) {
  for (int l = 0; l < 3; ++l) {
    double h = o->lb[l] / (double)o->nc[l];
    for (int i = 0; i < o->natm; ++i) {
      double x = XA(X, i, l);
      double w = x - (double)(int)floor(x * o->ilb[l]) * o->lb[l];
      XA(icell, i, l) = (int)(w / h) + 1;
    }
  }
}

static void push2(orc *o, int32_t j) {
  if (o->len2 >= o->cap2) {
    o->cap2 = o->cap2 ? 2 * o->cap2 : 1024;
    o->nlist2 = (int32_t *)realloc(o->nlist2, sizeof(int32_t) * (size_t)o->cap2);
  }
  o->nlist2[o->len2++] = j;
}
static void pushlj(orc *o, int32_t j) {
  if (o->lenlj >= o->caplj) {
    o->caplj = o->caplj ? 2 * o->caplj : 1024;
    o->nlistlj = (int32_t *)realloc(o->nlistlj, sizeof(int32_t) * (size_t)o->caplj);
  }
  o->nlistlj[o->lenlj++] = j;
}
static void push3(orc *o, int32_t a, int32_t b, int32_t c) {
  if (o->nbrs >= o->cap3) {
    o->cap3 = o->cap3 ? 2 * o->cap3 : 4096;
    o->nlist3 = (int32_t *)realloc(o->nlist3, sizeof(int32_t) * 3 * (size_t)o->cap3);
  }
  o->nlist3[3 * (size_t)o->nbrs + 0] = a;
  o->nlist3[3 * (size_t)o->nbrs + 1] = b;
  o->nlist3[3 * (size_t)o->nbrs + 2] = c;
  o->nbrs++;
}

static int imin3(int a, int b, int c) { int m = a < b ? a : b; return m < c ? m : c; }
static int imax3(int a, int b, int c) { int m = a > b ? a : b; return m > c ? m : c; }

/* si_nlist -> si_nlistG, si.f90:87-276, with chainlist nlist.f90:34-66.
 * At one rank P_big(i) == myid for all atoms, so the ownership filter of
 * si.f90:139,153,155 reduces to i != j; si_nlistL (par.f90:626-722) is the
 * identity and is not restated. */
int orc_si_nlist(orc *o, const double *X) {
  const int N1 = o->nc[0], N2 = o->nc[1], N3 = o->nc[2];
  const int natm = o->natm;
  const int S1 = N1 + 2, S2 = N2 + 2;
  size_t hsz = (size_t)S1 * S2 * (N3 + 2);
  int32_t *HOC = (int32_t *)calloc(hsz, sizeof(int32_t));
  int32_t *LL = (int32_t *)calloc((size_t)natm + 1, sizeof(int32_t));
  int32_t *ic = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)natm);
#define H(a, b, c) HOC[((size_t)(c) * S2 + (b)) * S1 + (a)]
  orc_atomcell(o, X, ic);
  for (int i = 1; i <= natm; ++i) {
    /* config mode: a removed atom is in no cell and so in no list */
    if (o->dead && o->dead[i - 1]) continue;
    int a = XA(ic, i - 1, 0), b = XA(ic, i - 1, 1), c = XA(ic, i - 1, 2);
    LL[i] = H(a, b, c);
    H(a, b, c) = i;
  }
  if (N1 > 2)
    for (int c = 0; c < N3 + 2; ++c)
      for (int b = 0; b < N2 + 2; ++b) { H(0, b, c) = H(N1, b, c); H(N1 + 1, b, c) = H(1, b, c); }
  if (N2 > 2)
    for (int c = 0; c < N3 + 2; ++c)
      for (int a = 0; a < N1 + 2; ++a) { H(a, 0, c) = H(a, N2, c); H(a, N2 + 1, c) = H(a, 1, c); }
  if (N3 > 2)
    for (int b = 0; b < N2 + 2; ++b)
      for (int a = 0; a < N1 + 2; ++a) { H(a, b, 0) = H(a, b, N3); H(a, b, N3 + 1) = H(a, b, 1); }

  double rv2 = o->rc * o->rc;
  double rv2lj = o->rclj * o->rclj;
  o->len2 = 0; o->lenlj = 0;
  int32_t *nss1 = o->nss, *nss2 = o->nss + natm;
  int32_t *njs1 = o->nsslj, *njs2 = o->nsslj + natm;
  for (int i = 0; i < natm; ++i) { nss2[i] = -1; nss1[i] = 0; njs2[i] = -1; njs1[i] = 0; }
  for (int c3 = 1; c3 <= N3; ++c3)
    for (int c2 = 1; c2 <= N2; ++c2)
      for (int c1 = 1; c1 <= N1; ++c1) {
        int i = H(c1, c2, c3);
        while (i != 0) {
          nss1[i - 1] = o->len2 + 1;
          njs1[i - 1] = o->lenlj + 1;
          for (int n1 = -1; n1 <= 1; ++n1)
            for (int n2 = -1; n2 <= 1; ++n2)
              for (int n3 = -1; n3 <= 1; ++n3) {
                int j = H(c1 + n1, c2 + n2, c3 + n3);
                while (j != 0) {
                  if (i != j) {
                    double xx[3];
                    for (int l = 0; l < 3; ++l) {
                      double d = XA(X, i - 1, l) - XA(X, j - 1, l);
                      xx[l] = d - (double)fnint(d * o->ilb[l]) * o->lb[l];
                    }
                    double rij2 = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
                    int ti = o->atype[i - 1], tj = o->atype[j - 1];
                    if ((ti > tj ? ti : tj) < 3) {
                      if (rij2 < rv2) push2(o, j);
                    } else {
                      if (rij2 < rv2lj) pushlj(o, j);
                    }
                  }
                  j = LL[j];
                }
              }
          nss2[i - 1] = o->len2;
          njs2[i - 1] = o->lenlj;
          i = LL[i];
        }
      }
#undef H
  /* keep a copy of the full lists (CSR by atom) for the parity exports */
  free(o->full2_cnt); free(o->fulllj_cnt); free(o->full2); free(o->fulllj);
  o->full2_cnt = (int32_t *)calloc((size_t)natm, sizeof(int32_t));
  o->fulllj_cnt = (int32_t *)calloc((size_t)natm, sizeof(int32_t));
  o->full2 = (int32_t *)malloc(sizeof(int32_t) * ((size_t)o->len2 + 1));
  o->fulllj = (int32_t *)malloc(sizeof(int32_t) * ((size_t)o->lenlj + 1));
  {
    size_t p = 0, q = 0;
    for (int i = 0; i < natm; ++i) {
      o->full2_cnt[i] = nss2[i] - nss1[i] + 1;
      for (int l = nss1[i]; l <= nss2[i]; ++l) o->full2[p++] = o->nlist2[l - 1];
      o->fulllj_cnt[i] = njs2[i] - njs1[i] + 1;
      for (int l = njs1[i]; l <= njs2[i]; ++l) o->fulllj[q++] = o->nlistlj[l - 1];
    }
  }

  /* triplets, si.f90:191-232 */
  o->nbrs = 0;
  for (int i = 1; i <= natm; ++i) {
    for (int l = nss1[i - 1]; l <= nss2[i - 1]; ++l) {
      int j = o->nlist2[l - 1];
      if (j > i) {
        for (int m = l + 1; m <= nss2[i - 1]; ++m) {
          int k = o->nlist2[m - 1];
          if (k > i)
            push3(o, imin3(i, j, k), j + k + i - imin3(i, j, k) - imax3(i, j, k),
                  imax3(i, j, k));
        }
        for (int m = nss1[j - 1]; m <= nss2[j - 1]; ++m) {
          int k = o->nlist2[m - 1];
          if (k > i) {
            int dup = 0;
            for (int n = nss1[i - 1]; n <= nss2[i - 1]; ++n)
              if (k == o->nlist2[n - 1]) { dup = 1; break; }
            if (dup) continue;
            push3(o, imin3(i, j, k), j + k + i - imin3(i, j, k) - imax3(i, j, k),
                  imax3(i, j, k));
          }
        }
      }
    }
  }

  /* half-list compaction, si.f90:234-260.  The Fortran DO trip count is
   * fixed on entry, hence the saved upper bound. */
  for (int pass = 0; pass < 2; ++pass) {
    int32_t *s1 = pass ? njs1 : nss1, *s2 = pass ? njs2 : nss2;
    int32_t *lst = pass ? o->nlistlj : o->nlist2;
    for (int i = 1; i <= natm; ++i) {
      int lo = s1[i - 1], hi = s2[i - 1];
      for (int l = lo; l <= hi; ++l) {
        int j = lst[l - 1];
        while (j < i && s2[i - 1] > l) {
          lst[l - 1] = lst[s2[i - 1] - 1];
          j = lst[l - 1];
          s2[i - 1] = s2[i - 1] - 1;
        }
        if (s2[i - 1] == l && j < i) s2[i - 1] = s2[i - 1] - 1;
      }
    }
  }
  free(HOC); free(LL); free(ic);
  int over = 0;
  if (o->nbrs >= o->ml3) over |= 1;
  if (o->len2 + 1 >= o->ml2) over |= 2;
  if (o->nlj > 0 && o->lenlj + 1 >= o->mllj) over |= 4;
  return over;
}

int32_t orc_nlist2_len(const orc *o) { return o->len2; }
int32_t orc_nlistlj_len(const orc *o) { return o->lenlj; }
int32_t orc_nbrs(const orc *o) { return o->nbrs; }
void orc_get_nlists(const orc *o, int32_t *nss, int32_t *nlist2, int32_t *nsslj,
                    int32_t *nlistlj, int32_t *nlist3) {
  size_t n = (size_t)o->natm;
  if (nss) memcpy(nss, o->nss, 2 * n * sizeof(int32_t));
  if (nsslj) memcpy(nsslj, o->nsslj, 2 * n * sizeof(int32_t));
  if (nlist2 && o->len2) memcpy(nlist2, o->nlist2, (size_t)o->len2 * sizeof(int32_t));
  if (nlistlj && o->lenlj) memcpy(nlistlj, o->nlistlj, (size_t)o->lenlj * sizeof(int32_t));
  if (nlist3 && o->nbrs) memcpy(nlist3, o->nlist3, 3 * (size_t)o->nbrs * sizeof(int32_t));
}
void orc_get_full_counts(const orc *o, int32_t *c2, int32_t *clj) {
  size_t n = (size_t)o->natm;
  if (c2) memcpy(c2, o->full2_cnt, n * sizeof(int32_t));
  if (clj) memcpy(clj, o->fulllj_cnt, n * sizeof(int32_t));
}
void orc_get_full_lists(const orc *o, int32_t *f2, int32_t *flj) {
  size_t a = 0, b = 0;
  for (int i = 0; i < o->natm; ++i) { a += o->full2_cnt[i]; b += o->fulllj_cnt[i]; }
  if (f2 && a) memcpy(f2, o->full2, a * sizeof(int32_t));
  if (flj && b) memcpy(flj, o->fulllj, b * sizeof(int32_t));
}

/* ---------------------------------------------------------------------- */
static void min_image(const orc *o, const double *X, int i, int j, double xx[3]) {
  for (int l = 0; l < 3; ++l) {
    double d = XA(X, i, l) - XA(X, j, l);
    xx[l] = d - (double)fnint(d * o->ilb[l]) * o->lb[l];
  }
}

static void vir_add(double *w, const double r[3], const double f[3]) {
  if (!w) return;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) w[3 * a + b] += r[a] * f[b];
}

/* Atoms a force or energy term leaves out: ions not fired yet (index above
 * Nsg+impact, si.f90:415,428,457) and, in config mode, removed atoms, which
 * stay in the lists of the last build until the next one. */
static int skipped(const orc *o, int i1, int lim) {
  return i1 > lim || (o->dead && o->dead[i1 - 1]);
}

/* si_force, si.f90:384-552.  The virial (absent in the reference) sums
 * r_ij (x) f_i for pairs and -x_ij (x) f_j - x_ik (x) f_k for triplets. */
void orc_si_force(orc *o, const double *Xp, double *F, double *W) {
  const int natm = o->natm;
  const int lim = o->nsg + o->impact; /* 1-based index limit */
  const double al = o->al, thrd = o->thrd, gam = o->gamma_;
  memset(F, 0, 3 * (size_t)natm * sizeof(double));
  if (W) memset(W, 0, 9 * sizeof(double));
  for (int i1 = 1; i1 <= natm; ++i1) {
    int i = i1 - 1;
    for (int l = o->nss[i]; l <= o->nss[natm + i]; ++l) {
      int j1 = o->nlist2[l - 1], j = j1 - 1;
      if (skipped(o, i1, lim) || skipped(o, j1, lim)) continue;
      double xx[3];
      min_image(o, Xp, i, j, xx);
      double rij = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
      int ti = o->atype[i] - 1, tj = o->atype[j] - 1;
      double rijs = rij * o->isigma[ti][tj];
      double df = orc_itable2bd(o, rijs);
      double fi[3];
      for (int c = 0; c < 3; ++c) {
        fi[c] = -(df * xx[c] / rij * o->epssig[ti][tj]);
        XA(F, i, c) = XA(F, i, c) + fi[c];
        XA(F, j, c) = XA(F, j, c) - fi[c];
      }
      vir_add(W, xx, fi);
    }
    for (int l = o->nsslj[i]; l <= o->nsslj[natm + i]; ++l) {
      int j1 = o->nlistlj[l - 1], j = j1 - 1;
      if (skipped(o, i1, lim) || skipped(o, j1, lim)) continue;
      double xx[3];
      min_image(o, Xp, i, j, xx);
      double rij2 = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
      if (rij2 < o->rlj2) {
        double rij = sqrt(rij2);
        int ti = o->atype[i] - 1, tj = o->atype[j] - 1;
        double aa = o->aamol[ti][tj];
        double e1 = exp(-(o->alpha_mol * rij / aa));
        double e2 = exp(-(o->beta_mol * rij / aa));
        double e3 = exp(-(o->gamma_mol * rij / aa));
        double pre = o->ie0 * o->Zmol[ti] * o->Zmol[tj] * ec * ec / rij;
        double s1 = (o->Amol * e1 + o->Bmol * e2 + o->Cmol * e3) / rij;
        double s2 = o->Amol * o->alpha_mol / aa * e1 +
                    o->Bmol * o->beta_mol / aa * e2 +
                    o->Cmol * o->gamma_mol / aa * e3;
        double fc = pre * (s1 + s2);
        double fi[3];
        for (int c = 0; c < 3; ++c) {
          fi[c] = fc * xx[c] / rij;
          XA(F, i, c) = XA(F, i, c) + fi[c];
          XA(F, j, c) = XA(F, j, c) - fi[c];
        }
        vir_add(W, xx, fi);
      }
    }
  }
  for (int l = 0; l < o->nbrs; ++l) {
    int i1 = o->nlist3[3 * (size_t)l], j1 = o->nlist3[3 * (size_t)l + 1],
        k1 = o->nlist3[3 * (size_t)l + 2];
    if (skipped(o, i1, lim) || skipped(o, j1, lim) || skipped(o, k1, lim)) continue;
    int i = i1 - 1, j = j1 - 1, k = k1 - 1;
    int ti = o->atype[i] - 1, tj = o->atype[j] - 1, tk = o->atype[k] - 1;
    double isigij = o->isigma[ti][tj], isigik = o->isigma[ti][tk],
           isigjk = o->isigma[tj][tk];
    double lepsijk = o->leps[ti][tj][tk];
    double xxij[3], xxik[3], xxjk[3], dxij[3], dxik[3], dxjk[3];
    min_image(o, Xp, i, j, xxij);
    double rij = 1. / sqrt(xxij[0] * xxij[0] + xxij[1] * xxij[1] + xxij[2] * xxij[2]);
    for (int c = 0; c < 3; ++c) dxij[c] = xxij[c] * rij;
    double rijs = isigij / rij;
    min_image(o, Xp, i, k, xxik);
    double rik = 1. / sqrt(xxik[0] * xxik[0] + xxik[1] * xxik[1] + xxik[2] * xxik[2]);
    for (int c = 0; c < 3; ++c) dxik[c] = xxik[c] * rik;
    double riks = isigik / rik;
    min_image(o, Xp, j, k, xxjk);
    double rjk = 1. / sqrt(xxjk[0] * xxjk[0] + xxjk[1] * xxjk[1] + xxjk[2] * xxjk[2]);
    for (int c = 0; c < 3; ++c) dxjk[c] = xxjk[c] * rjk;
    double rjks = isigjk / rjk;

    double ctjik = dxij[0] * dxik[0] + dxij[1] * dxik[1] + dxij[2] * dxik[2];
    double ctijk = -(dxij[0] * dxjk[0] + dxij[1] * dxjk[1] + dxij[2] * dxjk[2]);
    double ctikj = dxik[0] * dxjk[0] + dxik[1] * dxjk[1] + dxik[2] * dxjk[2];
    double dcjik[3], dcijk[3], dcikj[3];
    for (int c = 0; c < 3; ++c) {
      dcjik[c] = dxij[c] * rik + dxik[c] * rij - ctjik * (dxij[c] * rij + dxik[c] * rik);
      dcijk[c] = -dxjk[c] * rij - ctijk * dxij[c] * rij;
      dcikj[c] = +dxjk[c] * rik - ctikj * dxik[c] * rik;
    }
    double Eij = 0., Eik = 0., Ejk = 0.;
    double iij = 1. / (rijs - al), iik = 1. / (riks - al), ijk = 1. / (rjks - al);
    if (rijs < 0.999 * al) Eij = exp(gam * iij);
    if (riks < 0.999 * al) Eik = exp(gam * iik);
    if (rjks < 0.999 * al) Ejk = exp(gam * ijk);
    double fijk = (rijs < 0.999 * al && rjks < 0.999 * al) ? 1. : 0.;
    double fikj = (riks < 0.999 * al && rjks < 0.999 * al) ? 1. : 0.;
    double fjik = (rijs < 0.999 * al && riks < 0.999 * al) ? 1. : 0.;
    double dEij = -gam * iij * iij, dEik = -gam * iik * iik, dEjk = -gam * ijk * ijk;
    double Wjik = (ctjik + thrd) * (ctjik + thrd);
    double Wijk = (ctijk + thrd) * (ctijk + thrd);
    double Wikj = (ctikj + thrd) * (ctikj + thrd);
    double fi[3], fj[3], fk[3];
    for (int c = 0; c < 3; ++c) {
      fi[c] = -(lepsijk *
                (Eij * (fjik * Eik *
                            (Wjik * (dEij * dxij[c] * isigij + dEik * dxik[c] * isigik) +
                             (ctjik + thrd) * dcjik[c] * 2.) +
                        fijk * Ejk *
                            (dEij * Wijk * dxij[c] * isigij + (ctijk + thrd) * dcijk[c] * 2.)) +
                 fikj * Eik * Ejk *
                     (dEik * Wikj * dxik[c] * isigik + (ctikj + thrd) * dcikj[c] * 2.)));
    }
    for (int c = 0; c < 3; ++c) {
      dcjik[c] = -dxik[c] * rij + ctjik * dxij[c] * rij;
      dcijk[c] = -dxij[c] * rjk + dxjk[c] * rij + ctijk * (dxij[c] * rij - dxjk[c] * rjk);
      dcikj[c] = +dxik[c] * rjk - ctikj * dxjk[c] * rjk;
    }
    for (int c = 0; c < 3; ++c) {
      fj[c] = -(lepsijk *
                (+Eij * (fjik * Eik * (-dEij * Wjik * dxij[c] * isigij +
                                       (ctjik + thrd) * dcjik[c] * 2.) +
                         fijk * Ejk *
                             (Wijk * (-dEij * dxij[c] * isigij + dEjk * dxjk[c] * isigjk) +
                              (ctijk + thrd) * dcijk[c] * 2.)) +
                 fikj * Eik * Ejk *
                     (dEjk * Wikj * dxjk[c] * isigjk + (ctikj + thrd) * dcikj[c] * 2.)));
      fk[c] = -fi[c] - fj[c];
    }
    for (int c = 0; c < 3; ++c) {
      XA(F, i, c) = XA(F, i, c) + fi[c];
      XA(F, j, c) = XA(F, j, c) + fj[c];
      XA(F, k, c) = XA(F, k, c) + fk[c];
    }
    if (W) {
      double rj[3], rk[3];
      for (int c = 0; c < 3; ++c) { rj[c] = -xxij[c]; rk[c] = -xxik[c]; }
      vir_add(W, rj, fj);
      vir_add(W, rk, fk);
    }
  }
}

/* h, si.f90:371-380 */
static double hfun(const orc *o, double u, double v, double w) {
  if (u < 0.999 * o->al && v < 0.999 * o->al)
    return exp(o->gamma_ / (u - o->al) + o->gamma_ / (v - o->al)) * ((w + 1. / 3.) * (w + 1. / 3.));
  return 0.;
}

/* si_potential, si.f90:278-368.  lim < 0 keeps every listed term, as the
 * reference does; config mode passes Nsg+impact and leaves out the terms
 * the force leaves out (unfired ions, removed atoms). */
static void potential_impl(orc *o, const double *X, double *utot, double *u, int lim) {
  const int natm = o->natm;
  double U = 0.;
  memset(u, 0, (size_t)natm * sizeof(double));
#define SKIP(a) (lim >= 0 && skipped(o, (a) + 1, lim))
  for (int i = 0; i < natm; ++i) {
    for (int l = o->nss[i]; l <= o->nss[natm + i]; ++l) {
      int j = o->nlist2[l - 1] - 1;
      if (SKIP(i) || SKIP(j)) continue;
      double xx[3];
      min_image(o, X, i, j, xx);
      double rij = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
      int ti = o->atype[i] - 1, tj = o->atype[j] - 1;
      double r = rij * o->isigma[ti][tj];
      if (r < 0.999 * o->al) {
        double ul = 0.5 * o->eps[ti][tj] * o->A * (o->B * pow(r, -o->psi) - pow(r, -o->qsi)) *
                    exp(1. / (r - o->al));
        u[i] = u[i] + ul;
        u[j] = u[j] + ul;
        U = U + 2. * ul;
      }
    }
    for (int l = o->nsslj[i]; l <= o->nsslj[natm + i]; ++l) {
      int j = o->nlistlj[l - 1] - 1;
      if (SKIP(i) || SKIP(j)) continue;
      double xx[3];
      min_image(o, X, i, j, xx);
      double rij2 = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
      if (rij2 < o->rlj2) {
        double rij = sqrt(rij2);
        int ti = o->atype[i] - 1, tj = o->atype[j] - 1;
        double aa = o->aamol[ti][tj];
        double ul = o->ie0 * o->Zmol[ti] * o->Zmol[tj] * ec * ec / rij *
                    (o->Amol * exp(-(o->alpha_mol * rij / aa)) +
                     o->Bmol * exp(-(o->beta_mol * rij / aa)) +
                     o->Cmol * exp(-(o->gamma_mol * rij / aa)));
        u[i] = u[i] + ul / 2.;
        u[j] = u[j] + ul / 2.;
        U = U + ul;
      }
    }
  }
  for (int l = 0; l < o->nbrs; ++l) {
    int i = o->nlist3[3 * (size_t)l] - 1, j = o->nlist3[3 * (size_t)l + 1] - 1,
        k = o->nlist3[3 * (size_t)l + 2] - 1;
    if (SKIP(i) || SKIP(j) || SKIP(k)) continue;
    int ti = o->atype[i] - 1, tj = o->atype[j] - 1, tk = o->atype[k] - 1;
    double xij[3], xik[3], xjk[3];
    min_image(o, X, i, j, xij);
    double rij = sqrt(xij[0] * xij[0] + xij[1] * xij[1] + xij[2] * xij[2]);
    double rijs = rij * o->isigma[ti][tj];
    min_image(o, X, i, k, xik);
    double rik = sqrt(xik[0] * xik[0] + xik[1] * xik[1] + xik[2] * xik[2]);
    double riks = rik * o->isigma[ti][tk];
    min_image(o, X, j, k, xjk);
    double rjk = sqrt(xjk[0] * xjk[0] + xjk[1] * xjk[1] + xjk[2] * xjk[2]);
    double rjks = rjk * o->isigma[tj][tk];
    double ctjik = (xij[0] * xik[0] + xij[1] * xik[1] + xij[2] * xik[2]) / rij / rik;
    double ctijk = -(xij[0] * xjk[0] + xij[1] * xjk[1] + xij[2] * xjk[2]) / rij / rjk;
    double ctikj = (xik[0] * xjk[0] + xik[1] * xjk[1] + xik[2] * xjk[2]) / rik / rjk;
    double ul = 1. / 3. * o->leps[ti][tj][tk] *
                (hfun(o, rijs, riks, ctjik) + hfun(o, rijs, rjks, ctijk) +
                 hfun(o, riks, rjks, ctikj));
    u[i] = u[i] + ul;
    u[j] = u[j] + ul;
    u[k] = u[k] + ul;
    U = U + 3. * ul;
  }
#undef SKIP
  *utot = U;
}

void orc_si_potential(orc *o, const double *X, double *utot, double *u) {
  potential_impl(o, X, utot, u, -1);
}

/* ---------------------------------------------------------------------- */
/* temperature, temp.f90:183-205 */
void orc_temperature(const orc *o, double *temp) {
  double t = 0.;
  for (int i = 0; i < o->natm; ++i)
    if (o->atype[i] < 3) {
      double v1 = XA(o->V, i, 0), v2 = XA(o->V, i, 1), v3 = XA(o->V, i, 2);
      t = t + o->mass[i] * (v1 * v1 + v2 * v2 + v3 * v3);
    }
  *temp = t / 3. / (double)o->nsg / kB;
}

/* adjusttemp, temp.f90:160-179 */
void orc_adjusttemp(orc *o, double temp, double ttar) {
  double s = sqrt(1. + o->ts / o->Tau * (ttar / temp - 1.));
  for (int i = 0; i < o->natm; ++i)
    if (o->atype[i] < 3)
      for (int l = 0; l < 3; ++l) XA(o->V, i, l) = s * XA(o->V, i, l);
}

/* controlsidetemp, temp.f90:32-67 (Fortran precedence: .and. before .or.) */
void orc_controlsidetemp(orc *o, double ttar) {
  const double nm = 1.0e-9;
  double t = 0.;
  int kk = 0;
  int *lst = (int *)malloc(sizeof(int) * (size_t)o->natm);
  for (int i = 0; i < o->natm; ++i) {
    double x = XA(o->X, i, 0), y = XA(o->X, i, 1), z = XA(o->X, i, 2);
    int xs = (x < o->sidewidth * nm) || (x > (o->lb[0] - o->sidewidth * nm));
    int ys = (y < o->sidewidth * nm) || (y > (o->lb[1] - o->sidewidth * nm));
    int zc = (z < o->outz[0] * o->lb[2] / 10.0) || (z > o->outz[3] * o->lb[2] / 10.0);
    if (xs || (ys && zc)) {
      if (o->atype[i] < 3) {
        double v1 = XA(o->V, i, 0), v2 = XA(o->V, i, 1), v3 = XA(o->V, i, 2);
        t = t + o->mass[i] * (v1 * v1 + v2 * v2 + v3 * v3);
        lst[kk++] = i;
      }
    }
  }
  if (kk > 0) {
    t = t / 3. / (double)kk / kB;
    double s = sqrt(1. + o->ts / o->Tau * (ttar / t - 1.));
    for (int q = 0; q < kk; ++q)
      for (int l = 0; l < 3; ++l) XA(o->V, lst[q], l) = s * XA(o->V, lst[q], l);
  }
  free(lst);
}

/* freezsidesN, tools.f90:38-51 */
void orc_freezsides(orc *o) {
  for (int i = 0; i < o->natm; ++i)
    if (o->xsides[i] > 0)
      for (int l = 0; l < 3; ++l) { XA(o->X, i, l) = XA(o->Xi, i, l); XA(o->V, i, l) = 0.; }
}

/* rebox, tools.f90:26-36 */
void orc_rebox(orc *o) {
  for (int l = 0; l < 3; ++l)
    for (int i = 0; i < o->natm; ++i)
      XA(o->X, i, l) = fmod(XA(o->X, i, l) + 2. * o->lb[l], o->lb[l]);
}

/* frzsptrdatms, temp.f90:99-127 */
void orc_frzsptrdatms(orc *o) {
  const double buffer = 0.3;
  const double T = o->lb[2];
  for (int i = 0; i < o->natm; ++i) {
    double z = XA(o->X, i, 2);
    int zero = 0;
    if ((z > o->outz[0] * T / 10.0) && (z < o->outz[1] * T / 10.0)) {
      XA(o->X, i, 2) = o->atomz[0] * T / 10.0; zero = 1;
    } else if ((z > o->outz[2] * T / 10.0) && (z < o->outz[3] * T / 10.0)) {
      XA(o->X, i, 2) = o->atomz[1] * T / 10.0; zero = 1;
    } else if ((z > (o->atomz[0] - buffer) * T / 10.0) && (z < (o->atomz[0] + buffer) * T / 10.0)) {
      zero = 1;
    } else if ((z > (o->atomz[1] - buffer) * T / 10.0) && (z < (o->atomz[1] + buffer) * T / 10.0)) {
      zero = 1;
    } else if ((z > (o->ionz - buffer) * T / 10.0) && (z < (o->ionz + buffer) * T / 10.0)) {
      zero = 1;
    }
    if (zero)
      for (int l = 0; l < 3; ++l) { XA(o->V, i, l) = 0.0; XA(o->F, i, l) = 0.0; }
  }
}

/* calc_atm, timeint.f90:277-298 */
void orc_calc_atm(orc *o) {
  double mx = -HUGE_VAL;
  for (int i = 0; i < o->natm; ++i) {
    double v1 = XA(o->V, i, 0), v2 = XA(o->V, i, 1), v3 = XA(o->V, i, 2);
    double ke = 0.5 * o->mass[i] * (v1 * v1 + v2 * v2 + v3 * v3) / ec;
    if (ke > mx) mx = ke;
  }
  if (mx < o->ke_limit_1 && o->atom_is_fast) {
    orc_ionrun(o, 2);
    o->atom_is_fast = 0;
  }
}

/* onestep, timeint.f90:158-242 */
int orc_onestep(orc *o, int32_t whichctrl, double ttar, int32_t lt) {
  const int natm = o->natm;
  o->time = o->time + o->ts;
  o->restart_time = o->time;
  if (whichctrl == 0) {
    orc_adjusttemp(o, o->temp, ttar);
    orc_temperature(o, &o->temp);
  }
  if (whichctrl > 0) orc_controlsidetemp(o, ttar);
  for (int i = 0; i < natm; ++i)
    for (int l = 0; l < 3; ++l)
      XA(o->X, i, l) = XA(o->X, i, l) + XA(o->V, i, l) * o->ts +
                       0.5 * XA(o->F, i, l) * o->ts * o->ts / o->mass[i];
  orc_freezsides(o);
  orc_rebox(o);
  if (lt % o->atlist == 0) {
    orc_frzsptrdatms(o);
    orc_init_nlist(o);
  }
  if (lt % o->ntlist == 0) orc_si_nlist(o, o->X);
  if (lt % o->dslist == 0 && whichctrl == 1) orc_calc_atm(o);
  orc_si_force(o, o->X, o->Fp, NULL);
  for (int i = 0; i < natm; ++i)
    for (int l = 0; l < 3; ++l)
      XA(o->V, i, l) = XA(o->V, i, l) +
                       0.5 * (XA(o->F, i, l) + XA(o->Fp, i, l)) * o->ts / o->mass[i];
  memcpy(o->F, o->Fp, 3 * (size_t)natm * sizeof(double));
  int out = 0;
  if (whichctrl > 0) {
    if ((o->res_out > 0 && lt % o->res_out == 0) ||
        (o->atm_out > 0 && lt % o->atm_out == 0)) {
      orc_temperature(o, &o->temp);
      out = 1;
    }
  }
  return out;
}

/* tint prologue, timeint.f90:47-69 */
void orc_tint_init(orc *o) {
  o->lt = o->nt0;
  o->time = o->restart_time;
  memset(o->F, 0, 3 * (size_t)o->natm * sizeof(double));
  orc_temperature(o, &o->temp);
  orc_si_force(o, o->X, o->F, NULL);
}

/* Entry of TI (timeint.f90:130-131): whichctrl 0 recomputes temp. */
void orc_ti_begin(orc *o, int32_t whichctrl) {
  if (whichctrl == 0) orc_temperature(o, &o->temp);
}

/* One iteration of the loops in TI, timeint.f90:130-154. */
int orc_ti_iter(orc *o, int32_t whichctrl, double ttar) {
  if (whichctrl == 0) {
    if (!(fabs(o->temp - ttar) > o->Teps)) return 1;
    o->lt = o->lt + 1;
    orc_onestep(o, 0, ttar, o->lt);
    return 0;
  }
  if (whichctrl == 1) {
    o->lt = o->lt + 1;
    if (o->time > (double)o->impact * o->dti) {
      o->atom_is_slow = 1;
      return 1;
    }
    orc_onestep(o, 1, ttar, o->lt);
    return 0;
  }
  o->lt = o->lt + 1;
  orc_onestep(o, whichctrl, ttar, o->lt);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Config mode (the benchmark configurations' integrator).  It has no
 * reference counterpart; it reuses the reference's force (si_force) and
 * list build (si_nlist), with removed atoms out of the lists and terms.
 * Per step, in this order:
 *   1. each live atom: Berendsen factor of the previous step on the shells
 *      (x/y sides and bottom, SW atoms, not pinned), half kick, drift,
 *      pinned atoms back to Xi with V = 0, rebox, removal above the surface
 *      (its kinetic energy to the ledger), displacement since the build;
 *   2. interaction energy of the atoms removed in this step with the atoms
 *      that stay, from the lists of the last build: U(with them) - U(without);
 *   3. rebuild when the largest plus the second largest displacement since
 *      the build (as float32 lengths) exceed the skin;
 *   4. forces, second half kick (pinned atoms stay at rest), |v|max, kinetic
 *      energy and the shells' sum m v^2 and count;
 *   5. time += dt, dt = min(dt_max, dx_max/|v|max), the shells' factor
 *      sqrt(1 + dt/tau (Ttar/T_shell - 1)) for the next step. */

static int md_in_shell(const orc *o, const double x[3]) {
  const double w = o->md.shell_width;
  return x[0] < w || x[0] > o->lb[0] - w || x[1] < w || x[1] > o->lb[1] - w ||
         x[2] < o->bot_z + w;
}

static double md_potential(orc *o) {
  double U;
  double *u = (double *)malloc(sizeof(double) * (size_t)o->natm);
  potential_impl(o, o->X, &U, u, o->nsg + o->impact);
  free(u);
  return U;
}

static void md_build(orc *o) {
  orc_si_nlist(o, o->X);
  memcpy(o->Xb, o->X, 3 * (size_t)o->natm * sizeof(double));
  o->builds++;
}

/* forces of the current state, |v|max, kinetic energy; dt from |v|max */
static void md_refresh(orc *o) {
  orc_si_force(o, o->X, o->F, NULL);
  double vmax = 0., ke = 0.;
  for (int i = 0; i < o->natm; ++i) {
    if (o->dead[i]) continue;
    double v2 = XA(o->V, i, 0) * XA(o->V, i, 0) + XA(o->V, i, 1) * XA(o->V, i, 1) +
                XA(o->V, i, 2) * XA(o->V, i, 2);
    double vm = sqrt(v2);
    if (vm > vmax) vmax = vm;
    ke += 0.5 * o->mass[i] * v2;
  }
  o->md_vmax = vmax;
  o->md_ke = ke;
  o->md_dt = vmax > 0. ? fmin(o->md.dt_max, o->md.dx_max / vmax) : o->md.dt_max;
}

int orc_md_begin(orc *o, const orc_md_params *p, double surface_z, double bottom_z) {
  const size_t n = (size_t)o->natm;
  o->md = *p;
  o->surf_z = surface_z;
  o->bot_z = bottom_z;
  o->skin = o->rc - 0.999 * o->al * o->sigmaSi;
  if (!o->dead) o->dead = (int32_t *)calloc(n, sizeof(int32_t));
  if (!o->Xb) o->Xb = (double *)calloc(3 * n, sizeof(double));
  o->builds = 0;
  md_build(o);
  o->md_time = 0.;
  o->md_steps = 0;
  o->md_lambda = 1.;
  o->e_bath = o->e_out = o->e_in = 0.;
  md_refresh(o);
  return 0;
}

int orc_md_fire(orc *o, int32_t ion, const double x[3], const double v[3]) {
  if (!o->dead || ion < 1 || ion > o->nlj) return 1;
  const int i = o->nsg + ion - 1;
  const double m = o->mass[i];
  double w2 = 0., v2 = 0.;
  for (int c = 0; c < 3; ++c) {
    w2 += XA(o->V, i, c) * XA(o->V, i, c);
    v2 += v[c] * v[c];
    XA(o->X, i, c) = x[c];
    XA(o->V, i, c) = v[c];
  }
  o->e_in += 0.5 * m * (v2 - w2);
  md_build(o);
  /* the ion's interaction energy at its start point */
  o->impact = ion - 1;
  double u0 = md_potential(o);
  o->impact = ion;
  double u1 = md_potential(o);
  o->e_in += u1 - u0;
  md_refresh(o);
  return 0;
}

int orc_md_step(orc *o) {
  const int n = o->natm;
  const double dt = o->md_dt, lam = o->md_lambda;
  double e_bath = 0., e_out = 0.;
  float *disp = (float *)calloc((size_t)n, sizeof(float));
  int32_t *leaving = (int32_t *)calloc((size_t)n, sizeof(int32_t));
  int nleave = 0;
  double dmax = 0.;
  for (int i = 0; i < n; ++i) {
    if (o->dead[i]) continue;
    const int t = o->atype[i];
    const double m = o->mass[i];
    const int pinned = o->xsides[i] > 0;
    double x[3], v[3];
    for (int c = 0; c < 3; ++c) { x[c] = XA(o->X, i, c); v[c] = XA(o->V, i, c); }
    if (o->md.shell_width > 0. && t < 3 && !pinned && md_in_shell(o, x)) {
      e_bath += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * (1. - lam * lam);
      for (int c = 0; c < 3; ++c) v[c] = v[c] * lam;
    }
    for (int c = 0; c < 3; ++c) {
      v[c] = v[c] + 0.5 * XA(o->F, i, c) * dt / m;
      x[c] = x[c] + v[c] * dt;
    }
    if (pinned)
      for (int c = 0; c < 3; ++c) { x[c] = XA(o->Xi, i, c); v[c] = 0.; }
    for (int c = 0; c < 3; ++c) x[c] = fmod(x[c] + 2. * o->lb[c], o->lb[c]);
    if (x[2] - o->surf_z > o->md.remove_height && x[2] < o->lb[2] - o->md.remove_height) {
      o->dead[i] = 1;
      leaving[i] = 1;
      ++nleave;
      e_out += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      v[0] = v[1] = v[2] = 0.;
    }
    for (int c = 0; c < 3; ++c) { XA(o->X, i, c) = x[c]; XA(o->V, i, c) = v[c]; }
    double e[3];
    for (int c = 0; c < 3; ++c) {
      double d = x[c] - XA(o->Xb, i, c);
      e[c] = d - (double)fnint(d * o->ilb[c]) * o->lb[c];
    }
    disp[i] = (float)sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    if ((double)disp[i] > dmax) dmax = (double)disp[i];
  }
  /* second largest displacement, or the largest when two atoms share it */
  double d2 = 0.;
  int nmax = 0;
  for (int i = 0; i < n; ++i) {
    double d = (double)disp[i];
    if (d == dmax) ++nmax;
    else if (d < dmax && d > d2) d2 = d;
  }
  if (nmax >= 2) d2 = dmax;
  if (nleave > 0) {
    /* U of the state with the leaving atoms still in, minus U without them */
    for (int i = 0; i < n; ++i) if (leaving[i]) o->dead[i] = 0;
    double u_with = md_potential(o);
    for (int i = 0; i < n; ++i) if (leaving[i]) o->dead[i] = 1;
    double u_without = md_potential(o);
    e_out += u_with - u_without;
  }
  free(leaving);
  free(disp);
  if (dmax + d2 > o->skin) md_build(o);
  orc_si_force(o, o->X, o->Fp, NULL);
  double vmax = 0., ke = 0., ksh = 0., nsh = 0.;
  for (int i = 0; i < n; ++i) {
    if (o->dead[i]) continue;
    const int t = o->atype[i];
    const double m = o->mass[i];
    const int pinned = o->xsides[i] > 0;
    double v[3], x[3];
    for (int c = 0; c < 3; ++c) {
      v[c] = pinned ? 0. : XA(o->V, i, c) + 0.5 * XA(o->Fp, i, c) * dt / m;
      XA(o->V, i, c) = v[c];
      x[c] = XA(o->X, i, c);
    }
    double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double vm = sqrt(v2);
    if (vm > vmax) vmax = vm;
    ke += 0.5 * m * v2;
    if (o->md.shell_width > 0. && t < 3 && !pinned && md_in_shell(o, x)) {
      ksh += m * v2;
      nsh += 1.;
    }
  }
  memcpy(o->F, o->Fp, 3 * (size_t)n * sizeof(double));
  o->md_vmax = vmax;
  o->md_ke = ke;
  o->md_time += dt;
  o->md_steps++;
  o->e_bath += e_bath;
  o->e_out += e_out;
  const double ndt = vmax > 0. ? fmin(o->md.dt_max, o->md.dx_max / vmax) : o->md.dt_max;
  o->md_dt = ndt;
  double l = 1.;
  if (o->md.shell_width > 0. && nsh > 0. && ksh > 0.) {
    double ts = ksh / (3. * nsh * kB);
    l = sqrt(1. + ndt / o->md.tau * (o->md.ttar / ts - 1.));
  }
  o->md_lambda = l;
  return 0;
}

void orc_md_report(orc *o, double *out) {
  double rem = 0., rem_ga = 0., disp = 0., live = 0., ionz = -1e300;
  const double r2 = o->md.displaced_r * o->md.displaced_r;
  for (int i = 0; i < o->natm; ++i) {
    const int t = o->atype[i];
    if (o->dead[i]) {
      if (t < 3) rem += 1.; else rem_ga += 1.;
    } else {
      live += 1.;
      if (t < 3) {
        double e[3];
        for (int c = 0; c < 3; ++c) {
          double d = XA(o->X, i, c) - XA(o->Xi, i, c);
          e[c] = d - (double)fnint(d * o->ilb[c]) * o->lb[c];
        }
        if (e[0] * e[0] + e[1] * e[1] + e[2] * e[2] > r2) disp += 1.;
      }
    }
    if (t == 3 && o->impact > 0 && i == o->nsg + o->impact - 1)
      ionz = o->dead[i] ? 1e300 : XA(o->X, i, 2);
  }
  out[0] = o->md_time; out[1] = o->md_dt; out[2] = o->md_lambda; out[3] = o->md_ke;
  out[4] = (double)o->md_steps; out[5] = (double)o->builds;
  out[6] = rem; out[7] = rem_ga; out[8] = disp;
  out[9] = o->e_bath; out[10] = o->e_out; out[11] = o->e_in;
  out[12] = ionz; out[13] = o->md_vmax; out[14] = md_potential(o); out[15] = live;
}

void orc_md_dead(const orc *o, int32_t *dead) {
  if (o->dead) memcpy(dead, o->dead, sizeof(int32_t) * (size_t)o->natm);
  else memset(dead, 0, sizeof(int32_t) * (size_t)o->natm);
}
