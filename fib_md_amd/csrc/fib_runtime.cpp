// fib_runtime.cpp -- host runtime behind include/fibgpu.h.
//
// Owns the device buffers of a handle, derives the reference's constants
// (prms.f90 setupSiGe/setupMol, si.f90 init_nlist/mktables), sequences the
// kernels of one neighbor rebuild / force evaluation / time step and mirrors
// the reference's step-control state (Ts, Tau, cadences, lt, time) on the
// host so that the Fortran or Python host can drive the reference's TI loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fibgpu.h"
#include "fib_device.h"
#include "fib_kernels.h"

namespace {

// Phase timers: 0 sort, 1 nbuild, 2 force_a, 3 force_b, 4 integrate.
struct Timer {
  static constexpr int NP = FIB_NTIMERS;
  bool on = false;
  hipStream_t s = nullptr;
  std::vector<hipEvent_t> pool;
  struct Rec { int key; hipEvent_t a, b; };
  std::vector<Rec> recs;
  size_t used = 0;
  double acc[NP] = {};
  int cnt[NP] = {};
  KTimer hook{};
  hipEvent_t get() {
    if (used == pool.size()) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      pool.push_back(e);
    }
    return pool[used++];
  }
  uint32_t mask = 0;  // timers that record (bit per key)
  int begin(int key) {
    if (!on || !((mask >> key) & 1u)) return -1;
    hipEvent_t a = get();
    (void)hipEventRecord(a, s);
    recs.push_back({-1, a, nullptr});
    return (int)recs.size() - 1;
  }
  void end(int id, int key) {
    if (!on || id < 0) return;
    hipEvent_t b = get();
    (void)hipEventRecord(b, s);
    recs[id].key = key;
    recs[id].b = b;
    ++cnt[key];
    if (recs.size() > 4096) flush();
  }
  void flush() {
    if (recs.empty()) return;
    (void)hipStreamSynchronize(s);
    for (auto &r : recs) {
      float ms = 0;
      if (r.key >= 0 && r.b && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) acc[r.key] += ms;
    }
    recs.clear();
    used = 0;
  }
  void reset() {
    flush();
    for (int k = 0; k < NP; ++k) { acc[k] = 0; cnt[k] = 0; }
  }
  static int begin_cb(void *c, int key) { return static_cast<Timer *>(c)->begin(key); }
  static void end_cb(void *c, int id, int key) { static_cast<Timer *>(c)->end(id, key); }
};

static const char *const kTimerNames[KT_N] = {
    "k_bin", "sort", "k_nbuild", "list_aux", "k_legs",
    "k_tri", "k_force3", "k_extra", "k_pos", "reduce", "io", "md_step"};
static_assert(KT_N <= FIB_NTIMERS, "timer table");

struct RefCtl {
  bool have_run = false;
  fib_run_params rp{};
  int64_t lt = 0;
  double time = 0, temp = 0, ts = 0, tau = 0;
  int regime = 1;
  int atm_out = 0, ntlist = 1, atlist = 1, dslist = 1;
  int impact = 0;
  bool fast = true;
  bool begun = false;
  int whichctrl = 1;
};

}  // namespace

struct fib_handle {
  int dev = 0;
  hipStream_t stream = nullptr;
  int prec = FIB_PREC_FP64;
  fib_params P{};
  FConst fc{};
  std::vector<TabEntry> tab;
  std::vector<fib_instance_desc> desc;
  std::vector<DInst> inst;
  DevBufs d{};
  bool have_state = false, have_lists = false, have_force = false;
  // MD-step mode
  bool md_on = false;
  fib_md_params md{};
  StepArgs md_sa{};
  int *md_gate = nullptr;       // device [2 + 2*ninst]: rebuild gates of the step and
                                // the cell prefix of the gated instances
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool graph_valid = false;
  int last_flags = 0;
  std::string err;
  RefCtl ref;
  Timer tim;
  StepArgs base{};
  int n_builds = 0;
  // fib_md_run_begin / fib_md_run_end: steps enqueued, not yet checked
  bool run_pending = false;
  double run_target = 0;
  // asynchronous snapshots: per slot a device staging buffer (X then V,
  // (natm,3) each) and its pinned host copy, filled on copy_stream
  struct Snap {
    double *dev = nullptr, *host = nullptr;
    size_t cap = 0, n = 0;  // doubles held; atoms of the pending snapshot
    hipEvent_t ready = nullptr, copied = nullptr;
    bool pending = false;
  } snap[FIB_SNAPSHOT_SLOTS];
  hipStream_t copy_stream = nullptr;
};

#define HIPCHK(h, x)                                                              \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      (h)->err = std::string(#x) + ": " + hipGetErrorString(e_);                  \
      return FIB_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)

static int32_t fail(fib_handle *h, int32_t code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}

extern "C" void fib_default_params(fib_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->sw_A = 7.049556277;
  p->sw_B = 0.6022245584;
  p->sw_p = 4.0;
  p->sw_q = 0.0;
  p->sw_a = 1.8;
  p->sw_lambda_si = 21.0;
  p->sw_lambda_ge = 31.0;
  p->sw_gamma = 1.2;
  p->eps_si = 3.473928E-19;
  p->eps_ge = 3.085E-19;
  p->sigma_si = 2.0951E-10;
  p->sigma_ge = 2.0951E-10;
  p->mass_si = 46.637063E-27;
  p->mass_ge = 120.6049350E-27;
  p->mass_ga = 1.157777417E-25;
  p->z_si = 14.0;
  p->z_ge = 32.0;
  p->z_ga = 31.0;
  p->mol_c[0] = 0.35; p->mol_c[1] = 0.55; p->mol_c[2] = 0.10;
  p->mol_b[0] = 0.3; p->mol_b[1] = 1.2; p->mol_b[2] = 6.0;
  p->bohr = 0.053E-09;
  p->kB = 1.381E-23;
  p->eps0 = 8.854187E-12;
  p->ec = 1.60217646E-19;
  p->listfac = 1.1;
  p->ntable = 12000;
  p->mixed_fp64_moliere_r = 0.5E-10;
}

// Mixing tables (setupSiGe, prms.f90:452-480; setupMol :433-449), cutoffs
// (init_nlist, si.f90:56-61) and the pair-derivative table (mktables,
// si.f90:575-597).  Expressions keep the reference's operation order.
static void derive_constants(fib_handle *h) {
  const fib_params &p = h->P;
  FConst &c = h->fc;
  std::memset(&c, 0, sizeof(c));
  const double eSi = p.eps_si, eGe = p.eps_ge, sSi = p.sigma_si, sGe = p.sigma_ge;
  const double lSi = p.sw_lambda_si, lGe = p.sw_lambda_ge;
  c.eps[0][0] = eSi; c.eps[0][1] = 0.5 * (eSi + eGe); c.eps[1][1] = eGe; c.eps[1][0] = 0.5 * (eSi + eGe);
  c.isig[0][0] = 1. / sSi; c.isig[0][1] = 2. / (sSi + sGe); c.isig[1][1] = 1. / sGe;
  c.isig[1][0] = 2. / (sSi + sGe);
  const double m1 = 2. * lSi * eSi / 3. + lGe * eGe / 3.;
  const double m2 = lSi * eSi / 3. + 2. * lGe * eGe / 3.;
  c.leps[0][0][0] = lSi * eSi;
  c.leps[0][0][1] = m1; c.leps[0][1][0] = m1; c.leps[1][0][0] = m1;
  c.leps[0][1][1] = m2; c.leps[1][1][0] = m2; c.leps[1][0][1] = m2;
  c.leps[1][1][1] = lGe * eGe;
  c.epssig[0][0] = eSi / sSi;
  c.epssig[0][1] = (eSi + eGe) / (sSi + sGe);
  c.epssig[1][0] = (eSi + eGe) / (sSi + sGe);
  c.epssig[1][1] = eGe / sGe;
  c.A = p.sw_A; c.B = p.sw_B; c.p = p.sw_p; c.q = p.sw_q; c.al = p.sw_a; c.gamma = p.sw_gamma;
  c.cut_sw = 0.999 * p.sw_a;
  const double smax = std::max(sGe, sSi);
  const double rc = p.listfac * p.sw_a * smax;
  double rlj = 2.5 * smax;
  const double rclj = p.listfac * rlj;
  c.rv2 = rc * rc;
  c.rv2lj = rclj * rclj;
  c.rlj2 = rlj * rlj;
  for (int k = 0; k < 3; ++k) { c.molc[k] = p.mol_c[k]; c.molb[k] = p.mol_b[k]; }
  const double Pi = 4. * std::atan(1.0);
  const double Z[3] = {p.z_si, p.z_ge, p.z_ga};
  for (int k = 0; k < 3; ++k) c.zmol[k] = Z[k];
  c.aamol[0][2] = 0.885 * p.bohr * std::pow(std::sqrt(p.z_si) + std::sqrt(p.z_ga), -2. / 3.);
  c.aamol[1][2] = 0.885 * p.bohr * std::pow(std::sqrt(p.z_ge) + std::sqrt(p.z_ga), -2. / 3.);
  c.aamol[2][0] = c.aamol[0][2];
  c.aamol[2][1] = c.aamol[1][2];
  c.aamol[2][2] = 0.885 * p.bohr * std::pow(std::sqrt(p.z_ga) + std::sqrt(p.z_ga), -2. / 3.);
  c.ie0 = 1 / (4 * Pi * p.eps0);
  c.ec = p.ec;
  c.ie0ec2 = c.ie0 * p.ec * p.ec;
  c.mixed_r64 = p.mixed_fp64_moliere_r;
  // table (1-based abscissas u(i) = umax*(i-1)/(Ntable-1))
  const int nt = p.ntable;
  c.ntable = nt;
  c.umax = 2.5 * p.listfac * p.sw_a;
  c.tabn1 = (double)(nt - 1);
  const double du = c.umax / (double)(nt - 1);
  std::vector<double> u(nt + 4, 0.0), T(nt + 4, 0.0);
  int izero = nt + 1;
  for (int i = 2; i <= nt; ++i) {
    const double ui = c.umax * (double)(i - 1) / (double)(nt - 1);
    u[i] = ui;
    if (ui < .999 * p.sw_a) {
      const double t1 = (p.sw_B * std::pow(ui, -p.sw_p) - std::pow(ui, -p.sw_q)) / ((ui - p.sw_a) * (ui - p.sw_a));
      const double t2 = p.sw_p * p.sw_B * std::pow(ui, -p.sw_p - 1) - p.sw_q * std::pow(ui, -p.sw_q - 1);
      T[i] = -(p.sw_A * std::exp(1. / (ui - p.sw_a)) * (t1 + t2) / (du * du));
    } else {
      T[i] = 0.;
      if (izero > nt) izero = i;
    }
  }
  c.izero = izero;
  c.cut_sw2 = c.cut_sw * c.cut_sw * (1.0 + 1e-6);
  c.mixed_xleg = std::min(c.cut_sw, c.al - c.gamma / 30.0);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      c.rho[a][b] = (float)(c.isig[a][b] / c.isig[0][0]);
      c.fsw[a][b] = (float)(c.epssig[a][b] / (c.eps[0][0] * c.isig[0][0]));
    }
  c.skin = rc - c.cut_sw * std::max(p.sigma_si, p.sigma_ge);
  {
    const double ug = std::max(c.umax * (double)(izero - 1) / (double)(nt - 1), c.cut_sw);
    c.r2gate = (1.001 * ug) * (1.001 * ug);
  }
  h->tab.assign(nt + 3, TabEntry{0, 0, 0, 0, 0, 0});
  for (int i = 1; i <= nt; ++i) {
    TabEntry &e = h->tab[i];
    e.u0 = u[i]; e.u1 = u[i + 1]; e.u2 = u[i + 2];
    e.t0 = T[i]; e.t1 = T[i + 1]; e.t2 = T[i + 2];
  }
  StepArgs &b = h->base;
  std::memset(&b, 0, sizeof(b));
  b.mass[0] = p.mass_si; b.mass[1] = p.mass_ge; b.mass[2] = p.mass_ga;
  b.kB = p.kB;
  b.ec = p.ec;
}

static int32_t free_bufs(fib_handle *h) {
  DevBufs &d = h->d;
  void *ptrs[] = {d.I, d.ctl, d.ibase, d.binst, d.X[0], d.X[1], d.V[0], d.V[1], d.F[0], d.F[1], d.Xi[0], d.Xi[1],
                  d.inv, d.bkey, d.brank, d.bcnt, d.cstart, d.perm, d.bsum, d.cnt_sw, d.nbr_sw,
                  d.cnt_ion, d.nbr_ion, d.rk_sw, d.off_sw, d.G, d.RA, d.amask, d.cnt_leg, d.Xb, d.disp, d.EX, d.EXself, d.exmask, d.Fpart, d.upart, d.Fout, d.uout,
                  d.partials, d.fin_tmp, d.tab, d.xfer};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  std::memset(&d, 0, sizeof(d));
  return FIB_OK;
}

template <typename T>
static hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
}

static int32_t alloc_lists(fib_handle *h, int Wsw, int Wion) {
  DevBufs &d = h->d;
  if (d.nbr_sw) (void)hipFree(d.nbr_sw);
  if (d.nbr_ion) (void)hipFree(d.nbr_ion);
  void *old[] = {d.rk_sw, d.G, d.RA, d.EX};
  for (void *q : old)
    if (q) (void)hipFree(q);
  d.nbr_sw = nullptr; d.nbr_ion = nullptr; d.rk_sw = nullptr;
  d.G = nullptr; d.RA = nullptr;
  d.EX = nullptr;
  d.Wsw = Wsw;
  d.Wion = Wion;
  const size_t NA = (size_t)d.NA;
  // leg lookups (RA) are 32-bit indices into the slot-major force rows
  if ((uint64_t)Wsw * NA >= (1ull << 32))
    return fail(h, FIB_ERR_ARG, "too many atoms for one handle: atoms x list width >= 2^32");
  HIPCHK(h, dalloc(&d.nbr_sw, (size_t)Wsw * NA));
  HIPCHK(h, dalloc(&d.nbr_ion, (size_t)Wion * NA));
  HIPCHK(h, dalloc(&d.rk_sw, (size_t)Wsw * NA));
  // a leg is a listed pair, so Wsw slots always hold every leg of an atom
  const size_t gsz = h->prec == FIB_PREC_MIXED ? sizeof(float4) : sizeof(double4);
  HIPCHK(h, hipMalloc(&d.G, gsz * (size_t)Wsw * NA));
  HIPCHK(h, dalloc(&d.RA, (size_t)Wsw * NA));
  HIPCHK(h, hipMalloc(&d.EX, gsz * (size_t)Wsw * NA));
  return FIB_OK;
}

// Wider lists that keep what the current ones hold (rows copied into the
// new layout): instances that are not rebuilt keep their lists.  The
// per-evaluation arrays (G, RA, EX) are reallocated empty.
static int32_t regrow_lists(fib_handle *h, int Wsw, int Wion) {
  DevBufs &d = h->d;
  const size_t NA = (size_t)d.NA;
  if ((uint64_t)Wsw * NA >= (1ull << 32))
    return fail(h, FIB_ERR_ARG, "too many atoms for one handle: atoms x list width >= 2^32");
  int *nsw = nullptr, *nion = nullptr;
  uint16_t *rk = nullptr;
  HIPCHK(h, dalloc(&nsw, (size_t)Wsw * NA));
  HIPCHK(h, dalloc(&nion, (size_t)Wion * NA));
  HIPCHK(h, dalloc(&rk, (size_t)Wsw * NA));
  fibk::launch_copy_rows(d.NA, d.Wsw, d.nbr_sw, Wsw, nsw, h->stream);
  fibk::launch_copy_rows(d.NA, d.Wsw, d.rk_sw, Wsw, rk, h->stream);
  fibk::launch_copy_rows(d.NA, d.Wion, d.nbr_ion, Wion, nion, h->stream);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  void *old[] = {d.nbr_sw, d.nbr_ion, d.rk_sw, d.G, d.RA, d.EX};
  for (void *q : old)
    if (q) (void)hipFree(q);
  d.nbr_sw = nsw; d.nbr_ion = nion; d.rk_sw = rk;
  d.G = nullptr; d.RA = nullptr; d.EX = nullptr;
  d.Wsw = Wsw;
  d.Wion = Wion;
  const size_t gsz = h->prec == FIB_PREC_MIXED ? sizeof(float4) : sizeof(double4);
  HIPCHK(h, hipMalloc(&d.G, gsz * (size_t)Wsw * NA));
  HIPCHK(h, dalloc(&d.RA, (size_t)Wsw * NA));
  HIPCHK(h, hipMalloc(&d.EX, gsz * (size_t)Wsw * NA));
  h->graph_valid = false;  // the graph refers to the old arrays
  return FIB_OK;
}

// Derived layout of one instance (cells, buckets, padding) after the ones
// before it; false when the description is inconsistent.
static bool layout_instance(double rc, const fib_instance_desc &ds, long &atom_off,
                            long &bucket_off, long &cell_off, long &natm_before, DInst &in) {
  if (ds.natm < 1 || ds.nsg < 0 || ds.nlj < 0 || ds.nsg + ds.nlj != ds.natm ||
      !(ds.lb[0] > 0 && ds.lb[1] > 0 && ds.lb[2] > 0))
    return false;
  std::memset(&in, 0, sizeof(in));
  for (int a = 0; a < 3; ++a) {
    in.lb[a] = ds.lb[a];
    in.ilb[a] = 1. / ds.lb[a];
    const int n = (int)(ds.lb[a] / rc);
    in.nc[a] = n > 1 ? n : 1;
    in.h[a] = ds.lb[a] / (double)in.nc[a];
    in.next[a] = in.nc[a] <= 2 ? in.nc[a] + 1 : in.nc[a];
  }
  in.natm = ds.natm;
  in.nsg = ds.nsg;
  in.nlj = ds.nlj;
  in.natm_pad = (ds.natm + FIB_BLOCK - 1) / FIB_BLOCK * FIB_BLOCK;
  in.atom_off = (int)atom_off;
  in.pad_before = (int)(atom_off - natm_before);
  in.bucket_off = (int)bucket_off;
  in.nbucket = in.next[0] * in.next[1] * in.next[2] + 2;
  in.cell_real_off = (int)cell_off;
  in.block_off = (int)(atom_off / FIB_BLOCK);
  in.nblocks = in.natm_pad / FIB_BLOCK;
  atom_off += in.natm_pad;
  natm_before += ds.natm;
  bucket_off += in.nbucket;
  cell_off += (long)in.nc[0] * in.nc[1] * in.nc[2];
  // the list build keeps atom indices in 27 bits beside a stencil index
  return atom_off < (1L << 27) && bucket_off <= (1L << 30);
}

extern "C" int32_t fib_create(const fib_params *p, int32_t ninst, const fib_instance_desc *desc,
                              int32_t precision, fib_handle **out) {
  if (!out) return FIB_ERR_ARG;
  *out = nullptr;
  if (!p || ninst < 1 || !desc) return FIB_ERR_ARG;
  if (precision != FIB_PREC_FP64 && precision != FIB_PREC_MIXED) return FIB_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return FIB_ERR_NODEVICE;
  fib_handle *h = new fib_handle();
  if (hipGetDevice(&h->dev) != hipSuccess) { delete h; return FIB_ERR_NODEVICE; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, h->dev) != hipSuccess ||
      std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    delete h;
    return FIB_ERR_NODEVICE;
  }
  h->P = *p;
  h->prec = precision;
  derive_constants(h);
  const double smax = std::max(p->sigma_ge, p->sigma_si);
  const double rc = p->listfac * p->sw_a * smax;
  h->desc.assign(desc, desc + ninst);
  h->inst.resize(ninst);
  long atom_off = 0, bucket_off = 0, cell_off = 0, natm_before = 0;
  for (int k = 0; k < ninst; ++k) {
    if (!layout_instance(rc, desc[k], atom_off, bucket_off, cell_off, natm_before, h->inst[k])) {
      delete h;
      return FIB_ERR_ARG;
    }
  }
  DevBufs &d = h->d;
  d.NA = (int)atom_off;
  d.ninst = ninst;
  d.nbucket = (int)bucket_off;
  d.total_cells = (int)cell_off;
  const size_t NA = (size_t)d.NA;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return FIB_ERR_HIP;
  }
  h->tim.s = h->stream;
  h->tim.hook = KTimer{&h->tim, &Timer::begin_cb, &Timer::end_cb};
  h->d.kt = &h->tim.hook;
  int32_t rc_ = FIB_OK;
#define ALLOC(ptr, n)                                           \
  if (rc_ == FIB_OK && dalloc(&(ptr), (n)) != hipSuccess) {     \
    rc_ = FIB_ERR_HIP;                                          \
    h->err = "hipMalloc failed for " #ptr;                      \
  }
  ALLOC(d.I, (size_t)ninst);
  ALLOC(d.ctl, (size_t)ninst);
  for (int c = 0; c < 2; ++c) {
    ALLOC(d.X[c], NA);
    ALLOC(d.V[c], 3 * NA);
    ALLOC(d.F[c], 3 * NA);
    ALLOC(d.Xi[c], 3 * NA);
  }
  ALLOC(d.inv, NA);
  ALLOC(d.ibase, (size_t)ninst);
  ALLOC(d.binst, (size_t)NA / FIB_BLOCK);
  ALLOC(d.bkey, NA);
  ALLOC(d.brank, NA);
  ALLOC(d.bcnt, (size_t)d.nbucket);
  ALLOC(d.cstart, (size_t)d.nbucket + 1);
  ALLOC(d.perm, NA);
  ALLOC(d.bsum, (size_t)d.nbucket / 2048 + 2);
  ALLOC(d.cnt_sw, NA);
  ALLOC(d.cnt_ion, NA);
  ALLOC(d.amask, NA);
  ALLOC(d.cnt_leg, NA);
  ALLOC(d.off_sw, (size_t)32 * NA);
  ALLOC(d.Xb, NA);
  ALLOC(d.disp, NA);
  ALLOC(d.exmask, NA);
  if (rc_ == FIB_OK &&
      hipMalloc(&d.EXself, (h->prec == FIB_PREC_MIXED ? sizeof(float4) : sizeof(double4)) * NA) !=
          hipSuccess) {
    rc_ = FIB_ERR_HIP;
    h->err = "hipMalloc failed for d.EXself";
  }
  ALLOC(d.Fpart, 3 * NA);
  ALLOC(d.upart, NA);
  ALLOC(d.Fout, 3 * NA);
  ALLOC(d.uout, NA);
  d.pstride = d.NA / FIB_WAVE + 1;
  ALLOC(d.partials, (size_t)d.pstride * FIB_NRED * 3);
  d.max_pad = 0;
  for (const DInst &in : h->inst) d.max_pad = std::max(d.max_pad, in.natm_pad);
  ALLOC(d.fin_tmp, (size_t)ninst * (d.max_pad / FIB_WAVE / FIB_FIN_T + 1) * FIB_NRED);
  ALLOC(d.tab, h->tab.size());
  ALLOC(d.xfer, 3 * NA);
#undef ALLOC
  if (rc_ == FIB_OK) rc_ = alloc_lists(h, 32, 64);
  if (rc_ != FIB_OK) {
    free_bufs(h);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return rc_;
  }
  hipError_t e = hipMemcpy(d.I, h->inst.data(), sizeof(DInst) * ninst, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    std::vector<int> binst((size_t)NA / FIB_BLOCK);
    for (int k = 0; k < ninst; ++k)
      for (int b = 0; b < h->inst[k].nblocks; ++b) binst[(size_t)h->inst[k].block_off + b] = k;
    e = hipMemcpy(d.binst, binst.data(), sizeof(int) * binst.size(), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipMemset(d.ctl, 0, sizeof(DCtl) * ninst);
  if (e == hipSuccess) e = hipMemset(d.cnt_sw, 0, sizeof(int) * NA);
  if (e == hipSuccess) e = hipMemset(d.cnt_ion, 0, sizeof(int) * NA);
  if (e == hipSuccess) e = hipMemset(d.amask, 0, sizeof(uint64_t) * NA);
  if (e == hipSuccess) e = hipMemset(d.cnt_leg, 0, sizeof(int) * NA);
  if (e == hipSuccess) e = hipMemset(d.exmask, 0, sizeof(uint64_t) * NA);
  if (e == hipSuccess) e = hipMemset(d.off_sw, 0, (size_t)32 * NA);
  if (e == hipSuccess) e = hipMemset(d.disp, 0, sizeof(float) * NA);
  if (e == hipSuccess)
    e = hipMemcpy(d.tab, h->tab.data(), sizeof(TabEntry) * h->tab.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    free_bufs(h);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return FIB_ERR_HIP;
  }
  // every slot starts as padding of its instance so that all kernels see a
  // consistent meta word before fib_set_state
  for (int k = 0; k < ninst; ++k) {
    DInst in = h->inst[k];
    in.natm = 0;
    fibk::launch_import(d, k, in, nullptr, nullptr, nullptr, nullptr, nullptr, h->stream);
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    free_bufs(h);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return FIB_ERR_HIP;
  }
  *out = h;
  return FIB_OK;
}

extern "C" int32_t fib_destroy(fib_handle *h) {
  if (!h) return FIB_ERR_ARG;
  (void)hipStreamSynchronize(h->stream);
  if (h->copy_stream) {
    (void)hipStreamSynchronize(h->copy_stream);
    (void)hipStreamDestroy(h->copy_stream);
  }
  for (auto &sn : h->snap) {
    if (sn.dev) (void)hipFree(sn.dev);
    if (sn.host) (void)hipHostFree(sn.host);
    if (sn.ready) (void)hipEventDestroy(sn.ready);
    if (sn.copied) (void)hipEventDestroy(sn.copied);
  }
  for (auto e : h->tim.pool) (void)hipEventDestroy(e);
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->md_gate) (void)hipFree(h->md_gate);
  free_bufs(h);
  (void)hipStreamDestroy(h->stream);
  delete h;
  return FIB_OK;
}

// Message of the calling thread's last failed handle-free call.
static thread_local std::string g_free_err = "null handle";

extern "C" const char *fib_last_error(const fib_handle *h) {
  return h ? h->err.c_str() : g_free_err.c_str();
}

extern "C" int32_t fib_get_cells(const fib_handle *h, int32_t inst, int32_t nc[3]) {
  if (!h || inst < 0 || inst >= (int)h->inst.size() || !nc) return FIB_ERR_ARG;
  for (int a = 0; a < 3; ++a) nc[a] = h->inst[inst].nc[a];
  return FIB_OK;
}

static bool bad_inst(const fib_handle *h, int32_t inst) {
  return !h || inst < 0 || inst >= (int)h->inst.size();
}

// Steps enqueued by fib_md_run_begin are not checked yet: calls that read
// or change the state must wait for fib_md_run_end.
static bool run_in_flight(fib_handle *h, const char *who) {
  if (!h || !h->run_pending) return false;
  h->err = std::string(who) + ": steps in flight; call fib_md_run_end first";
  return true;
}

extern "C" int32_t fib_set_state(fib_handle *h, int32_t inst, const double *X, const double *V,
                                 const int32_t *atype, const int32_t *xsides, const double *Xi) {
  if (run_in_flight(h, "fib_set_state")) return FIB_ERR_STATE;
  if (bad_inst(h, inst) || !X || !atype) return fail(h, FIB_ERR_ARG, "fib_set_state: bad argument");
  const DInst &in = h->inst[inst];
  const size_t n = (size_t)in.natm;
  for (size_t i = 0; i < n; ++i) {
    const int t = atype[i];
    const bool ion = i >= (size_t)in.nsg;
    if (t < 1 || t > 3 || (ion != (t == 3)))
      return fail(h, FIB_ERR_ARG, "fib_set_state: atype must be 1/2 for atoms 1..nsg and 3 for ions");
  }
  double *dx = nullptr, *dv = nullptr, *dxi = nullptr;
  int32_t *dt = nullptr, *ds = nullptr;
  HIPCHK(h, dalloc(&dx, 3 * n));
  HIPCHK(h, hipMemcpy(dx, X, 3 * n * sizeof(double), hipMemcpyHostToDevice));
  if (V) {
    HIPCHK(h, dalloc(&dv, 3 * n));
    HIPCHK(h, hipMemcpy(dv, V, 3 * n * sizeof(double), hipMemcpyHostToDevice));
  }
  if (Xi) {
    HIPCHK(h, dalloc(&dxi, 3 * n));
    HIPCHK(h, hipMemcpy(dxi, Xi, 3 * n * sizeof(double), hipMemcpyHostToDevice));
  }
  HIPCHK(h, dalloc(&dt, n));
  HIPCHK(h, hipMemcpy(dt, atype, n * sizeof(int32_t), hipMemcpyHostToDevice));
  if (xsides) {
    HIPCHK(h, dalloc(&ds, n));
    HIPCHK(h, hipMemcpy(ds, xsides, n * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  fibk::launch_import(h->d, inst, in, dx, dv, dt, ds, dxi, h->stream);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  (void)hipFree(dx);
  if (dv) (void)hipFree(dv);
  if (dxi) (void)hipFree(dxi);
  (void)hipFree(dt);
  if (ds) (void)hipFree(ds);
  h->have_state = true;
  h->md_on = false;  // MD-step mode restarts from fib_md_begin
  h->have_lists = false;
  h->have_force = false;
  return FIB_OK;
}

extern "C" int32_t fib_get_state(fib_handle *h, int32_t inst, double *X, double *V, double *F) {
  if (bad_inst(h, inst)) return fail(h, FIB_ERR_ARG, "fib_get_state: bad instance");
  if (run_in_flight(h, "fib_get_state")) return FIB_ERR_STATE;
  if (!h->have_state) return fail(h, FIB_ERR_STATE, "fib_get_state: no state");
  const DInst &in = h->inst[inst];
  const size_t n = (size_t)in.natm;
  DevBufs &d = h->d;
  if (X) {
    fibk::launch_export_x(d, inst, in, d.xfer, h->stream);
    HIPCHK(h, hipMemcpyAsync(X, d.xfer, 3 * n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  if (V) {
    fibk::launch_export_vec(d, inst, in, d.V[d.cur], d.xfer, h->stream);
    HIPCHK(h, hipMemcpyAsync(V, d.xfer, 3 * n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  if (F) {
    fibk::launch_export_vec(d, inst, in, d.F[d.cur], d.xfer, h->stream);
    HIPCHK(h, hipMemcpyAsync(F, d.xfer, 3 * n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  return FIB_OK;
}

extern "C" int32_t fib_set_impact(fib_handle *h, int32_t inst, int32_t impact) {
  if (run_in_flight(h, "fib_set_impact")) return FIB_ERR_STATE;
  if (bad_inst(h, inst) || impact < 0) return fail(h, FIB_ERR_ARG, "fib_set_impact: bad argument");
  h->inst[inst].impact = impact;
  HIPCHK(h, hipMemcpyAsync(&h->d.I[inst], &h->inst[inst], sizeof(DInst), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FIB_OK;
}

extern "C" int32_t fib_bin_cells(fib_handle *h, int32_t inst, int32_t *icell) {
  if (run_in_flight(h, "fib_bin_cells")) return FIB_ERR_STATE;
  if (bad_inst(h, inst) || !icell) return fail(h, FIB_ERR_ARG, "fib_bin_cells: bad argument");
  if (!h->have_state) return fail(h, FIB_ERR_STATE, "fib_bin_cells: no state");
  const size_t n = (size_t)h->inst[inst].natm;
  int32_t *dout = nullptr;
  HIPCHK(h, dalloc(&dout, 3 * n));
  fibk::launch_cellid(h->d, inst, dout, h->stream);
  HIPCHK(h, hipMemcpyAsync(icell, dout, 3 * n * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  (void)hipFree(dout);
  return FIB_OK;
}

// Sort + list build.  Grows the ELL widths and rebuilds when a list
// overflowed (the reference only prints, si.f90:264-272).
static int32_t build_lists(fib_handle *h) {
  DevBufs &d = h->d;
  fibk::launch_bin(d, h->stream);
  fibk::launch_scan(d, h->stream);
  fibk::launch_sort(d, h->stream);
  d.cur = 1 - d.cur;
  for (int attempt = 0; attempt < 4; ++attempt) {
    // k_nbuild reports max list lengths into DCtl::maxcnt_*
    fibk::launch_nbuild(d, h->fc, h->stream);
    std::vector<DCtl> c(d.ninst);
    HIPCHK(h, hipMemcpyAsync(c.data(), d.ctl, sizeof(DCtl) * d.ninst, hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int msw = 0, mion = 0, nonfin = 0;
    for (auto &x : c) {
      msw = std::max(msw, (int)x.maxcnt_sw);
      mion = std::max(mion, (int)x.maxcnt_ion);
      nonfin |= x.nonfinite;
    }
    if (nonfin) {
      for (int k = 0; k < d.ninst; ++k) {
        int zero = 0;
        HIPCHK(h, hipMemcpy(&d.ctl[k].nonfinite, &zero, sizeof(int), hipMemcpyHostToDevice));
      }
    }
    if (nonfin) return fail(h, FIB_ERR_NONFINITE, "non-finite position detected while binning");
    if (msw <= d.Wsw && mion <= d.Wion) {
      fibk::launch_mark_build(d, h->stream);
      h->graph_valid = false;  // the sort moved the arrays a graph refers to
      h->have_lists = true;
      h->have_force = false;
      ++h->n_builds;
      return FIB_OK;
    }
    int nsw = d.Wsw, nion = d.Wion;
    while (nsw < msw) nsw *= 2;
    while (nion < mion) nion *= 2;
    if (nsw > 64) return fail(h, FIB_ERR_OVERFLOW, "SW neighbor list exceeds 64 entries per atom");
    int32_t r = alloc_lists(h, nsw, nion);
    if (r != FIB_OK) return r;
  }
  return fail(h, FIB_ERR_OVERFLOW, "neighbor list did not fit after regrowth");
}

extern "C" int32_t fib_build_nlist(fib_handle *h) {
  if (run_in_flight(h, "fib_build_nlist")) return FIB_ERR_STATE;
  if (!h) return FIB_ERR_ARG;
  if (!h->have_state) return fail(h, FIB_ERR_STATE, "fib_build_nlist: no state");
  return build_lists(h);
}

extern "C" int32_t fib_get_nlist(fib_handle *h, int32_t inst, int32_t *nss, int32_t *nlist2,
                                 int32_t *len2, int32_t *nsslj, int32_t *nlistlj,
                                 int32_t *lenlj) {
  if (run_in_flight(h, "fib_get_nlist")) return FIB_ERR_STATE;
  if (bad_inst(h, inst)) return fail(h, FIB_ERR_ARG, "fib_get_nlist: bad instance");
  if (!h->have_lists) return fail(h, FIB_ERR_STATE, "fib_get_nlist: no lists");
  const DInst &in = h->inst[inst];
  const DevBufs &d = h->d;
  const int n = in.natm, np = in.natm_pad, off = in.atom_off;
  std::vector<double4> X(np);
  std::vector<int> csw(np), cio(np);
  HIPCHK(h, hipMemcpy(X.data(), d.X[d.cur] + off, sizeof(double4) * np, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(csw.data(), d.cnt_sw + off, sizeof(int) * np, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(cio.data(), d.cnt_ion + off, sizeof(int) * np, hipMemcpyDeviceToHost));
  std::vector<int> orig(np, -1);
  for (int q = 0; q < np; ++q) {
    uint64_t w;
    std::memcpy(&w, &X[q].w, 8);
    if (!((uint32_t)w & META_PAD)) orig[q] = (int)(w >> 32);
  }
  auto fetch = [&](const int *lst, int W, const std::vector<int> &cnt, std::vector<std::vector<int>> &half) -> int32_t {
    std::vector<int> rows((size_t)W * np);
    hipError_t e = hipMemcpy(rows.data(), lst + (size_t)off * W, sizeof(int) * (size_t)W * np,
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) { h->err = hipGetErrorString(e); return FIB_ERR_HIP; }
    half.assign(n, {});
    for (int q = 0; q < np; ++q) {
      if (orig[q] < 0) continue;
      const int c = std::min(cnt[q], W);
      for (int k = 0; k < c; ++k) {
        const int j = rows[(size_t)q * W + k] - off;
        if (j < 0 || j >= np || orig[j] < 0) continue;
        if (orig[j] > orig[q]) half[orig[q]].push_back(orig[j] + 1);
      }
    }
    return FIB_OK;
  };
  std::vector<std::vector<int>> h2, hl;
  int32_t r = fetch(d.nbr_sw, d.Wsw, csw, h2);
  if (r != FIB_OK) return r;
  r = fetch(d.nbr_ion, d.Wion, cio, hl);
  if (r != FIB_OK) return r;
  auto emit = [&](const std::vector<std::vector<int>> &half, int32_t *ss, int32_t *lst, int32_t *len) {
    int pos = 0;
    for (int i = 0; i < n; ++i) {
      if (ss) { ss[i] = pos + 1; ss[n + i] = pos + (int)half[i].size(); }
      if (lst)
        for (int v : half[i]) lst[pos++] = v;
      else
        pos += (int)half[i].size();
    }
    if (len) *len = pos;
  };
  emit(h2, nss, nlist2, len2);
  emit(hl, nsslj, nlistlj, lenlj);
  return FIB_OK;
}

// mode 0: forces into dst (default d.Fout); mode 1: MD velocity update.
static int32_t run_force(fib_handle *h, int flags, int mode, const StepArgs &sa,
                         double *dst = nullptr) {
  const bool en = (flags & FIB_EVAL_ENERGY) != 0, vir = (flags & FIB_EVAL_VIRIAL) != 0;
  const bool mixed = h->prec == FIB_PREC_MIXED;
  DevBufs &d = h->d;
  double *Fo = mode == 0 ? (dst ? dst : d.Fout) : d.F[d.cur];
  fibk::launch_force_a(d, h->fc, mixed, en, vir, h->stream);
  fibk::launch_force_b(d, h->fc, mixed, en && mode == 0, mode, sa, Fo, h->stream);
  return FIB_OK;
}

extern "C" int32_t fib_compute(fib_handle *h, int32_t flags) {
  if (run_in_flight(h, "fib_compute")) return FIB_ERR_STATE;
  if (!h) return FIB_ERR_ARG;
  if (!h->have_lists) return fail(h, FIB_ERR_STATE, "fib_compute: build the lists first");
  int32_t r = run_force(h, flags, 0, h->base);
  if (r != FIB_OK) return r;
  HIPCHK(h, hipGetLastError());
  h->have_force = true;
  h->last_flags = flags;
  return FIB_OK;
}

extern "C" int32_t fib_get_forces(fib_handle *h, int32_t inst, double *F, double *u, double *utot,
                                  double virial[9]) {
  if (run_in_flight(h, "fib_get_forces")) return FIB_ERR_STATE;
  if (bad_inst(h, inst)) return fail(h, FIB_ERR_ARG, "fib_get_forces: bad instance");
  if (!h->have_force) return fail(h, FIB_ERR_STATE, "fib_get_forces: no force evaluation");
  const DInst &in = h->inst[inst];
  const size_t n = (size_t)in.natm;
  DevBufs &d = h->d;
  if (F) {
    fibk::launch_export_vec(d, inst, in, d.Fout, d.xfer, h->stream);
    HIPCHK(h, hipMemcpyAsync(F, d.xfer, 3 * n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  if (u) {
    if (!(h->last_flags & FIB_EVAL_ENERGY)) return fail(h, FIB_ERR_STATE, "energy not evaluated");
    fibk::launch_export_scalar(d, inst, in, d.uout, d.xfer, h->stream);
    HIPCHK(h, hipMemcpyAsync(u, d.xfer, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  DCtl c;
  HIPCHK(h, hipMemcpyAsync(&c, &d.ctl[inst], sizeof(DCtl), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (utot) *utot = c.utot;
  if (virial)
    for (int q = 0; q < 9; ++q) virial[q] = c.virial[q];
  return FIB_OK;
}

extern "C" int32_t fib_force(fib_handle *h, int32_t inst, int32_t impact, double *F, double *u,
                             double *utot, double virial[9]) {
  if (run_in_flight(h, "fib_force")) return FIB_ERR_STATE;
  int32_t r = fib_set_impact(h, inst, impact);
  if (r != FIB_OK) return r;
  int flags = FIB_EVAL_FORCE;
  if (u || utot) flags |= FIB_EVAL_ENERGY;
  if (virial) flags |= FIB_EVAL_VIRIAL;
  r = fib_compute(h, flags);
  if (r != FIB_OK) return r;
  return fib_get_forces(h, inst, F, u, utot, virial);
}

// ---------------------------------------------------------------------------
// reference step control (instance 0)

static void ionrun(fib_handle *h, int tic) {
  RefCtl &R = h->ref;
  const fib_run_params &p = R.rp;
  if (tic == 1) {
    R.ts = p.ts_i; R.atm_out = p.atm_out_i; R.ntlist = p.ntlist_i; R.atlist = p.atlist_i;
    R.tau = p.tau_i; R.dslist = p.dslist_i; R.regime = 1;
  } else {
    R.ts = p.ts_r; R.atm_out = p.atm_out_r; R.ntlist = p.ntlist_r; R.atlist = p.atlist_r;
    R.tau = p.tau_r; R.dslist = p.dslist_r; R.regime = 2;
  }
}

static StepArgs step_args(fib_handle *h, int mode) {
  StepArgs sa = h->base;
  const RefCtl &R = h->ref;
  sa.ts = R.ts;
  sa.tau = R.tau;
  sa.ttar = R.rp.ttar;
  sa.temp = R.temp;
  sa.sidewidth = R.rp.sidewidth;
  for (int k = 0; k < 4; ++k) sa.outz[k] = R.rp.outz[k];
  sa.atomz[0] = R.rp.atomz[0];
  sa.atomz[1] = R.rp.atomz[1];
  sa.ionz = R.rp.ionz;
  sa.mode = mode;
  sa.frz = 0;
  return sa;
}

extern "C" int32_t fib_set_run(fib_handle *h, int32_t inst, const fib_run_params *rp) {
  if (run_in_flight(h, "fib_set_run")) return FIB_ERR_STATE;
  if (bad_inst(h, inst) || inst != 0 || !rp) return fail(h, FIB_ERR_ARG, "fib_set_run: bad argument");
  if (rp->ntlist_i < 1 || rp->ntlist_r < 1 || rp->atlist_i < 1 || rp->atlist_r < 1 ||
      rp->dslist_i < 1 || rp->dslist_r < 1)
    return fail(h, FIB_ERR_ARG, "fib_set_run: cadences must be >= 1");
  h->ref.rp = *rp;
  h->ref.have_run = true;
  ionrun(h, 1);
  return FIB_OK;
}

static int32_t read_ctl(fib_handle *h, int inst, DCtl *c) {
  HIPCHK(h, hipMemcpyAsync(c, &h->d.ctl[inst], sizeof(DCtl), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FIB_OK;
}

static double temp_of(const fib_handle *h, const DCtl &c) {
  return c.glob_sum / 3. / (double)h->inst[0].nsg / h->P.kB;
}

extern "C" int32_t fib_temperature(fib_handle *h, int32_t inst, double *temp) {
  if (run_in_flight(h, "fib_temperature")) return FIB_ERR_STATE;
  if (bad_inst(h, inst) || !temp) return fail(h, FIB_ERR_ARG, "fib_temperature: bad argument");
  if (!h->have_state) return fail(h, FIB_ERR_STATE, "fib_temperature: no state");
  StepArgs sa = h->ref.have_run ? step_args(h, 1) : h->base;
  fibk::launch_sums(h->d, sa, h->stream);
  DCtl c;
  int32_t r = read_ctl(h, inst, &c);
  if (r != FIB_OK) return r;
  *temp = c.glob_sum / 3. / (double)h->inst[inst].nsg / h->P.kB;
  return FIB_OK;
}

extern "C" int32_t fib_tint_init(fib_handle *h, int64_t lt0, double time0) {
  if (run_in_flight(h, "fib_tint_init")) return FIB_ERR_STATE;
  if (!h) return FIB_ERR_ARG;
  if (!h->ref.have_run) return fail(h, FIB_ERR_STATE, "fib_tint_init: call fib_set_run first");
  if (!h->have_lists) return fail(h, FIB_ERR_STATE, "fib_tint_init: build the lists first");
  RefCtl &R = h->ref;
  R.lt = lt0;
  R.time = time0;
  double t = 0;
  int32_t r = fib_temperature(h, 0, &t);
  if (r != FIB_OK) return r;
  R.temp = t;
  // initial force into F (timeint.f90:69): force-only pass, then copy
  StepArgs sa = step_args(h, 1);
  r = run_force(h, FIB_EVAL_FORCE, 0, sa, h->d.F[h->d.cur]);
  if (r != FIB_OK) return r;
  fibk::launch_sums(h->d, sa, h->stream);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FIB_OK;
}

extern "C" int32_t fib_ti_begin(fib_handle *h, int32_t whichctrl, int32_t impact) {
  if (run_in_flight(h, "fib_ti_begin")) return FIB_ERR_STATE;
  if (!h || whichctrl < 0 || whichctrl > 2) return FIB_ERR_ARG;
  if (!h->ref.have_run) return fail(h, FIB_ERR_STATE, "fib_ti_begin: call fib_set_run first");
  RefCtl &R = h->ref;
  R.whichctrl = whichctrl;
  R.begun = true;
  if (whichctrl == 0) {
    R.tau = 1.E-14;  // timeint.f90:85
    double t = 0;
    int32_t r = fib_temperature(h, 0, &t);
    if (r != FIB_OK) return r;
    R.temp = t;
  } else if (whichctrl == 1) {
    R.fast = true;
    ionrun(h, 1);
    R.impact = impact;
    int32_t r = fib_set_impact(h, 0, impact);
    if (r != FIB_OK) return r;
  } else {
    ionrun(h, 2);
  }
  return FIB_OK;
}

extern "C" int32_t fib_knock(fib_handle *h, int32_t knc, const double xion[3],
                             const double vion[3]) {
  if (run_in_flight(h, "fib_knock")) return FIB_ERR_STATE;
  if (!h || !xion || !vion) return FIB_ERR_ARG;
  const DInst &in = h->inst[0];
  if (knc < 1 || knc > in.nlj) return fail(h, FIB_ERR_ARG, "fib_knock: ion number out of range");
  int p = -1;
  HIPCHK(h, hipMemcpy(&p, h->d.inv + in.atom_off + in.nsg + knc - 1, sizeof(int),
                      hipMemcpyDeviceToHost));
  fibk::launch_set_atom(h->d, p, xion, vion, h->base, h->stream);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FIB_OK;
}

extern "C" int32_t fib_step(fib_handle *h, const fib_step_ctl *ctl, fib_step_report *rep) {
  if (run_in_flight(h, "fib_step")) return FIB_ERR_STATE;
  if (!h || !ctl || !rep) return FIB_ERR_ARG;
  if (!h->ref.have_run || !h->ref.begun) return fail(h, FIB_ERR_STATE, "fib_step: call fib_ti_begin first");
  if (!h->have_lists) return fail(h, FIB_ERR_STATE, "fib_step: build the lists first");
  if (h->inst.size() != 1) return fail(h, FIB_ERR_STATE, "fib_step drives a single instance");
  RefCtl &R = h->ref;
  const int w = ctl->whichctrl;
  if (w != R.whichctrl) return fail(h, FIB_ERR_STATE, "fib_step: whichctrl differs from fib_ti_begin");
  std::memset(rep, 0, sizeof(*rep));
  const fib_run_params &P = R.rp;
  const int builds0 = h->n_builds;
  int steps = 0;
  for (;;) {
    if (ctl->max_steps > 0 && steps >= ctl->max_steps) break;
    if (w == 0) {
      if (!(std::fabs(R.temp - P.ttar) > P.teps)) { rep->ended = 1; break; }
      R.lt += 1;
    } else if (w == 1) {
      R.lt += 1;
      if (R.time > (double)R.impact * P.dti) { rep->ended = 1; break; }
    } else {
      R.lt += 1;
    }
    const int64_t lt = R.lt;
    // onestep (timeint.f90:158-242)
    R.time = R.time + R.ts;
    StepArgs sa = step_args(h, w == 0 ? 0 : 1);
    sa.frz = (lt % R.atlist == 0) ? 1 : 0;
    fibk::launch_pos(h->d, sa, h->stream);
    if (w == 0) {
      DCtl c;
      int32_t r = read_ctl(h, 0, &c);
      if (r != FIB_OK) return r;
      R.temp = c.eq_sum / 3. / (double)h->inst[0].nsg / h->P.kB;
    }
    if (lt % R.ntlist == 0) {
      int32_t r = build_lists(h);
      if (r != FIB_OK) return r;
    }
    if (lt % R.dslist == 0 && w == 1) {
      DCtl c;
      int32_t r = read_ctl(h, 0, &c);
      if (r != FIB_OK) return r;
      if (c.kemax < P.ke_limit_1 && R.fast) {
        ionrun(h, 2);
        R.fast = false;
      }
    }
    StepArgs sb = step_args(h, 1);
    int32_t r = run_force(h, FIB_EVAL_FORCE, 1, sb);
    if (r != FIB_OK) return r;
    ++steps;
    // iostuff (out.f90:129-147), called by onestep for whichctrl > 0 only:
    // on either cadence the temperature is recomputed
    int out = 0;
    if (w > 0) {
      const int atoms = (R.atm_out > 0 && lt % R.atm_out == 0);
      out = (P.res_out > 0 && lt % P.res_out == 0) || atoms;
      if (out) {
        DCtl c;
        r = read_ctl(h, 0, &c);
        if (r != FIB_OK) return r;
        R.temp = temp_of(h, c);
        rep->atoms_due = atoms;
      }
    }
    if (out) {
      rep->output_due = 1;
      if (ctl->stop_on_output) break;
    }
  }
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->tim.flush();
  rep->lt = R.lt;
  rep->time = R.time;
  rep->temp = R.temp;
  rep->ts = R.ts;
  rep->steps_done = steps;
  rep->regime = R.regime;
  rep->nlist_builds = h->n_builds - builds0;
  h->have_force = false;
  return FIB_OK;
}

// ---------------------------------------------------------------------------
// MD-step mode

static void enqueue_md_step(fib_handle *h) {
  DevBufs &d = h->d;
  const StepArgs &sa = h->md_sa;
  fibk::launch_md_pos(d, h->fc, sa, h->md_gate, h->stream);
  // list rebuild, skipped on the device unless the gate is set; the sort is
  // in place so every array stays where the step (or its graph) expects it
  d.gate = h->md_gate;
  fibk::launch_bin(d, h->stream);
  fibk::launch_scan(d, h->stream);
  fibk::launch_sort(d, h->stream, true);
  fibk::launch_nbuild(d, h->fc, h->stream);
  fibk::launch_mark_build(d, h->stream);
  d.gate = nullptr;
  const bool mixed = h->prec == FIB_PREC_MIXED;
  // no stale-list pass: a step either rebuilt an instance's lists or left
  // its two largest displacements within the skin, so every pair in range
  // is listed and k_extra would find nothing
  fibk::launch_force_a(d, h->fc, mixed, false, false, h->stream, false);
  fibk::launch_force_b(d, h->fc, mixed, false, 2, sa, d.F[d.cur], h->stream);
  fibk::launch_md_ctl(d, sa, h->stream);
}

// forces of the current state into F, and dt from the current |v|max
static int32_t md_refresh(fib_handle *h, bool reset) {
  DevBufs &d = h->d;
  int32_t r = run_force(h, FIB_EVAL_FORCE, 0, h->md_sa, d.F[d.cur]);
  if (r != FIB_OK) return r;
  fibk::launch_sums(d, h->md_sa, h->stream);
  fibk::launch_md_init(d, h->md_sa, reset, h->stream);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return FIB_OK;
}

// Layout of md_gate: [0] any gate, [1 + k] gate of instance k, [1 + n + k]
// cell prefix of the gated instances (n + 1 entries), then the removal list
// [count, FIB_LEAVE_CAP atoms], then the halt flag.
static size_t md_leave_idx(const fib_handle *h) { return 2 * h->inst.size() + 2; }
static size_t md_halt_idx(const fib_handle *h) { return md_leave_idx(h) + 1 + FIB_LEAVE_CAP; }

static int32_t md_halted(fib_handle *h, bool &halted) {
  int v = 0;
  HIPCHK(h, hipMemcpy(&v, h->d.halt, sizeof(int), hipMemcpyDeviceToHost));
  halted = v != 0;
  return FIB_OK;
}

// Gives up on a list that cannot grow: the halt flag is cleared so that the
// handle's other calls work; the state is that of the halted step.
static int32_t md_abort(fib_handle *h, const std::string &msg) {
  fibk::launch_clear_overflow(h->d, h->stream, true);
  (void)hipStreamSynchronize(h->stream);
  return fail(h, FIB_ERR_OVERFLOW, msg);
}

// A gated build overflowed its list width (the halt flag is set, the stream
// idle).  Grow the widths to fit the longest list, copy the lists of the
// instances not rebuilt into the wider layout, repeat the gated list build
// (binning and sort are done; positions are where the build found them) and,
// when the build belonged to an MD step, finish that step.
static int32_t md_recover(fib_handle *h, bool finish_step) {
  DevBufs &d = h->d;
  for (int attempt = 0; attempt < 4; ++attempt) {
    std::vector<DCtl> c(h->inst.size());
    HIPCHK(h, hipMemcpy(c.data(), d.ctl, sizeof(DCtl) * c.size(), hipMemcpyDeviceToHost));
    int msw = 0, mion = 0;
    for (auto &x : c) {
      msw = std::max(msw, (int)x.maxcnt_sw);
      mion = std::max(mion, (int)x.maxcnt_ion);
    }
    int nsw = d.Wsw, nion = d.Wion;
    while (nsw < msw) nsw *= 2;
    while (nion < mion) nion *= 2;
    if (nsw > 64) return md_abort(h, "SW neighbor list exceeds 64 entries per atom");
    if (nsw != d.Wsw || nion != d.Wion) {
      int32_t r = regrow_lists(h, nsw, nion);
      if (r != FIB_OK) return r;
    }
    fibk::launch_clear_overflow(d, h->stream);
    const int one = 1;
    HIPCHK(h, hipMemcpyAsync(h->md_gate, &one, sizeof(int), hipMemcpyHostToDevice, h->stream));
    d.gate = h->md_gate;
    fibk::launch_nbuild(d, h->fc, h->stream);
    fibk::launch_mark_build(d, h->stream, false);
    d.gate = nullptr;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    bool halted = false;
    int32_t r = md_halted(h, halted);
    if (r != FIB_OK) return r;
    if (halted) continue;
    if (finish_step) {
      const bool mixed = h->prec == FIB_PREC_MIXED;
      fibk::launch_force_a(d, h->fc, mixed, false, false, h->stream, false);
      fibk::launch_force_b(d, h->fc, mixed, false, 2, h->md_sa, d.F[d.cur], h->stream);
      fibk::launch_md_ctl(d, h->md_sa, h->stream);
      HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    h->graph_valid = false;
    return FIB_OK;
  }
  return md_abort(h, "neighbor list did not fit after regrowth");
}

extern "C" int32_t fib_md_begin(fib_handle *h, const fib_md_params *p, const double *surface_z,
                                const double *bottom_z) {
  if (!h || !p || !surface_z || !bottom_z) return FIB_ERR_ARG;
  if (run_in_flight(h, "fib_md_begin")) return FIB_ERR_STATE;
  if (!(p->dt_max > 0) || !(p->dx_max > 0) || !(p->remove_height > 0) || p->shell_width < 0 ||
      (p->shell_width > 0 && !(p->tau > 0 && p->ttar > 0)) || p->displaced_r < 0)
    return fail(h, FIB_ERR_ARG, "fib_md_begin: bad parameters");
  if (!h->have_state) return fail(h, FIB_ERR_STATE, "fib_md_begin: set the state first");
  if (!h->md_gate) {
    const size_t ng = md_halt_idx(h) + 1;
    HIPCHK(h, hipMalloc(&h->md_gate, sizeof(int) * ng));
    HIPCHK(h, hipMemset(h->md_gate, 0, sizeof(int) * ng));
  }
  for (size_t k = 0; k < h->inst.size(); ++k) {
    h->inst[k].surf_z = surface_z[k];
    h->inst[k].bot_z = bottom_z[k];
  }
  HIPCHK(h, hipMemcpy(h->d.I, h->inst.data(), sizeof(DInst) * h->inst.size(),
                      hipMemcpyHostToDevice));
  h->md = *p;
  StepArgs &sa = h->md_sa;
  sa = h->base;
  sa.dt_max = p->dt_max;
  sa.dx_max = p->dx_max;
  sa.remove_dz = p->remove_height;
  sa.shell_w = p->shell_width;
  sa.ttar = p->ttar;
  sa.tau = p->tau > 0 ? p->tau : 1.0;
  sa.disp_r = p->displaced_r;
  sa.skin = h->fc.skin;
  // a fresh start: no removal pending, nothing halted
  HIPCHK(h, hipMemset(h->md_gate + md_leave_idx(h), 0, sizeof(int)));
  HIPCHK(h, hipMemset(h->md_gate + md_halt_idx(h), 0, sizeof(int)));
  h->d.halt = h->md_gate + md_halt_idx(h);
  if (!h->have_lists) {
    int32_t r = build_lists(h);
    if (r != FIB_OK) return r;
  }
  h->md_on = true;
  h->graph_valid = false;
  return md_refresh(h, true);
}

extern "C" int32_t fib_md_fire_many(fib_handle *h, int32_t count, const int32_t *inst,
                                    const int32_t *ion, const double *x, const double *v) {
  if (!h || count < 0 || (count > 0 && (!inst || !ion || !x || !v))) return FIB_ERR_ARG;
  if (!h->md_on) return fail(h, FIB_ERR_STATE, "fib_md_fire: call fib_md_begin first");
  if (run_in_flight(h, "fib_md_fire")) return FIB_ERR_STATE;
  if (count == 0) return FIB_OK;
  const int n = (int)h->inst.size();
  std::vector<int> fired(n, 0);
  for (int q = 0; q < count; ++q) {
    if (bad_inst(h, inst[q])) return fail(h, FIB_ERR_ARG, "fib_md_fire: bad instance");
    if (ion[q] < 1 || ion[q] > h->inst[inst[q]].nlj)
      return fail(h, FIB_ERR_ARG, "fib_md_fire: ion out of range");
    if (fired[inst[q]]++) return fail(h, FIB_ERR_ARG, "fib_md_fire: one ion per instance and call");
  }
  DevBufs &d = h->d;
  // sorted positions of the ions (the device's inverse map), then place them
  std::vector<int> pos(count);
  for (int q = 0; q < count; ++q) {
    const DInst &in = h->inst[inst[q]];
    HIPCHK(h, hipMemcpy(&pos[q], d.inv + in.atom_off + in.nsg + ion[q] - 1, sizeof(int),
                        hipMemcpyDeviceToHost));
  }
  for (int q = 0; q < count; ++q) {
    fibk::launch_set_atom(d, pos[q], x + 3 * (size_t)q, v + 3 * (size_t)q, h->md_sa, h->stream);
    DInst &in = h->inst[inst[q]];
    in.impact = ion[q];
    HIPCHK(h, hipMemcpyAsync(&d.I[inst[q]], &in, sizeof(DInst), hipMemcpyHostToDevice,
                             h->stream));
  }
  // the ions jumped: rebuild the fired instances only (gated, in place, so
  // the other instances keep their order, lists and rebuild bookkeeping)
  std::vector<int> g(2 + 2 * (size_t)n, 0);
  int acc = 0;
  g[0] = 1;
  for (int k = 0; k < n; ++k) {
    g[1 + k] = fired[k];
    g[1 + n + k] = acc;
    if (fired[k]) acc += h->inst[k].nc[0] * h->inst[k].nc[1] * h->inst[k].nc[2];
  }
  g[1 + 2 * (size_t)n] = acc;
  HIPCHK(h, hipMemcpyAsync(h->md_gate, g.data(), sizeof(int) * g.size(), hipMemcpyHostToDevice,
                           h->stream));
  d.gate = h->md_gate;
  fibk::launch_bin(d, h->stream);
  fibk::launch_scan(d, h->stream);
  fibk::launch_sort(d, h->stream, true);
  fibk::launch_nbuild(d, h->fc, h->stream);
  fibk::launch_mark_build(d, h->stream);
  d.gate = nullptr;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  bool halted = false;
  int32_t r = md_halted(h, halted);
  if (r != FIB_OK) return r;
  if (halted) {
    r = md_recover(h, false);
    if (r != FIB_OK) return r;
  }
  ++h->n_builds;
  h->have_lists = true;
  // the ledger takes each ion's potential energy at its start point with its
  // kinetic energy
  for (int q = 0; q < count; ++q) {
    const DInst &in = h->inst[inst[q]];
    fibk::launch_fire_energy(d, h->fc, in.atom_off + in.nsg + ion[q] - 1, h->stream);
  }
  return md_refresh(h, false);
}

extern "C" int32_t fib_md_fire(fib_handle *h, int32_t inst, int32_t ion, const double x[3],
                               const double v[3]) {
  if (!h || !x || !v) return FIB_ERR_ARG;
  return fib_md_fire_many(h, 1, &inst, &ion, x, v);
}

static int32_t md_steps_done(fib_handle *h, double &steps) {
  DCtl c;
  HIPCHK(h, hipMemcpy(&c, h->d.ctl, sizeof(DCtl), hipMemcpyDeviceToHost));
  steps = c.cfg_steps;
  return FIB_OK;
}

// Enqueue `left` MD steps on the handle's stream (the captured graph, made
// when missing, or ordinary launches), with the step timer around each.
static int32_t md_enqueue_steps(fib_handle *h, int left) {
  if (h->md.use_graph) {
    if (!h->graph_valid) {
      if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
      if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
      const KTimer *kt = h->d.kt;
      h->d.kt = nullptr;  // no timing events inside the graph
      HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      enqueue_md_step(h);
      hipError_t e = hipStreamEndCapture(h->stream, &h->graph);
      h->d.kt = kt;
      HIPCHK(h, e);
      HIPCHK(h, hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
      h->graph_valid = true;
    }
    for (int k = 0; k < left; ++k) {
      const int t = h->tim.begin(KT_STEP);
      HIPCHK(h, hipGraphLaunch(h->graph_exec, h->stream));
      h->tim.end(t, KT_STEP);
    }
  } else {
    for (int k = 0; k < left; ++k) {
      const int t = h->tim.begin(KT_STEP);
      enqueue_md_step(h);
      h->tim.end(t, KT_STEP);
    }
  }
  HIPCHK(h, hipGetLastError());
  return FIB_OK;
}

// Wait for the enqueued steps; while a gated build of some step overflowed
// (the steps after it did nothing), grow the lists, finish that step and run
// the rest; then the per-instance error flags.
static int32_t md_finish(fib_handle *h) {
  for (int rounds = 0;; ++rounds) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    bool halted = false;
    int32_t r = md_halted(h, halted);
    if (r != FIB_OK) return r;
    if (!halted) break;
    r = md_recover(h, true);
    if (r != FIB_OK) return r;
    if (rounds > 64) return md_abort(h, "fib_md_run: lists keep overflowing");
    double now = 0;
    r = md_steps_done(h, now);
    if (r != FIB_OK) return r;
    const int left = (int)(h->run_target - now + 0.5);
    if (left <= 0) break;
    r = md_enqueue_steps(h, left);
    if (r != FIB_OK) return r;
  }
  h->have_force = false;
  std::vector<DCtl> c(h->inst.size());
  HIPCHK(h, hipMemcpy(c.data(), h->d.ctl, sizeof(DCtl) * c.size(), hipMemcpyDeviceToHost));
  for (auto &x : c) {
    if (x.overflow) return fail(h, FIB_ERR_OVERFLOW, "fib_md_run: a neighbour list outgrew its width");
    if (x.nonfinite) return fail(h, FIB_ERR_NONFINITE, "fib_md_run: non-finite position");
  }
  for (auto &x : c)
    if (x.leave_overflow) {
      // the ledger missed the interaction energy of the atoms past the
      // removal list's capacity in that step; no atom stays marked leaving
      fibk::launch_clear_leaving(h->d, h->stream);
      fibk::launch_clear_overflow(h->d, h->stream, false);
      HIPCHK(h, hipStreamSynchronize(h->stream));
      return fail(h, FIB_ERR_OVERFLOW, "fib_md_run: more atoms left the simulation in one step "
                                       "than the removal list holds (" +
                                       std::to_string(FIB_LEAVE_CAP) + "); the energy ledger "
                                       "misses their interaction energy");
    }
  return FIB_OK;
}

extern "C" int32_t fib_md_run_begin(fib_handle *h, int32_t nsteps) {
  if (!h || nsteps < 0) return FIB_ERR_ARG;
  if (!h->md_on) return fail(h, FIB_ERR_STATE, "fib_md_run: call fib_md_begin first");
  if (run_in_flight(h, "fib_md_run")) return FIB_ERR_STATE;
  if (nsteps == 0) return FIB_OK;
  // every instance steps together: instance 0's count is the batch's
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double s0 = 0;
  int32_t r = md_steps_done(h, s0);
  if (r != FIB_OK) return r;
  h->run_target = s0 + nsteps;
  r = md_enqueue_steps(h, nsteps);
  if (r != FIB_OK) {
    (void)hipStreamSynchronize(h->stream);
    return r;
  }
  h->run_pending = true;
  return FIB_OK;
}

extern "C" int32_t fib_md_run_end(fib_handle *h) {
  if (!h) return FIB_ERR_ARG;
  if (!h->run_pending) return FIB_OK;
  h->run_pending = false;
  return md_finish(h);
}

extern "C" int32_t fib_md_run(fib_handle *h, int32_t nsteps) {
  int32_t r = fib_md_run_begin(h, nsteps);
  if (r != FIB_OK) return r;
  return fib_md_run_end(h);
}

extern "C" int32_t fib_snapshot_begin(fib_handle *h, int32_t inst, int32_t slot) {
  if (bad_inst(h, inst) || slot < 0 || slot >= FIB_SNAPSHOT_SLOTS)
    return fail(h, FIB_ERR_ARG, "fib_snapshot_begin: bad instance or slot");
  if (!h->have_state) return fail(h, FIB_ERR_STATE, "fib_snapshot_begin: no state");
  if (run_in_flight(h, "fib_snapshot_begin")) return FIB_ERR_STATE;
  auto &sn = h->snap[slot];
  if (!h->copy_stream) HIPCHK(h, hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
  if (!sn.ready) {
    HIPCHK(h, hipEventCreateWithFlags(&sn.ready, hipEventDisableTiming));
    HIPCHK(h, hipEventCreateWithFlags(&sn.copied, hipEventDisableTiming));
  }
  // the slot's previous copy must have left the staging buffer
  if (sn.pending) HIPCHK(h, hipEventSynchronize(sn.copied));
  sn.pending = false;
  const DInst &in = h->inst[inst];
  const size_t n = (size_t)in.natm, need = 6 * n;
  if (need > sn.cap) {
    if (sn.dev) { (void)hipFree(sn.dev); sn.dev = nullptr; }
    if (sn.host) { (void)hipHostFree(sn.host); sn.host = nullptr; }
    sn.cap = 0;
    HIPCHK(h, hipMalloc((void **)&sn.dev, need * sizeof(double)));
    HIPCHK(h, hipHostMalloc((void **)&sn.host, need * sizeof(double), hipHostMallocDefault));
    sn.cap = need;
  }
  DevBufs &d = h->d;
  fibk::launch_export_x(d, inst, in, sn.dev, h->stream);
  fibk::launch_export_vec(d, inst, in, d.V[d.cur], sn.dev + 3 * n, h->stream);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventRecord(sn.ready, h->stream));
  HIPCHK(h, hipStreamWaitEvent(h->copy_stream, sn.ready, 0));
  HIPCHK(h, hipMemcpyAsync(sn.host, sn.dev, need * sizeof(double), hipMemcpyDeviceToHost,
                           h->copy_stream));
  HIPCHK(h, hipEventRecord(sn.copied, h->copy_stream));
  sn.n = n;
  sn.pending = true;
  return FIB_OK;
}

extern "C" int32_t fib_snapshot_wait(fib_handle *h, int32_t slot, double *X, double *V) {
  if (!h || slot < 0 || slot >= FIB_SNAPSHOT_SLOTS)
    return fail(h, FIB_ERR_ARG, "fib_snapshot_wait: bad slot");
  auto &sn = h->snap[slot];
  if (!sn.pending) return fail(h, FIB_ERR_STATE, "fib_snapshot_wait: no snapshot in this slot");
  HIPCHK(h, hipEventSynchronize(sn.copied));
  sn.pending = false;
  if (X) std::memcpy(X, sn.host, 3 * sn.n * sizeof(double));
  if (V) std::memcpy(V, sn.host + 3 * sn.n, 3 * sn.n * sizeof(double));
  return FIB_OK;
}

extern "C" int32_t fib_md_use_graph(fib_handle *h, int32_t on) {
  if (!h) return FIB_ERR_ARG;
  if (!h->md_on) return fail(h, FIB_ERR_STATE, "fib_md_use_graph: call fib_md_begin first");
  if (run_in_flight(h, "fib_md_use_graph")) return FIB_ERR_STATE;
  h->md.use_graph = on ? 1 : 0;
  return FIB_OK;
}

extern "C" int32_t fib_set_list_width(fib_handle *h, int32_t wsw, int32_t wion) {
  if (run_in_flight(h, "fib_set_list_width")) return FIB_ERR_STATE;
  if (!h) return FIB_ERR_ARG;
  if (wsw < 4 || wsw > 64 || wsw % 4 || wion < 4 || wion % 4)
    return fail(h, FIB_ERR_ARG, "fib_set_list_width: widths must be multiples of 4, SW at most 64");
  if (h->md_on) return fail(h, FIB_ERR_STATE, "fib_set_list_width: not in MD-step mode");
  int32_t r = alloc_lists(h, wsw, wion);
  if (r != FIB_OK) return r;
  h->have_lists = false;
  h->have_force = false;
  h->graph_valid = false;
  return FIB_OK;
}

extern "C" int32_t fib_set_tiles(fib_handle *h, int32_t tri_ns, int32_t tri_batch,
                                 int32_t nb_cells) {
  if (!h) return FIB_ERR_ARG;
  if (run_in_flight(h, "fib_set_tiles")) return FIB_ERR_STATE;
  if ((tri_ns != 0 && tri_ns != 128 && tri_ns != 256) ||
      (tri_batch != 0 && tri_batch != 2 && tri_batch != 4) ||
      (nb_cells != 0 && nb_cells != 4 && nb_cells != 8 && nb_cells != 16))
    return fail(h, FIB_ERR_ARG, "fib_set_tiles: tri_ns 128|256, tri_batch 2|4, nb_cells 4|8|16 "
                                "(0: default)");
  h->d.tri_ns = tri_ns;
  h->d.tri_tb = tri_batch;
  h->d.nb_cells = nb_cells;
  h->graph_valid = false;  // the captured step holds the old launches
  return FIB_OK;
}

extern "C" int32_t fib_get_tiles(const fib_handle *h, int32_t tiles[3]) {
  if (!h || !tiles) return FIB_ERR_ARG;
  tiles[0] = h->d.tri_ns ? h->d.tri_ns : 256;
  tiles[1] = h->d.tri_tb ? h->d.tri_tb : 4;
  tiles[2] = h->d.nb_cells ? h->d.nb_cells : 8;
  return FIB_OK;
}

extern "C" int32_t fib_get_list_width(const fib_handle *h, int32_t *wsw, int32_t *wion) {
  if (!h || !wsw || !wion) return FIB_ERR_ARG;
  *wsw = h->d.Wsw;
  *wion = h->d.Wion;
  return FIB_OK;
}

extern "C" int32_t fib_md_get_report(fib_handle *h, int32_t inst, fib_md_report *r) {
  if (!h || !r) return FIB_ERR_ARG;
  if (!h->md_on) return fail(h, FIB_ERR_STATE, "fib_md_get_report: call fib_md_begin first");
  if (run_in_flight(h, "fib_md_get_report")) return FIB_ERR_STATE;
  if (bad_inst(h, inst)) return fail(h, FIB_ERR_ARG, "fib_md_get_report: bad instance");
  fibk::launch_md_report(h->d, h->md_sa, h->stream);
  DCtl c;
  HIPCHK(h, hipMemcpyAsync(&c, &h->d.ctl[inst], sizeof(DCtl), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::memset(r, 0, sizeof(*r));
  r->time = c.cfg_time;
  r->dt = c.cfg_dt;
  r->ion_z = c.ion_z;
  r->ke = c.ke_sum;
  r->steps = (int64_t)c.cfg_steps;
  r->removed = (int32_t)c.n_removed;
  r->ions_removed = (int32_t)c.n_ga_removed;
  r->displaced = (int32_t)c.n_displaced;
  r->builds = (int32_t)c.builds;
  r->e_bath = c.e_bath;
  r->e_removed = c.e_out;
  r->e_fired = c.e_in;
  r->dt_min = c.cfg_dt_min < 1e299 ? c.cfg_dt_min : 0.0;
  return FIB_OK;
}

extern "C" int32_t fib_md_energy(fib_handle *h, double *epot) {
  if (!h || !epot) return FIB_ERR_ARG;
  if (!h->md_on) return fail(h, FIB_ERR_STATE, "fib_md_energy: call fib_md_begin first");
  if (run_in_flight(h, "fib_md_energy")) return FIB_ERR_STATE;
  // an evaluation with energies into the output buffer: the step's own
  // forces (F of the current state) stay as they are
  int32_t r = run_force(h, FIB_EVAL_FORCE | FIB_EVAL_ENERGY, 0, h->md_sa, h->d.Fout);
  if (r != FIB_OK) return r;
  HIPCHK(h, hipGetLastError());
  std::vector<DCtl> c(h->inst.size());
  HIPCHK(h, hipMemcpyAsync(c.data(), h->d.ctl, sizeof(DCtl) * c.size(), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t k = 0; k < c.size(); ++k) epot[k] = c[k].utot;
  h->have_force = true;
  h->last_flags = FIB_EVAL_FORCE | FIB_EVAL_ENERGY;
  return FIB_OK;
}

extern "C" int32_t fib_timing_enable(fib_handle *h, int32_t on) {
  if (run_in_flight(h, "fib_timing_enable")) return FIB_ERR_STATE;
  if (!h) return FIB_ERR_ARG;
  h->tim.reset();
  h->tim.on = on != 0;
  h->tim.mask = (uint32_t)on;
  return FIB_OK;
}

extern "C" int32_t fib_get_timing(fib_handle *h, fib_timing *t) {
  if (run_in_flight(h, "fib_get_timing")) return FIB_ERR_STATE;
  if (!h || !t) return FIB_ERR_ARG;
  h->tim.flush();
  std::memset(t, 0, sizeof(*t));
  for (int k = 0; k < FIB_NTIMERS; ++k) {
    t->ms[k] = h->tim.acc[k];
    t->calls[k] = h->tim.cnt[k];
  }
  return FIB_OK;
}

extern "C" const char *fib_timer_name(int32_t k) {
  return (k >= 0 && k < KT_N) ? kTimerNames[k] : nullptr;
}

extern "C" int32_t fib_synchronize(fib_handle *h) {
  if (run_in_flight(h, "fib_synchronize")) return FIB_ERR_STATE;
  if (!h) return FIB_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipGetLastError());
  return FIB_OK;
}


// ---------------------------------------------------------------------------
// Handle-free force evaluation inside a caller-provided device arena.  The
// arrays of one handle (one instance) are carved out of `arena` and the
// usual pipeline runs: import, binning and sort, list build, force passes,
// export in the caller's atom order.  No device memory is allocated.  Used
// by fib_force_noalloc (fib_noalloc.hip), which brings its own arena.

namespace {
// Host copy of the pair table of the last parameter block, kept for the
// asynchronous upload (its memory must outlive the call).
struct TableCache {
  fib_params P{};
  bool have = false;
  std::vector<TabEntry> tab;
  FConst fc{};
};
thread_local TableCache g_tab;

struct Arena {
  char *base;
  size_t size, used = 0;
  template <typename T>
  T *take(size_t n) {
    const size_t b = (used + 255) & ~(size_t)255;
    const size_t e = b + std::max<size_t>(n, 1) * sizeof(T);
    if (e > size) return nullptr;
    used = e;
    return reinterpret_cast<T *>(base + b);
  }
};
}  // namespace

int32_t fib_arena_force(const fib_params *p, int32_t precision, int32_t natm, int32_t nsg,
                        int32_t nlj, const double lb[3], int32_t impact, const double *X,
                        const int32_t *atype, double *F, double *utot, void *arena,
                        size_t arena_bytes, int32_t max_atoms, hipStream_t s) {
  auto bad = [](int32_t code, const std::string &m) { g_free_err = m; return code; };
  if (!p || !lb || !X || !atype || !F || !arena)
    return bad(FIB_ERR_ARG, "fib_force_noalloc: null argument");
  if (precision != FIB_PREC_FP64 && precision != FIB_PREC_MIXED)
    return bad(FIB_ERR_ARG, "fib_force_noalloc: bad precision");
  if (natm > max_atoms)
    return bad(FIB_ERR_ARG, "fib_force_noalloc: " + std::to_string(natm) + " atoms, at most " +
                                std::to_string(max_atoms) + " without allocation");
  if (impact < 0 || impact > nlj) return bad(FIB_ERR_ARG, "fib_force_noalloc: bad impact");
  // constants and the pair table: derived on the host (no device memory)
  if (!g_tab.have || std::memcmp(&g_tab.P, p, sizeof(fib_params)) != 0) {
    fib_handle tmp;
    tmp.P = *p;
    derive_constants(&tmp);
    g_tab.P = *p;
    g_tab.tab = tmp.tab;
    g_tab.fc = tmp.fc;
    g_tab.have = true;
  }
  const double smax = std::max(p->sigma_ge, p->sigma_si);
  const double rc = p->listfac * p->sw_a * smax;
  fib_instance_desc ds{};
  ds.natm = natm;
  ds.nsg = nsg;
  ds.nlj = nlj;
  for (int a = 0; a < 3; ++a) ds.lb[a] = lb[a];
  DInst in;
  long atom_off = 0, bucket_off = 0, cell_off = 0, natm_before = 0;
  if (!layout_instance(rc, ds, atom_off, bucket_off, cell_off, natm_before, in))
    return bad(FIB_ERR_ARG, "fib_force_noalloc: bad atom counts or box");
  in.impact = impact;
  DevBufs d{};
  d.NA = (int)atom_off;
  d.ninst = 1;
  d.nbucket = (int)bucket_off;
  d.total_cells = (int)cell_off;
  d.Wsw = 32;
  d.Wion = 64;
  d.pstride = d.NA / FIB_WAVE + 1;
  d.max_pad = in.natm_pad;
  const size_t NA = (size_t)d.NA;
  const bool mixed = precision == FIB_PREC_MIXED;
  const size_t g4 = mixed ? sizeof(float4) : sizeof(double4);
  Arena A{static_cast<char *>(arena), arena_bytes};
  d.I = A.take<DInst>(1);
  d.ctl = A.take<DCtl>(1);
  for (int c = 0; c < 2; ++c) {
    d.X[c] = A.take<double4>(NA);
    d.V[c] = A.take<double>(3 * NA);
    d.F[c] = A.take<double>(3 * NA);
    d.Xi[c] = A.take<double>(3 * NA);
  }
  d.inv = A.take<int>(NA);
  d.ibase = A.take<int>(1);
  d.binst = A.take<int>(NA / FIB_BLOCK);
  d.bkey = A.take<int>(NA);
  d.brank = A.take<int>(NA);
  d.bcnt = A.take<int>((size_t)d.nbucket);
  d.cstart = A.take<int>((size_t)d.nbucket + 1);
  d.perm = A.take<int>(NA);
  d.bsum = A.take<int>((size_t)d.nbucket / 2048 + 2);
  d.cnt_sw = A.take<int>(NA);
  d.cnt_ion = A.take<int>(NA);
  d.amask = A.take<uint64_t>(NA);
  d.cnt_leg = A.take<int>(NA);
  d.off_sw = A.take<uint8_t>(32 * NA);
  d.Xb = A.take<double4>(NA);
  d.disp = A.take<float>(NA);
  d.exmask = A.take<uint64_t>(NA);
  d.EXself = A.take<char>(g4 * NA);
  d.Fpart = A.take<double>(3 * NA);
  d.upart = A.take<double>(NA);
  d.Fout = A.take<double>(3 * NA);
  d.uout = A.take<double>(NA);
  d.partials = A.take<double>((size_t)d.pstride * FIB_NRED * 3);
  d.fin_tmp = A.take<double>((d.max_pad / FIB_WAVE / FIB_FIN_T + 1) * FIB_NRED);
  d.tab = A.take<TabEntry>(g_tab.tab.size());
  d.nbr_sw = A.take<int>((size_t)d.Wsw * NA);
  d.nbr_ion = A.take<int>((size_t)d.Wion * NA);
  d.rk_sw = A.take<uint16_t>((size_t)d.Wsw * NA);
  d.G = A.take<char>(g4 * d.Wsw * NA);
  d.RA = A.take<uint32_t>((size_t)d.Wsw * NA);
  d.EX = A.take<char>(g4 * d.Wsw * NA);
  if (!d.EX)
    return bad(FIB_ERR_ARG, "fib_force_noalloc: the configuration needs more than the " +
                                std::to_string(arena_bytes >> 20) + " MiB arena (" +
                                std::to_string(d.nbucket) + " cell buckets)");
  // start state: instance record and zeroed bookkeeping, written by kernels
  // (no memset or host copy of the arrays)
  fibk::launch_put_inst(d, in, s);
  fibk::launch_fill_zero(d.binst, NA / FIB_BLOCK * sizeof(int), s);
  fibk::launch_fill_zero(d.ctl, sizeof(DCtl), s);
  fibk::launch_fill_zero(d.cnt_sw, NA * sizeof(int), s);
  fibk::launch_fill_zero(d.cnt_ion, NA * sizeof(int), s);
  fibk::launch_fill_zero(d.amask, NA * sizeof(uint64_t), s);
  fibk::launch_fill_zero(d.cnt_leg, NA * sizeof(int), s);
  fibk::launch_fill_zero(d.exmask, NA * sizeof(uint64_t), s);
  fibk::launch_fill_zero(d.off_sw, 32 * NA, s);
  fibk::launch_fill_zero(d.disp, NA * sizeof(float), s);
  if (!mixed) {
    hipError_t e = hipMemcpyAsync(d.tab, g_tab.tab.data(), sizeof(TabEntry) * g_tab.tab.size(),
                                  hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return bad(FIB_ERR_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
  }
  fibk::launch_import(d, 0, in, X, nullptr, atype, nullptr, nullptr, s);
  fibk::launch_bin(d, s);
  fibk::launch_scan(d, s);
  fibk::launch_sort(d, s);
  d.cur = 1 - d.cur;
  fibk::launch_nbuild(d, g_tab.fc, s);
  // the widths cannot grow here: a longer list is an error
  DCtl c;
  hipError_t e = hipMemcpyAsync(&c, d.ctl, sizeof(DCtl), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return bad(FIB_ERR_HIP, std::string("fib_force_noalloc: ") + hipGetErrorString(e));
  if (c.nonfinite) return bad(FIB_ERR_NONFINITE, "fib_force_noalloc: non-finite position");
  if (c.maxcnt_sw > d.Wsw || c.maxcnt_ion > d.Wion)
    return bad(FIB_ERR_OVERFLOW, "fib_force_noalloc: a neighbour list is longer than its "
                                 "fixed width (SW 32, ion 64)");
  fibk::launch_force_a(d, g_tab.fc, mixed, utot != nullptr, false, s);
  fibk::launch_force_b(d, g_tab.fc, mixed, false, 0, StepArgs{}, d.Fout, s);
  fibk::launch_export_vec(d, 0, in, d.Fout, F, s);
  if (utot) fibk::launch_copy_utot(d, utot, s);
  e = hipGetLastError();
  if (e != hipSuccess) return bad(FIB_ERR_HIP, std::string("fib_force_noalloc: ") + hipGetErrorString(e));
  return FIB_OK;
}
