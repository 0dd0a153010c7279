// fib_kernels.hip -- gfx950 kernels of the FIB_MD force-and-integrate path.
//
// Pipeline per neighbor rebuild (replaces si_nlist, si.f90:87-276):
//   k_bin       cell of every atom with the reference's FP64 arithmetic
//               (nlist.f90:29), atomic histogram per cell bucket
//   k_scan_*    exclusive scan of the bucket histogram
//   k_scatter   counting-sort scatter; k_fixup restores a deterministic
//               order inside each bucket (old sorted position)
//   k_reorder   gathers every per-atom array into cell order
//   k_nbuild    one wavefront per cell: the 27-cell halo is staged in LDS,
//               the reference pair predicate (si.f90:156-170) is evaluated
//               lane-parallel and survivors are compacted with ballots
//   k_rev       slot of i inside each neighbour's list (for the force gather)
// Per force evaluation (replaces si_force, si.f90:384-552):
//   k_force_a   owner-computes: each lane owns one atom, sums its SW pair and
//               Moliere forces and every three-body term centred on it in
//               registers; the share of a centre term acting on neighbour a
//               is written to G[slot(a)] -- no atomics anywhere
//   k_force_b   gathers the G entries that neighbours wrote for this atom and
//               (in MD mode) applies the velocity-Verlet velocity update
// Per step (replaces the inline Verlet of onestep, timeint.f90:165-236):
//   k_pos       thermostat scaling, position update, side freeze, rebox and
//               the sputter-layer rules in one pass
//
// Reductions (energy, virial, thermostat sums, max KE) use per-block
// partials; the last block of an instance sums them in block order, so every
// result is deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fib_device.h"
#include "fib_kernels.h"

namespace fibk {

// ---------------------------------------------------------------------------
// helpers
__device__ __forceinline__ uint64_t wbits(const double4 &x) {
  return (uint64_t)__double_as_longlong(x.w);
}
__device__ __forceinline__ uint32_t meta_of(const double4 &x) { return (uint32_t)wbits(x); }
__device__ __forceinline__ int32_t orig_of(const double4 &x) { return (int32_t)(wbits(x) >> 32); }
__device__ __forceinline__ int type_of(uint32_t m) { return (int)(m & META_TYPE_MASK); }
__device__ __forceinline__ int inst_of(uint32_t m) { return (int)(m >> META_INST_SHIFT); }

// Workgroups are dispatched round-robin over the 8 XCDs; remap so each XCD
// walks a contiguous range of atoms and neighbour gathers stay in its L2.
__device__ __forceinline__ int xcd_remap(int w, int nb) {
  const int nx = 8;
  int q = nb / nx, r = nb % nx;
  int x = w % nx, loc = w / nx;
  return (x < r) ? x * (q + 1) + loc : r * (q + 1) + (x - r) * q + loc;
}

// Reference cell index of one coordinate (nlist.f90:29), no contraction.
__device__ __forceinline__ int ref_cell(double x, double ilb, double lb, double h) {
#pragma clang fp contract(off)
  double t = x * ilb;
  double f = floor(t);
  double w = x - f * lb;
  double q = w / h;
  return (int)q + 1;
}

// Reference minimum-image squared distance (si.f90:156-157), no contraction.
__device__ __forceinline__ double ref_r2(const double4 &a, const double4 &b, const DInst &in) {
#pragma clang fp contract(off)
  double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
  double x0 = d0 - round(d0 * in.ilb[0]) * in.lb[0];
  double x1 = d1 - round(d1 * in.ilb[1]) * in.lb[1];
  double x2 = d2 - round(d2 * in.ilb[2]) * in.lb[2];
  return x0 * x0 + x1 * x1 + x2 * x2;
}

// Minimum-image difference a - b (force kernels: contraction allowed).
__device__ __forceinline__ void min_image(const double4 &a, const double4 &b, const DInst &in,
                                          double &x0, double &x1, double &x2) {
  double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
  x0 = d0 - round(d0 * in.ilb[0]) * in.lb[0];
  x1 = d1 - round(d1 * in.ilb[1]) * in.lb[1];
  x2 = d2 - round(d2 * in.ilb[2]) * in.lb[2];
}

__device__ __forceinline__ int find_inst_by_cell(const DInst *I, int ninst, int gcell) {
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (I[mid].cell_real_off <= gcell) lo = mid; else hi = mid - 1;
  }
  return lo;
}
__device__ __forceinline__ int find_inst_by_bucket(const DInst *I, int ninst, int b) {
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (I[mid].bucket_off <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Compensated FP32 accumulation (TwoSum): s + c carries the exact sum of the
// added terms up to O(u^2) relative error.
__device__ __forceinline__ void kadd(float &s, float &c, float v) {
  float t = s + v;
  float bp = t - s;
  c += (s - (t - bp)) + (v - bp);
  s = t;
}

// ---------------------------------------------------------------------------
// Block reduction of NV doubles (sum or max per slot) into per-block
// partials, layout partials[k * pstride + block].  k_finalize combines them
// per instance in a fixed order, so results are deterministic.
template <int NV, int BS>
__device__ void block_partials(double (&v)[NV], const int (&op)[NV], int lblock, int pstride,
                               double *__restrict__ partials) {
  __shared__ double red[BS / FIB_WAVE][NV];
  const int lane = threadIdx.x % FIB_WAVE, wid = threadIdx.x / FIB_WAVE;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
    for (int o = FIB_WAVE / 2; o > 0; o >>= 1) {
      double y = __shfl_xor(x, o);
      x = op[k] ? fmax(x, y) : x + y;
    }
    v[k] = x;
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wid][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int k = threadIdx.x;
    double x = red[0][k];
    for (int w = 1; w < BS / FIB_WAVE; ++w) x = op[k] ? fmax(x, red[w][k]) : x + red[w][k];
    partials[(size_t)k * pstride + lblock] = x;
  }
}

#define FIN_T 256
// One block per instance: combine the instance's partials in a fixed order.
__global__ __launch_bounds__(FIN_T) void k_finalize(const DInst *__restrict__ I,
                                                    const double *__restrict__ partials,
                                                    int pstride, RedSpec rs,
                                                    DCtl *__restrict__ ctl) {
  __shared__ double red[FIN_T / FIB_WAVE];
  const int inst = blockIdx.x;
  const int blk0 = I[inst].atom_off / rs.bs, nblk = I[inst].natm_pad / rs.bs;
  const int lane = threadIdx.x % FIB_WAVE, wid = threadIdx.x / FIB_WAVE;
  double *out = reinterpret_cast<double *>(&ctl[inst]);
  for (int k = 0; k < rs.nv; ++k) {
    const bool mx = rs.op[k] != 0;
    const double *p = partials + (size_t)k * pstride + blk0;
    double x = mx ? -1.0e300 : 0.0;
    for (int b = threadIdx.x; b < nblk; b += FIN_T) x = mx ? fmax(x, p[b]) : x + p[b];
    for (int o = FIB_WAVE / 2; o > 0; o >>= 1) {
      double y = __shfl_xor(x, o);
      x = mx ? fmax(x, y) : x + y;
    }
    if (lane == 0) red[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = red[0];
      for (int w = 1; w < FIN_T / FIB_WAVE; ++w) t = mx ? fmax(t, red[w]) : t + red[w];
      if (rs.dst[k] >= 0) out[rs.dst[k]] = t;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// binning
__global__ void k_bin(const double4 *__restrict__ X, int NA, const DInst *__restrict__ I,
                      int *__restrict__ bkey, int *__restrict__ brank, int *__restrict__ bcnt,
                      DCtl *__restrict__ ctl) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NA) return;
  double4 x = X[p];
  uint32_t m = meta_of(x);
  if (m & META_PAD) { bkey[p] = -1; return; }
  const DInst &in = I[inst_of(m)];
  int local;
  if (m & META_DEAD) {
    local = in.nbucket - 1;
  } else if (!(isfinite(x.x) && isfinite(x.y) && isfinite(x.z))) {
    local = in.nbucket - 1;
    atomicOr(&ctl[inst_of(m)].nonfinite, 1);
  } else {
    int c0 = ref_cell(x.x, in.ilb[0], in.lb[0], in.h[0]);
    int c1 = ref_cell(x.y, in.ilb[1], in.lb[1], in.h[1]);
    int c2 = ref_cell(x.z, in.ilb[2], in.lb[2], in.h[2]);
    // Cell N+1 with N > 2 is overwritten by the periodic halo copy of
    // chainlist (nlist.f90:53-64): such atoms drop out of every list.
    bool drop = (c0 == in.nc[0] + 1 && in.nc[0] > 2) || (c1 == in.nc[1] + 1 && in.nc[1] > 2) ||
                (c2 == in.nc[2] + 1 && in.nc[2] > 2) || c0 < 1 || c1 < 1 || c2 < 1 ||
                c0 > in.nc[0] + 1 || c1 > in.nc[1] + 1 || c2 > in.nc[2] + 1;
    if (drop) local = in.nbucket - 2;
    else local = ((c2 - 1) * in.next[1] + (c1 - 1)) * in.next[0] + (c0 - 1);
  }
  int b = in.bucket_off + local;
  bkey[p] = b;
  brank[p] = atomicAdd(&bcnt[b], 1);
}

// Export of the reference cell ids in original atom order.
__global__ void k_cellid(const double4 *__restrict__ X, int NA, const DInst *__restrict__ I,
                         int inst, int32_t *__restrict__ out) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NA) return;
  double4 x = X[p];
  uint32_t m = meta_of(x);
  if ((m & META_PAD) || inst_of(m) != inst) return;
  const DInst &in = I[inst];
  int o = orig_of(x);
  out[o] = ref_cell(x.x, in.ilb[0], in.lb[0], in.h[0]);
  out[in.natm + o] = ref_cell(x.y, in.ilb[1], in.lb[1], in.h[1]);
  out[2 * in.natm + o] = ref_cell(x.z, in.ilb[2], in.lb[2], in.h[2]);
}

// ---------------------------------------------------------------------------
// exclusive scan (3 phases), chunk = 256 threads x 8 elements
#define SCAN_T 256
#define SCAN_E 8
#define SCAN_CH (SCAN_T * SCAN_E)

__device__ __forceinline__ int block_scan_excl(int v, int &total) {
  __shared__ int ws[SCAN_T / FIB_WAVE];
  const int lane = threadIdx.x % FIB_WAVE, wid = threadIdx.x / FIB_WAVE;
  int x = v;
  for (int o = 1; o < FIB_WAVE; o <<= 1) {
    int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == FIB_WAVE - 1) ws[wid] = x;
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < SCAN_T / FIB_WAVE; ++w) {
    if (w < wid) off += ws[w];
    tot += ws[w];
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

__global__ void k_scan_reduce(const int *__restrict__ in, int n, int *__restrict__ bsum) {
  int base = blockIdx.x * SCAN_CH;
  int s = 0;
  for (int e = 0; e < SCAN_E; ++e) {
    int idx = base + e * SCAN_T + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  int tot;
  block_scan_excl(s, tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void k_scan_top(int *__restrict__ bsum, int nb) {
  int carry = 0;
  for (int base = 0; base < nb; base += SCAN_T) {
    int idx = base + threadIdx.x;
    int v = idx < nb ? bsum[idx] : 0;
    int tot;
    int ex = block_scan_excl(v, tot);
    if (idx < nb) bsum[idx] = ex + carry;
    carry += tot;
    __syncthreads();
  }
}

// out[b] = exclusive prefix + pad_before of b's instance
__global__ void k_scan_down(const int *__restrict__ in, int n, const int *__restrict__ bsum,
                            const DInst *__restrict__ I, int ninst, int *__restrict__ out) {
  int base = blockIdx.x * SCAN_CH + threadIdx.x * SCAN_E;
  int vals[SCAN_E];
  int s = 0;
  for (int e = 0; e < SCAN_E; ++e) {
    int idx = base + e;
    vals[e] = idx < n ? in[idx] : 0;
    s += vals[e];
  }
  int tot;
  int ex = block_scan_excl(s, tot) + bsum[blockIdx.x];
  for (int e = 0; e < SCAN_E; ++e) {
    int idx = base + e;
    if (idx < n) {
      int k = find_inst_by_bucket(I, ninst, idx);
      out[idx] = ex + I[k].pad_before;
    }
    ex += vals[e];
  }
}

__global__ void k_scatter(const int *__restrict__ bkey, const int *__restrict__ brank, int NA,
                          const int *__restrict__ cstart, int *__restrict__ perm) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NA) return;
  int b = bkey[p];
  if (b < 0) return;
  perm[cstart[b] + brank[p]] = p;
}

// Deterministic order inside a bucket: ascending previous position.
__global__ void k_fixup(const int *__restrict__ cstart, const int *__restrict__ bcnt, int nb,
                        int *__restrict__ perm) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  int n = bcnt[b];
  if (n < 2) return;
  int *a = perm + cstart[b];
  for (int i = 1; i < n; ++i) {
    int v = a[i], j = i - 1;
    while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; --j; }
    a[j + 1] = v;
  }
}

__global__ void k_reorder(int NA, const int *__restrict__ perm, const DInst *__restrict__ I,
                          const double4 *__restrict__ X0, double4 *__restrict__ X1,
                          const double *__restrict__ V0, double *__restrict__ V1,
                          const double *__restrict__ F0, double *__restrict__ F1,
                          const double *__restrict__ Xi0, double *__restrict__ Xi1,
                          int *__restrict__ inv) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NA) return;
  uint32_t mp = meta_of(X0[p]);
  if (mp & META_PAD) return;
  int s = perm[p];
  double4 x = X0[s];
  X1[p] = x;
  for (int c = 0; c < 3; ++c) {
    size_t a = (size_t)c * NA;
    V1[a + p] = V0[a + s];
    F1[a + p] = F0[a + s];
    Xi1[a + p] = Xi0[a + s];
  }
  const DInst &in = I[inst_of(mp)];
  inv[in.atom_off + orig_of(x)] = p;
}

// ---------------------------------------------------------------------------
// neighbour build: one 64-lane block per real cell
#define NB_CHUNK 512

__global__ __launch_bounds__(FIB_WAVE) void k_nbuild(
    const double4 *__restrict__ X, const DInst *__restrict__ I, int ninst, int total_cells,
    const int *__restrict__ cstart, const int *__restrict__ bcnt, double rv2, double rv2lj,
    int NA, int Wsw, int Wion, int *__restrict__ cnt_sw, int *__restrict__ nbr_sw,
    int *__restrict__ cnt_ion, int *__restrict__ nbr_ion, DCtl *__restrict__ ctl) {
  __shared__ double4 s_x[NB_CHUNK];
  __shared__ int s_j[NB_CHUNK];
  __shared__ int s_rs[27], s_off[28];
  const int lane = threadIdx.x;
  const int gcell = blockIdx.x;
  if (gcell >= total_cells) return;
  const int k = find_inst_by_cell(I, ninst, gcell);
  const DInst &in = I[k];
  const int N0 = in.nc[0], N1 = in.nc[1];
  const int loc = gcell - in.cell_real_off;
  const int cx = loc % N0, cy = (loc / N0) % N1, cz = loc / (N0 * N1);
  const int bi = in.bucket_off + (cz * in.next[1] + cy) * in.next[0] + cx;
  const int i0 = cstart[bi], ni = bcnt[bi];
  if (ni == 0) return;

  // 27 stencil ranges.  N > 2: periodic.  N <= 2: no wrap; index N is the
  // phantom layer holding atoms whose cell id came out as N+1.
  int rs = 0, rc = 0;
  if (lane < 27) {
    int d[3] = {lane % 3 - 1, (lane / 3) % 3 - 1, lane / 9 - 1};
    int c[3] = {cx, cy, cz};
    bool ok = true;
    int e[3];
    for (int a = 0; a < 3; ++a) {
      int n = in.nc[a];
      int v = c[a] + d[a];
      if (n > 2) v = (v + n) % n;
      else if (v < 0 || v > n) ok = false;
      e[a] = v;
    }
    if (ok) {
      int b = in.bucket_off + (e[2] * in.next[1] + e[1]) * in.next[0] + e[0];
      rs = cstart[b];
      rc = bcnt[b];
    }
  }
  // inclusive scan of rc over lanes
  int sc = rc;
  for (int o = 1; o < FIB_WAVE; o <<= 1) {
    int y = __shfl_up(sc, o);
    if (lane >= o) sc += y;
  }
  if (lane < 27) { s_rs[lane] = rs; s_off[lane + 1] = sc; }
  if (lane == 0) s_off[0] = 0;
  const int total = __shfl(sc, 26);
  __syncthreads();

  const double4 *Xc = X;
  for (int ii = 0; ii < ni; ++ii) {
    const int i = i0 + ii;
    const double4 xi = Xc[i];
    const int ti = type_of(meta_of(xi));
    int nsw = 0, nion = 0;
    for (int base = 0; base < total; base += NB_CHUNK) {
      const int n = min(NB_CHUNK, total - base);
      // (re)stage the chunk; a single chunk is loaded once per cell
      if (ii == 0 || total > NB_CHUNK) {
        __syncthreads();
        for (int q = lane; q < n; q += FIB_WAVE) {
          int g = base + q;
          int r = 0;
          while (s_off[r + 1] <= g) ++r;
          int j = s_rs[r] + (g - s_off[r]);
          s_j[q] = j;
          s_x[q] = Xc[j];
        }
        __syncthreads();
      }
      for (int q0 = 0; q0 < n; q0 += FIB_WAVE) {
        const int q = q0 + lane;
        bool sw = false, io = false;
        int j = -1;
        if (q < n) {
          j = s_j[q];
          if (j != i) {
            const double4 xj = s_x[q];
            const int tj = type_of(meta_of(xj));
            const double r2 = ref_r2(xi, xj, in);
            if (max(ti, tj) < 3) sw = r2 < rv2;
            else io = r2 < rv2lj;
          }
        }
        uint64_t msw = __ballot(sw), mio = __ballot(io);
        uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        if (sw) {
          int pos = nsw + __popcll(msw & below);
          if (pos < Wsw) nbr_sw[(size_t)pos * NA + i] = j;
        }
        if (io) {
          int pos = nion + __popcll(mio & below);
          if (pos < Wion) nbr_ion[(size_t)pos * NA + i] = j;
        }
        nsw += __popcll(msw);
        nion += __popcll(mio);
      }
    }
    if (lane == 0) {
      cnt_sw[i] = nsw;
      cnt_ion[i] = nion;
    }
  }
}

// Atoms outside every scanned cell (dropped, dead, or phantom i-atoms) get
// empty lists; k_nbuild only writes the atoms of real cells.
__global__ void k_clear_counts(int NA, int *__restrict__ cnt_sw, int *__restrict__ cnt_ion) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NA) return;
  cnt_sw[p] = 0;
  cnt_ion[p] = 0;
}

// Reverse slots plus the longest lists of the build (overflow check).
__global__ __launch_bounds__(FIB_BLOCK) void k_rev(int NA, int Wsw, const int *__restrict__ cnt_sw,
                                                   const int *__restrict__ cnt_ion,
                                                   const int *__restrict__ nbr_sw,
                                                   uint8_t *__restrict__ rev, int pstride,
                                                   double *__restrict__ partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int cs = cnt_sw[i];
  const int c = min(cs, Wsw);
  for (int k = 0; k < c; ++k) {
    int j = nbr_sw[(size_t)k * NA + i];
    int cj = min(cnt_sw[j], Wsw);
    int r = 255;
    for (int q = 0; q < cj; ++q)
      if (nbr_sw[(size_t)q * NA + j] == i) { r = q; break; }
    rev[(size_t)k * NA + i] = (uint8_t)r;
  }
  double v[2] = {(double)cs, (double)cnt_ion[i]};
  const int op[2] = {1, 1};
  block_partials<2, FIB_BLOCK>(v, op, blockIdx.x, pstride, partials);
}

// ---------------------------------------------------------------------------
// force phase A
template <typename T>
struct Slot {  // one active SW leg of the centre atom
  T e[3];      // unit vector from neighbour to centre
  T rinv;      // 1/r (FP64: m^-1; mixed: sigma^-1 units)
  T E;         // exp(gamma/(r/sigma - a))
  T g;         // d ln E / dr
  int k;       // slot in the buffered list | type << 8
};

template <bool MIXED>
struct Prec;
template <>
struct Prec<false> {
  typedef double T;
  typedef double4 T4;
  static constexpr int MAXS = 8;
};
template <>
struct Prec<true> {
  typedef float T;
  typedef float4 T4;
  static constexpr int MAXS = 12;
};

// One ordered three-body term: centre c, legs a and b.  Returns the force on
// a (F_a = -dh/dx_a) and h.  Derivation: h = leps E(ra) E(rb) (cos+1/3)^2,
// d cos/dx_a = -(e_b - cos e_a)/r_a, d r_a/dx_a = -e_a.
template <typename T>
__device__ __forceinline__ void vterm(T leps, const T *ea, T rinva, T Ea, T ga, const T *eb,
                                      T Eb, T *Fa, T &hh) {
  T cs = ea[0] * eb[0] + ea[1] * eb[1] + ea[2] * eb[2];
  T D = cs + T(1.0 / 3.0);
  T Eab = leps * Ea * Eb;
  hh = Eab * D * D;
  T c1 = hh * ga;
  T c2 = T(2) * Eab * D * rinva;
#pragma unroll
  for (int c = 0; c < 3; ++c) Fa[c] = c1 * ea[c] + c2 * (eb[c] - cs * ea[c]);
}

template <bool MIXED, bool EN, bool VIR>
__global__ __launch_bounds__(FIB_WAVE) void k_force_a(
    const double4 *__restrict__ X, const DInst *__restrict__ I, const FConst fc,
    const TabEntry *__restrict__ tab, int NA, int Wsw, int Wion, const int *__restrict__ cnt_sw,
    const int *__restrict__ nbr_sw, const int *__restrict__ cnt_ion,
    const int *__restrict__ nbr_ion, typename Prec<MIXED>::T4 *__restrict__ G,
    uint64_t *__restrict__ amask, double *__restrict__ Fpart, double *__restrict__ upart,
    int pstride, double *__restrict__ partials, DCtl *__restrict__ ctl) {
  typedef typename Prec<MIXED>::T T;
  typedef typename Prec<MIXED>::T4 T4;
  constexpr int MAXS = Prec<MIXED>::MAXS;
  __shared__ Slot<T> s_sl[MAXS][FIB_WAVE];

  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int lane = threadIdx.x;
  const int i = lb * FIB_WAVE + lane;
  const double4 xi = X[i];
  const uint32_t mi = meta_of(xi);
  const int kin = inst_of(mi);
  const DInst &in = I[kin];
  const bool live = !(mi & (META_PAD | META_DEAD));
  const int ti = type_of(mi);

  double fx = 0, fy = 0, fz = 0;     // FP64 accumulation (FP64 path)
  float sx = 0, sy = 0, sz = 0, cx = 0, cy = 0, cz = 0;  // compensated FP32 (mixed)
  double ui = 0, etot = 0;
  double W[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) W[q] = 0;
  uint64_t am = 0;
  int na = 0;

  auto addF = [&](double vx, double vy, double vz) {
    if (MIXED) { kadd(sx, cx, (float)vx); kadd(sy, cy, (float)vy); kadd(sz, cz, (float)vz); }
    else { fx += vx; fy += vy; fz += vz; }
  };

  if (live && ti < 3) {
    const int cs = min(cnt_sw[i], Wsw);
    for (int k = 0; k < cs; ++k) {
      const int j = nbr_sw[(size_t)k * NA + i];
      const double4 xj = X[j];
      const int tj = type_of(meta_of(xj));
      double d0, d1, d2;
      min_image(xi, xj, in, d0, d1, d2);
      const double isg = fc.isig[ti - 1][tj - 1];
      if (!MIXED) {
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        const double r = sqrt(r2);
        const double rinv = 1.0 / r;
        const double rs = r * isg;
        // reference pair force: 3-point Lagrange on the derivative table
        const int t1 = (int)(rs / fc.umax * fc.tabn1) + 1;
        if (t1 < fc.izero) {
          const TabEntry te = tab[t1];
          const double e1 = rs - te.u0, e2 = rs - te.u1, e3 = rs - te.u2;
          const double df = 0.5 * e2 * e3 * te.t0 - e1 * e3 * te.t1 + 0.5 * e1 * e2 * te.t2;
          const double fs = -df * fc.epssig[ti - 1][tj - 1] * rinv;
          const double f0 = fs * d0, f1 = fs * d1, f2 = fs * d2;
          fx += f0; fy += f1; fz += f2;
          if (VIR) {
            W[0] += 0.5 * d0 * f0; W[1] += 0.5 * d0 * f1; W[2] += 0.5 * d0 * f2;
            W[3] += 0.5 * d1 * f0; W[4] += 0.5 * d1 * f1; W[5] += 0.5 * d1 * f2;
            W[6] += 0.5 * d2 * f0; W[7] += 0.5 * d2 * f1; W[8] += 0.5 * d2 * f2;
          }
        }
        if (rs < fc.cut_sw) {
          const double inv_ra = 1.0 / (rs - fc.al);
          const double E = exp(fc.gamma * inv_ra);
          if (EN) {
            const double rs2 = rs * rs;
            const double ul = 0.5 * fc.eps[ti - 1][tj - 1] * fc.A * (fc.B / (rs2 * rs2) - 1.0) *
                              exp(inv_ra);
            ui += ul;
            etot += ul;
          }
          if (na < MAXS) {
            Slot<T> &s = s_sl[na][lane];
            s.e[0] = d0 * rinv; s.e[1] = d1 * rinv; s.e[2] = d2 * rinv;
            s.rinv = rinv; s.E = E; s.g = -fc.gamma * inv_ra * inv_ra * isg;
            s.k = k | (tj << 8);
          }
          am |= 1ull << k;
          ++na;
        }
      } else {
        // mixed: FP64 difference, reduced units (sigma) in FP32
        const float r0 = (float)(d0 * isg), r1 = (float)(d1 * isg), r2f = (float)(d2 * isg);
        const float rr = r0 * r0 + r1 * r1 + r2f * r2f;
        const float rs = sqrtf(rr);
        const float rinv = 1.0f / rs;
        if (rs < (float)fc.cut_sw) {
          const float al = (float)fc.al;
          const float inv_ra = 1.0f / (rs - al);
          const float ex = __expf(inv_ra);
          const float B = (float)fc.B;
          const float A = (float)fc.A;
          const float ir2 = rinv * rinv;
          const float ir4 = ir2 * ir2;
          // phi2(r) = A (B r^-4 - 1) e^{1/(r-a)};  dphi/dr
          const float dphi = A * ex * (-4.0f * B * ir4 * rinv - (B * ir4 - 1.0f) * inv_ra * inv_ra);
          const float es = (float)fc.epssig[ti - 1][tj - 1];
          // force on i along +d: -dphi * eps/sigma * e
          const float fs = -dphi * rinv;  // units eps/sigma per sigma
          const float f0 = fs * r0, f1 = fs * r1, f2 = fs * r2f;
          const double fsc = (double)es;
          addF(f0 * fsc, f1 * fsc, f2 * fsc);
          if (VIR) {
            const double sg = 1.0 / isg;
            const double w0 = 0.5 * sg * fsc;
            W[0] += w0 * r0 * f0; W[1] += w0 * r0 * f1; W[2] += w0 * r0 * f2;
            W[3] += w0 * r1 * f0; W[4] += w0 * r1 * f1; W[5] += w0 * r1 * f2;
            W[6] += w0 * r2f * f0; W[7] += w0 * r2f * f1; W[8] += w0 * r2f * f2;
          }
          if (EN) {
            const double ul = 0.5 * fc.eps[ti - 1][tj - 1] * (double)(A * (B * ir4 - 1.0f) * ex);
            ui += ul;
            etot += ul;
          }
          const float E = __expf((float)fc.gamma * inv_ra);
          if (na < MAXS) {
            Slot<T> &s = s_sl[na][lane];
            s.e[0] = r0 * rinv; s.e[1] = r1 * rinv; s.e[2] = r2f * rinv;
            s.rinv = rinv; s.E = E; s.g = -(float)fc.gamma * inv_ra * inv_ra;
            s.k = k | (tj << 8);
          }
          am |= 1ull << k;
          ++na;
        }
      }
    }
  }

  // Moliere pairs (si.f90:426-446): both atoms must be fired.
  if (live) {
    const int ci = min(cnt_ion[i], Wion);
    const bool ifired = !(ti == 3 && orig_of(xi) >= in.nsg + in.impact);
    for (int k = 0; ifired && k < ci; ++k) {
      const int j = nbr_ion[(size_t)k * NA + i];
      const double4 xj = X[j];
      const uint32_t mj = meta_of(xj);
      const int tj = type_of(mj);
      if (tj == 3 && orig_of(xj) >= in.nsg + in.impact) continue;
      if (mj & META_DEAD) continue;
      double d0, d1, d2;
      min_image(xi, xj, in, d0, d1, d2);
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      if (!(r2 < fc.rlj2)) continue;
      const double r = sqrt(r2);
      const double aa = fc.aamol[ti - 1][tj - 1];
      const double zz = fc.ie0 * fc.zmol[ti - 1] * fc.zmol[tj - 1] * fc.ec * fc.ec;
      if (!MIXED || r < fc.mixed_r64) {
        const double x = r / aa;
        const double e1 = exp(-fc.molb[0] * x), e2 = exp(-fc.molb[1] * x),
                     e3 = exp(-fc.molb[2] * x);
        const double sa = fc.molc[0] * e1 + fc.molc[1] * e2 + fc.molc[2] * e3;
        const double sb = (fc.molc[0] * fc.molb[0] * e1 + fc.molc[1] * fc.molb[1] * e2 +
                           fc.molc[2] * fc.molb[2] * e3) / aa;
        const double fcm = zz / r * (sa / r + sb);
        const double rinv = 1.0 / r;
        const double f0 = fcm * d0 * rinv, f1 = fcm * d1 * rinv, f2 = fcm * d2 * rinv;
        addF(f0, f1, f2);
        if (EN) { const double ul = 0.5 * zz / r * sa; ui += ul; etot += ul; }
        if (VIR) {
          W[0] += 0.5 * d0 * f0; W[1] += 0.5 * d0 * f1; W[2] += 0.5 * d0 * f2;
          W[3] += 0.5 * d1 * f0; W[4] += 0.5 * d1 * f1; W[5] += 0.5 * d1 * f2;
          W[6] += 0.5 * d2 * f0; W[7] += 0.5 * d2 * f1; W[8] += 0.5 * d2 * f2;
        }
      } else {
        // FP32: lengths in units of aa, force scaled by zz/aa^2
        const float x = (float)(r / aa);
        const float e1 = __expf(-(float)fc.molb[0] * x), e2 = __expf(-(float)fc.molb[1] * x),
                    e3 = __expf(-(float)fc.molb[2] * x);
        const float c0 = (float)fc.molc[0], c1 = (float)fc.molc[1], c2 = (float)fc.molc[2];
        const float sa = c0 * e1 + c1 * e2 + c2 * e3;
        const float sb = c0 * (float)fc.molb[0] * e1 + c1 * (float)fc.molb[1] * e2 +
                         c2 * (float)fc.molb[2] * e3;
        const float ix = 1.0f / x;
        const float fr = ix * (sa * ix + sb) * ix;  // (zz/aa^2) units, divided by x for d
        const double scale = zz / (aa * aa) / aa;   // times d (m) gives N
        const double f0 = (double)(fr * (float)(d0 / aa)) * scale * aa;
        const double f1 = (double)(fr * (float)(d1 / aa)) * scale * aa;
        const double f2 = (double)(fr * (float)(d2 / aa)) * scale * aa;
        addF(f0, f1, f2);
        if (EN) { const double ul = 0.5 * zz / aa * (double)(sa * ix); ui += ul; etot += ul; }
        if (VIR) {
          W[0] += 0.5 * d0 * f0; W[1] += 0.5 * d0 * f1; W[2] += 0.5 * d0 * f2;
          W[3] += 0.5 * d1 * f0; W[4] += 0.5 * d1 * f1; W[5] += 0.5 * d1 * f2;
          W[6] += 0.5 * d2 * f0; W[7] += 0.5 * d2 * f1; W[8] += 0.5 * d2 * f2;
        }
      }
    }
  }

  // Three-body terms centred on i (si.f90:451-549 restated per vertex).
  // na == 1 still writes a zero G entry: the neighbour gathers it.
  if (live && ti < 3 && na >= 1) {
    const int cs = min(cnt_sw[i], Wsw);
    // unit factor: FP64 path works in m and N; mixed path in sigma and eps/sigma
    const double fscale = MIXED ? fc.eps[0][0] * fc.isig[0][0] : 1.0;
    T fcx = 0, fcy = 0, fcz = 0;
    if (na <= MAXS) {
      for (int a = 0; a < na; ++a) {
        const Slot<T> sa = s_sl[a][lane];
        const int ta = sa.k >> 8;
        T Ga[3] = {0, 0, 0};
        T Ua = 0;
        for (int b = 0; b < na; ++b) {
          if (b == a) continue;
          const Slot<T> sb = s_sl[b][lane];
          const int tb = sb.k >> 8;
          const T lep = MIXED ? (T)(fc.leps[ti - 1][ta - 1][tb - 1] / fc.eps[0][0])
                              : (T)fc.leps[ti - 1][ta - 1][tb - 1];
          T Fa[3], hh;
          vterm<T>(lep, sa.e, sa.rinv, sa.E, sa.g, sb.e, sb.E, Fa, hh);
          Ga[0] += Fa[0]; Ga[1] += Fa[1]; Ga[2] += Fa[2];
          if (EN) Ua += hh * T(1.0 / 3.0);
          if (EN) { ui += (double)hh * (1.0 / 6.0) * (MIXED ? fc.eps[0][0] : 1.0);
                    etot += (double)hh * 0.5 * (MIXED ? fc.eps[0][0] : 1.0); }
        }
        fcx -= Ga[0]; fcy -= Ga[1]; fcz -= Ga[2];
        if (VIR) {
          // r_ca = x_a - x_c = -e_a / rinv_a
          const double rl = -1.0 / (double)sa.rinv * (MIXED ? 1.0 / fc.isig[0][0] : 1.0);
          const double g0 = (double)Ga[0] * fscale, g1 = (double)Ga[1] * fscale,
                       g2 = (double)Ga[2] * fscale;
          const double e0 = rl * sa.e[0], e1 = rl * sa.e[1], e2 = rl * sa.e[2];
          W[0] += e0 * g0; W[1] += e0 * g1; W[2] += e0 * g2;
          W[3] += e1 * g0; W[4] += e1 * g1; W[5] += e1 * g2;
          W[6] += e2 * g0; W[7] += e2 * g1; W[8] += e2 * g2;
        }
        const int ka = sa.k & 0xff;
        T4 gv;
        gv.x = Ga[0]; gv.y = Ga[1]; gv.z = Ga[2];
        gv.w = EN ? Ua : T(0);
        G[(size_t)ka * NA + i] = gv;
      }
    } else {
      // more active legs than LDS slots: recompute leg geometry on the fly
      auto leg = [&](int k, T *e, T &rinv, T &E, T &g, int &tj) {
        const int j = nbr_sw[(size_t)k * NA + i];
        const double4 xj = X[j];
        tj = type_of(meta_of(xj));
        double d0, d1, d2;
        min_image(xi, xj, in, d0, d1, d2);
        const double isg = fc.isig[ti - 1][tj - 1];
        if (!MIXED) {
          const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
          const double ri = 1.0 / r;
          const double rs = r * isg;
          const double inv_ra = 1.0 / (rs - fc.al);
          e[0] = d0 * ri; e[1] = d1 * ri; e[2] = d2 * ri;
          rinv = ri; E = exp(fc.gamma * inv_ra); g = -fc.gamma * inv_ra * inv_ra * isg;
        } else {
          const float r0 = (float)(d0 * isg), r1 = (float)(d1 * isg), r2f = (float)(d2 * isg);
          const float rs = sqrtf(r0 * r0 + r1 * r1 + r2f * r2f);
          const float ri = 1.0f / rs;
          const float inv_ra = 1.0f / (rs - (float)fc.al);
          e[0] = r0 * ri; e[1] = r1 * ri; e[2] = r2f * ri;
          rinv = ri; E = __expf((float)fc.gamma * inv_ra); g = -(float)fc.gamma * inv_ra * inv_ra;
        }
      };
      for (int ka = 0; ka < cs; ++ka) {
        if (!((am >> ka) & 1ull)) continue;
        T ea[3], ra, Ea, ga;
        int ta;
        leg(ka, ea, ra, Ea, ga, ta);
        T Ga[3] = {0, 0, 0};
        T Ua = 0;
        for (int kb = 0; kb < cs; ++kb) {
          if (kb == ka || !((am >> kb) & 1ull)) continue;
          T eb[3], rb, Eb, gb;
          int tb;
          leg(kb, eb, rb, Eb, gb, tb);
          const T lep = MIXED ? (T)(fc.leps[ti - 1][ta - 1][tb - 1] / fc.eps[0][0])
                              : (T)fc.leps[ti - 1][ta - 1][tb - 1];
          T Fa[3], hh;
          vterm<T>(lep, ea, ra, Ea, ga, eb, Eb, Fa, hh);
          Ga[0] += Fa[0]; Ga[1] += Fa[1]; Ga[2] += Fa[2];
          if (EN) Ua += hh * T(1.0 / 3.0);
          if (EN) { ui += (double)hh * (1.0 / 6.0) * (MIXED ? fc.eps[0][0] : 1.0);
                    etot += (double)hh * 0.5 * (MIXED ? fc.eps[0][0] : 1.0); }
        }
        fcx -= Ga[0]; fcy -= Ga[1]; fcz -= Ga[2];
        if (VIR) {
          const double rl = -1.0 / (double)ra * (MIXED ? 1.0 / fc.isig[0][0] : 1.0);
          const double g0 = (double)Ga[0] * fscale, g1 = (double)Ga[1] * fscale,
                       g2 = (double)Ga[2] * fscale;
          const double e0 = rl * ea[0], e1 = rl * ea[1], e2 = rl * ea[2];
          W[0] += e0 * g0; W[1] += e0 * g1; W[2] += e0 * g2;
          W[3] += e1 * g0; W[4] += e1 * g1; W[5] += e1 * g2;
          W[6] += e2 * g0; W[7] += e2 * g1; W[8] += e2 * g2;
        }
        T4 gv;
        gv.x = Ga[0]; gv.y = Ga[1]; gv.z = Ga[2];
        gv.w = EN ? Ua : T(0);
        G[(size_t)ka * NA + i] = gv;
      }
    }
    addF((double)fcx * fscale, (double)fcy * fscale, (double)fcz * fscale);
  }

  if (MIXED) {
    fx = (double)sx + (double)cx;
    fy = (double)sy + (double)cy;
    fz = (double)sz + (double)cz;
  }
  Fpart[i] = fx;
  Fpart[(size_t)NA + i] = fy;
  Fpart[2 * (size_t)NA + i] = fz;
  amask[i] = am;
  if (EN) upart[i] = ui;

  if (EN || VIR) {
    double v[10];
    v[0] = etot;
#pragma unroll
    for (int q = 0; q < 9; ++q) v[1 + q] = W[q];
    const int op[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    block_partials<10, FIB_WAVE>(v, op, lb, pstride, partials);
  }
}

// ---------------------------------------------------------------------------
// force phase B: gather neighbour contributions; MODE 0 = store forces,
// MODE 1 = reference velocity update (timeint.f90:231-236) + thermostat sums.
template <bool MIXED, bool EN, int MODE>
__global__ __launch_bounds__(FIB_BLOCK) void k_force_b(
    const double4 *__restrict__ X, const DInst *__restrict__ I, int NA, int Wsw,
    const int *__restrict__ nbr_sw, const uint8_t *__restrict__ rev,
    const typename Prec<MIXED>::T4 *__restrict__ G, const uint64_t *__restrict__ amask,
    const double *__restrict__ Fpart, const double *__restrict__ upart, double fscale,
    double escale, double *__restrict__ Fout, double *__restrict__ uout, double *__restrict__ V,
    StepArgs sa, int pstride, double *__restrict__ partials, DCtl *__restrict__ ctl) {
  typedef typename Prec<MIXED>::T4 T4;
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int i = lb * FIB_BLOCK + threadIdx.x;
  const double4 xi = X[i];
  const uint32_t mi = meta_of(xi);
  const int kin = inst_of(mi);
  const DInst &in = I[kin];
  const bool live = !(mi & (META_PAD | META_DEAD));
  double f0 = 0, f1 = 0, f2 = 0, u = 0;
  if (live) {
    f0 = Fpart[i]; f1 = Fpart[(size_t)NA + i]; f2 = Fpart[2 * (size_t)NA + i];
    if (EN) u = upart[i];
    uint64_t m = amask[i];
    double g0 = 0, g1 = 0, g2 = 0, gu = 0;
    while (m) {
      const int k = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const int j = nbr_sw[(size_t)k * NA + i];
      const int r = rev[(size_t)k * NA + i];
      // r == 255: j never lists i (a phantom-layer atom, see k_nbuild)
      if (r >= Wsw) continue;
      const T4 g = G[(size_t)r * NA + j];
      g0 += (double)g.x; g1 += (double)g.y; g2 += (double)g.z;
      if (EN) gu += (double)g.w;
    }
    f0 += g0 * fscale; f1 += g1 * fscale; f2 += g2 * fscale;
    if (EN) u += gu * escale;
  }
  if (MODE == 0) {
    if (live) {
      Fout[i] = f0; Fout[(size_t)NA + i] = f1; Fout[2 * (size_t)NA + i] = f2;
      if (EN) uout[i] = u;
    }
    return;
  }
  // MODE 1: V += 0.5*(F+Fp)*Ts/m ; F = Fp ; thermostat / temperature sums
  double v[5] = {0, 0, 0, 0, 0};
  if (live) {
#pragma clang fp contract(off)
    const int t = type_of(mi);
    const double m = sa.mass[t - 1];
    double vv[3];
    const double fn[3] = {f0, f1, f2};
    for (int c = 0; c < 3; ++c) {
      const size_t a = (size_t)c * NA + i;
      vv[c] = V[a] + 0.5 * (Fout[a] + fn[c]) * sa.ts / m;
      V[a] = vv[c];
      Fout[a] = fn[c];
    }
    if (t < 3) {
      const double m2 = m * (vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
      v[0] = m2;  // global temperature sum (temp.f90:196-198)
      const double nm = 1.0e-9;
      const bool xs = (xi.x < sa.sidewidth * nm) || (xi.x > (in.lb[0] - sa.sidewidth * nm));
      const bool ys = (xi.y < sa.sidewidth * nm) || (xi.y > (in.lb[1] - sa.sidewidth * nm));
      const bool zc = (xi.z < sa.outz[0] * in.lb[2] / 10.0) || (xi.z > sa.outz[3] * in.lb[2] / 10.0);
      if (xs || (ys && zc)) { v[1] = m2; v[2] = 1.0; }
    }
    v[3] = 0.5 * m * (vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
    v[4] = sqrt(vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
  }
  const int op[5] = {0, 0, 0, 0, 1};
  block_partials<5, FIB_BLOCK>(v, op, lb, pstride, partials);
}


// ---------------------------------------------------------------------------
// onestep head (timeint.f90:165-203): Berendsen scaling (adjusttemp
// temp.f90:160-179 or controlsidetemp :32-67), X += V*Ts + 0.5*F*Ts^2/m,
// freezsidesN (tools.f90:38-51), rebox (:26-36), frzsptrdatms (temp.f90:99-127).
// Reductions: sum m v^2 of the rescaled V (equilibration temperature) and
// max 0.5 m v^2/ec (calc_atm's test, timeint.f90:284-293).
__global__ __launch_bounds__(FIB_BLOCK) void k_pos(double4 *__restrict__ X,
                                                   const DInst *__restrict__ I, int NA,
                                                   double *__restrict__ V, double *__restrict__ F,
                                                   const double *__restrict__ Xi, StepArgs sa,
                                                   int pstride, double *__restrict__ partials,
                                                   DCtl *__restrict__ ctl) {
#pragma clang fp contract(off)
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int i = lb * FIB_BLOCK + threadIdx.x;
  double4 x = X[i];
  const uint32_t mi = meta_of(x);
  const int kin = inst_of(mi);
  const DInst &in = I[kin];
  const bool live = !(mi & (META_PAD | META_DEAD));
  double red[2] = {0.0, -1.0};
  if (live) {
    const int t = type_of(mi);
    const double m = sa.mass[t - 1];
    double v[3], f[3], xx[3] = {x.x, x.y, x.z};
    for (int c = 0; c < 3; ++c) {
      v[c] = V[(size_t)c * NA + i];
      f[c] = F[(size_t)c * NA + i];
    }
    if (sa.mode == 0) {
      if (t < 3) {
        const double s = sqrt(1. + sa.ts / sa.tau * (sa.ttar / sa.temp - 1.));
        for (int c = 0; c < 3; ++c) v[c] = s * v[c];
        red[0] = m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      }
    } else if (sa.mode == 1) {
      const DCtl &c0 = ctl[kin];
      if (t < 3 && c0.side_cnt > 0.0) {
        const double nm = 1.0e-9;
        const bool xs = (xx[0] < sa.sidewidth * nm) || (xx[0] > (in.lb[0] - sa.sidewidth * nm));
        const bool ys = (xx[1] < sa.sidewidth * nm) || (xx[1] > (in.lb[1] - sa.sidewidth * nm));
        const bool zc = (xx[2] < sa.outz[0] * in.lb[2] / 10.0) ||
                        (xx[2] > sa.outz[3] * in.lb[2] / 10.0);
        if (xs || (ys && zc)) {
          const double tside = c0.side_sum / 3. / c0.side_cnt / sa.kB;
          const double s = sqrt(1. + sa.ts / sa.tau * (sa.ttar / tside - 1.));
          for (int c = 0; c < 3; ++c) v[c] = s * v[c];
        }
      }
    }
    for (int c = 0; c < 3; ++c) xx[c] = xx[c] + v[c] * sa.ts + 0.5 * f[c] * sa.ts * sa.ts / m;
    if (mi & META_FROZEN) {
      for (int c = 0; c < 3; ++c) {
        xx[c] = Xi[(size_t)c * NA + i];
        v[c] = 0.;
      }
    }
    for (int c = 0; c < 3; ++c) xx[c] = fmod(xx[c] + 2. * in.lb[c], in.lb[c]);
    if (sa.frz) {
      const double T = in.lb[2];
      const double z = xx[2];
      const double buffer = 0.3;
      bool zero = false;
      if ((z > sa.outz[0] * T / 10.0) && (z < sa.outz[1] * T / 10.0)) {
        xx[2] = sa.atomz[0] * T / 10.0; zero = true;
      } else if ((z > sa.outz[2] * T / 10.0) && (z < sa.outz[3] * T / 10.0)) {
        xx[2] = sa.atomz[1] * T / 10.0; zero = true;
      } else if ((z > (sa.atomz[0] - buffer) * T / 10.0) && (z < (sa.atomz[0] + buffer) * T / 10.0)) {
        zero = true;
      } else if ((z > (sa.atomz[1] - buffer) * T / 10.0) && (z < (sa.atomz[1] + buffer) * T / 10.0)) {
        zero = true;
      } else if ((z > (sa.ionz - buffer) * T / 10.0) && (z < (sa.ionz + buffer) * T / 10.0)) {
        zero = true;
      }
      if (zero)
        for (int c = 0; c < 3; ++c) { v[c] = 0.0; f[c] = 0.0; }
    }
    red[1] = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / sa.ec;
    x.x = xx[0]; x.y = xx[1]; x.z = xx[2];
    X[i] = x;
    for (int c = 0; c < 3; ++c) {
      V[(size_t)c * NA + i] = v[c];
      F[(size_t)c * NA + i] = f[c];
    }
  }
  const int op[2] = {0, 1};
  block_partials<2, FIB_BLOCK>(red, op, lb, pstride, partials);
}

// Thermostat / temperature sums of the current state (same as k_force_b's
// MODE 1 reduction without the velocity update).
__global__ __launch_bounds__(FIB_BLOCK) void k_sums(const double4 *__restrict__ X,
                                                    const DInst *__restrict__ I, int NA,
                                                    const double *__restrict__ V, StepArgs sa,
                                                    int pstride, double *__restrict__ partials,
                                                    DCtl *__restrict__ ctl) {
  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int i = lb * FIB_BLOCK + threadIdx.x;
  const double4 xi = X[i];
  const uint32_t mi = meta_of(xi);
  const int kin = inst_of(mi);
  const DInst &in = I[kin];
  const bool live = !(mi & (META_PAD | META_DEAD));
  double v[5] = {0, 0, 0, 0, 0};
  if (live) {
    const int t = type_of(mi);
    const double m = sa.mass[t - 1];
    const double a = V[i], b = V[(size_t)NA + i], c = V[2 * (size_t)NA + i];
    const double vv = a * a + b * b + c * c;
    if (t < 3) {
      const double m2 = m * vv;
      v[0] = m2;
      const double nm = 1.0e-9;
      const bool xs = (xi.x < sa.sidewidth * nm) || (xi.x > (in.lb[0] - sa.sidewidth * nm));
      const bool ys = (xi.y < sa.sidewidth * nm) || (xi.y > (in.lb[1] - sa.sidewidth * nm));
      const bool zc = (xi.z < sa.outz[0] * in.lb[2] / 10.0) || (xi.z > sa.outz[3] * in.lb[2] / 10.0);
      if (xs || (ys && zc)) { v[1] = m2; v[2] = 1.0; }
    }
    v[3] = 0.5 * m * vv;
    v[4] = sqrt(vv);
  }
  const int op[5] = {0, 0, 0, 0, 1};
  block_partials<5, FIB_BLOCK>(v, op, lb, pstride, partials);
}

// ---------------------------------------------------------------------------
// host <-> device layout conversion
__global__ void k_import(int inst, DInst in, int NA, const double *__restrict__ hx,
                         const double *__restrict__ hv, const int32_t *__restrict__ at,
                         const int32_t *__restrict__ xs, const double *__restrict__ hxi,
                         double4 *__restrict__ X0, double4 *__restrict__ X1,
                         double *__restrict__ V0, double *__restrict__ V1,
                         double *__restrict__ F0, double *__restrict__ F1,
                         double *__restrict__ Xi0, double *__restrict__ Xi1,
                         int *__restrict__ inv) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= in.natm_pad) return;
  const int p = in.atom_off + o;
  double4 x;
  uint64_t w;
  double v[3] = {0, 0, 0}, xi[3] = {0, 0, 0};
  if (o < in.natm) {
    x.x = hx[o]; x.y = hx[in.natm + o]; x.z = hx[2 * in.natm + o];
    uint32_t m = (uint32_t)at[o] & META_TYPE_MASK;
    if (xs && xs[o] > 0) m |= META_FROZEN;
    m |= (uint32_t)inst << META_INST_SHIFT;
    w = ((uint64_t)(uint32_t)o << 32) | m;
    for (int c = 0; c < 3; ++c) {
      v[c] = hv ? hv[(size_t)c * in.natm + o] : 0.0;
      xi[c] = hxi ? hxi[(size_t)c * in.natm + o] : (c == 0 ? x.x : (c == 1 ? x.y : x.z));
    }
  } else {
    x.x = x.y = x.z = 0.0;
    w = ((uint64_t)(uint32_t)o << 32) | META_PAD | ((uint32_t)inst << META_INST_SHIFT);
  }
  x.w = __longlong_as_double((long long)w);
  X0[p] = x; X1[p] = x;
  for (int c = 0; c < 3; ++c) {
    const size_t a = (size_t)c * NA + p;
    V0[a] = v[c]; V1[a] = v[c];
    F0[a] = 0.0; F1[a] = 0.0;
    Xi0[a] = xi[c]; Xi1[a] = xi[c];
  }
  inv[p] = p;
}

__global__ void k_export_x(DInst in, const double4 *__restrict__ X, double *__restrict__ out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= in.natm_pad) return;
  const double4 x = X[in.atom_off + o];
  if (meta_of(x) & META_PAD) return;
  const int g = orig_of(x);
  out[g] = x.x; out[in.natm + g] = x.y; out[2 * in.natm + g] = x.z;
}

__global__ void k_export_vec(DInst in, int NA, const double4 *__restrict__ X,
                             const double *__restrict__ src, double *__restrict__ out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= in.natm_pad) return;
  const int p = in.atom_off + o;
  const double4 x = X[p];
  if (meta_of(x) & META_PAD) return;
  const int g = orig_of(x);
  for (int c = 0; c < 3; ++c) out[(size_t)c * in.natm + g] = src[(size_t)c * NA + p];
}

__global__ void k_export_scalar(DInst in, const double4 *__restrict__ X,
                                const double *__restrict__ src, double *__restrict__ out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= in.natm_pad) return;
  const int p = in.atom_off + o;
  const double4 x = X[p];
  if (meta_of(x) & META_PAD) return;
  out[orig_of(x)] = src[p];
}

__global__ void k_set_atom(int NA, double4 *__restrict__ X, double *__restrict__ V, int p,
                           double x0, double x1, double x2, double v0, double v1, double v2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double4 x = X[p];
  x.x = x0; x.y = x1; x.z = x2;
  X[p] = x;
  V[p] = v0; V[(size_t)NA + p] = v1; V[2 * (size_t)NA + p] = v2;
}

// ---------------------------------------------------------------------------
// launchers
static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

static void finalize(const DevBufs &d, int nv, int bs, const int *op, const int *dst,
                     hipStream_t s) {
  RedSpec rs;
  rs.nv = nv;
  rs.bs = bs;
  for (int k = 0; k < FIB_NRED; ++k) {
    rs.op[k] = k < nv ? op[k] : 0;
    rs.dst[k] = k < nv ? dst[k] : -1;
  }
  hipLaunchKernelGGL(k_finalize, dim3(d.ninst), dim3(FIN_T), 0, s, d.I, d.partials, d.pstride,
                     rs, d.ctl);
}

static void finalize_thermo(const DevBufs &d, hipStream_t s) {
  const int op[5] = {0, 0, 0, 0, 1};
  const int dst[5] = {DCTL_IDX(glob_sum), DCTL_IDX(side_sum), DCTL_IDX(side_cnt),
                      DCTL_IDX(ke_sum), DCTL_IDX(vmax)};
  finalize(d, 5, FIB_BLOCK, op, dst, s);
}

void launch_bin(const DevBufs &d, hipStream_t s) {
  (void)hipMemsetAsync(d.bcnt, 0, sizeof(int) * (size_t)d.nbucket, s);
  hipLaunchKernelGGL(k_bin, dim3(nblk(d.NA, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, d.X[d.cur], d.NA,
                     d.I, d.bkey, d.brank, d.bcnt, d.ctl);
}

void launch_scan(const DevBufs &d, hipStream_t s) {
  const int nb = nblk(d.nbucket, SCAN_CH);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_T), 0, s, d.bcnt, d.nbucket, d.bsum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_T), 0, s, d.bsum, nb);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_T), 0, s, d.bcnt, d.nbucket, d.bsum, d.I,
                     d.ninst, d.cstart);
}

void launch_sort(const DevBufs &d, hipStream_t s) {
  hipLaunchKernelGGL(k_scatter, dim3(nblk(d.NA, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, d.bkey,
                     d.brank, d.NA, d.cstart, d.perm);
  hipLaunchKernelGGL(k_fixup, dim3(nblk(d.nbucket, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, d.cstart,
                     d.bcnt, d.nbucket, d.perm);
  const int c = d.cur, n = 1 - d.cur;
  hipLaunchKernelGGL(k_reorder, dim3(nblk(d.NA, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, d.NA, d.perm,
                     d.I, d.X[c], d.X[n], d.V[c], d.V[n], d.F[c], d.F[n], d.Xi[c], d.Xi[n],
                     d.inv);
}

void launch_nbuild(const DevBufs &d, const FConst &fc, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_counts, dim3(nblk(d.NA, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, d.NA,
                     d.cnt_sw, d.cnt_ion);
  if (d.total_cells > 0)
    hipLaunchKernelGGL(k_nbuild, dim3(d.total_cells), dim3(FIB_WAVE), 0, s, d.X[d.cur], d.I,
                       d.ninst, d.total_cells, d.cstart, d.bcnt, fc.rv2, fc.rv2lj, d.NA, d.Wsw,
                       d.Wion, d.cnt_sw, d.nbr_sw, d.cnt_ion, d.nbr_ion, d.ctl);
  hipLaunchKernelGGL(k_rev, dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s, d.NA, d.Wsw,
                     d.cnt_sw, d.cnt_ion, d.nbr_sw, d.rev, d.pstride, d.partials);
  const int op[2] = {1, 1};
  const int dst[2] = {DCTL_IDX(maxcnt_sw), DCTL_IDX(maxcnt_ion)};
  finalize(d, 2, FIB_BLOCK, op, dst, s);
}

void launch_cellid(const DevBufs &d, int inst, int32_t *dout, hipStream_t s) {
  hipLaunchKernelGGL(k_cellid, dim3(nblk(d.NA, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, d.X[d.cur],
                     d.NA, d.I, inst, dout);
}

template <bool MIXED, bool EN, bool VIR>
static void force_a(const DevBufs &d, const FConst &fc, hipStream_t s) {
  typedef typename Prec<MIXED>::T4 T4;
  hipLaunchKernelGGL((k_force_a<MIXED, EN, VIR>), dim3(d.NA / FIB_WAVE), dim3(FIB_WAVE), 0, s,
                     d.X[d.cur], d.I, fc, d.tab, d.NA, d.Wsw, d.Wion, d.cnt_sw, d.nbr_sw,
                     d.cnt_ion, d.nbr_ion, (T4 *)d.G, d.amask, d.Fpart, d.upart, d.pstride,
                     d.partials, d.ctl);
  if (EN || VIR) {
    int op[10], dst[10];
    for (int k = 0; k < 10; ++k) { op[k] = 0; dst[k] = DCTL_IDX(utot) + k; }
    finalize(d, 10, FIB_WAVE, op, dst, s);
  }
}

template <bool MIXED, bool EN, int MODE>
static void force_b(const DevBufs &d, const FConst &fc, const StepArgs &sa, double *Fo,
                    hipStream_t s) {
  typedef typename Prec<MIXED>::T4 T4;
  const double fscale = MIXED ? fc.eps[0][0] * fc.isig[0][0] : 1.0;
  const double escale = MIXED ? fc.eps[0][0] : 1.0;
  hipLaunchKernelGGL((k_force_b<MIXED, EN, MODE>), dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s,
                     d.X[d.cur], d.I, d.NA, d.Wsw, d.nbr_sw, d.rev, (const T4 *)d.G, d.amask,
                     d.Fpart, d.upart, fscale, escale, Fo, d.uout, d.V[d.cur], sa, d.pstride,
                     d.partials, d.ctl);
  if (MODE == 1) finalize_thermo(d, s);
}

void launch_force_a(const DevBufs &d, const FConst &fc, bool mixed, bool energy, bool virial,
                    hipStream_t s) {
  if (mixed) {
    if (energy) { if (virial) force_a<true, true, true>(d, fc, s); else force_a<true, true, false>(d, fc, s); }
    else { if (virial) force_a<true, false, true>(d, fc, s); else force_a<true, false, false>(d, fc, s); }
  } else {
    if (energy) { if (virial) force_a<false, true, true>(d, fc, s); else force_a<false, true, false>(d, fc, s); }
    else { if (virial) force_a<false, false, true>(d, fc, s); else force_a<false, false, false>(d, fc, s); }
  }
}

void launch_force_b(const DevBufs &d, const FConst &fc, bool mixed, bool energy, int mode,
                    const StepArgs &sa, double *Fout, hipStream_t s) {
  if (mixed) {
    if (mode == 0) { if (energy) force_b<true, true, 0>(d, fc, sa, Fout, s); else force_b<true, false, 0>(d, fc, sa, Fout, s); }
    else force_b<true, false, 1>(d, fc, sa, Fout, s);
  } else {
    if (mode == 0) { if (energy) force_b<false, true, 0>(d, fc, sa, Fout, s); else force_b<false, false, 0>(d, fc, sa, Fout, s); }
    else force_b<false, false, 1>(d, fc, sa, Fout, s);
  }
}

void launch_pos(const DevBufs &d, const StepArgs &sa, hipStream_t s) {
  hipLaunchKernelGGL(k_pos, dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s, d.X[d.cur], d.I, d.NA,
                     d.V[d.cur], d.F[d.cur], d.Xi[d.cur], sa, d.pstride, d.partials, d.ctl);
  const int op[2] = {0, 1};
  const int dst[2] = {DCTL_IDX(eq_sum), DCTL_IDX(kemax)};
  finalize(d, 2, FIB_BLOCK, op, dst, s);
}

void launch_sums(const DevBufs &d, const StepArgs &sa, hipStream_t s) {
  hipLaunchKernelGGL(k_sums, dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s, d.X[d.cur], d.I, d.NA,
                     d.V[d.cur], sa, d.pstride, d.partials, d.ctl);
  finalize_thermo(d, s);
}

void launch_import(const DevBufs &d, int inst, const DInst &in, const double *X,
                   const double *V, const int32_t *atype, const int32_t *xsides,
                   const double *Xi, hipStream_t s) {
  hipLaunchKernelGGL(k_import, dim3(nblk(in.natm_pad, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, inst, in,
                     d.NA, X, V, atype, xsides, Xi, d.X[0], d.X[1], d.V[0], d.V[1], d.F[0],
                     d.F[1], d.Xi[0], d.Xi[1], d.inv);
}

void launch_export_x(const DevBufs &d, int inst, const DInst &in, double *out, hipStream_t s) {
  (void)inst;
  hipLaunchKernelGGL(k_export_x, dim3(nblk(in.natm_pad, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, in,
                     d.X[d.cur], out);
}

void launch_export_vec(const DevBufs &d, int inst, const DInst &in, const double *src,
                       double *out, hipStream_t s) {
  (void)inst;
  hipLaunchKernelGGL(k_export_vec, dim3(nblk(in.natm_pad, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, in,
                     d.NA, d.X[d.cur], src, out);
}

void launch_export_scalar(const DevBufs &d, int inst, const DInst &in, const double *src,
                          double *out, hipStream_t s) {
  (void)inst;
  hipLaunchKernelGGL(k_export_scalar, dim3(nblk(in.natm_pad, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s,
                     in, d.X[d.cur], src, out);
}

void launch_set_atom(const DevBufs &d, int p, const double x[3], const double v[3],
                     hipStream_t s) {
  hipLaunchKernelGGL(k_set_atom, dim3(1), dim3(FIB_WAVE), 0, s, d.NA, d.X[d.cur], d.V[d.cur], p,
                     x[0], x[1], x[2], v[0], v[1], v[2]);
}

}  // namespace fibk
