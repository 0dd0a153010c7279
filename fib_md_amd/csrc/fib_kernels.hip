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

#define FIN_T 1024
// One block per instance: combine the instance's partials in a fixed order
// (thread t sums blocks t, t+FIN_T, ... then a fixed tree), all NV slots at
// once so the loads of every slot are in flight together.
template <int NV>
__global__ __launch_bounds__(FIN_T) void k_finalize(const DInst *__restrict__ I,
                                                    const double *__restrict__ partials,
                                                    int pstride, RedSpec rs,
                                                    DCtl *__restrict__ ctl) {
  __shared__ double red[FIN_T / FIB_WAVE][NV];
  const int inst = blockIdx.x;
  const int blk0 = I[inst].atom_off / rs.bs, nblk = I[inst].natm_pad / rs.bs;
  const int lane = threadIdx.x % FIB_WAVE, wid = threadIdx.x / FIB_WAVE;
  double x[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) x[k] = rs.op[k] ? -1.0e300 : 0.0;
  for (int b = threadIdx.x; b < nblk; b += FIN_T) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double y = partials[(size_t)k * pstride + blk0 + b];
      x[k] = rs.op[k] ? fmax(x[k], y) : x[k] + y;
    }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    for (int o = FIB_WAVE / 2; o > 0; o >>= 1) {
      const double y = __shfl_xor(x[k], o);
      x[k] = rs.op[k] ? fmax(x[k], y) : x[k] + y;
    }
    if (lane == 0) red[wid][k] = x[k];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int k = threadIdx.x;
    double t = red[0][k];
    for (int w = 1; w < FIN_T / FIB_WAVE; ++w) t = rs.op[k] ? fmax(t, red[w][k]) : t + red[w][k];
    if (rs.dst[k] >= 0) reinterpret_cast<double *>(&ctl[inst])[rs.dst[k]] = t;
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
#define NB_CHUNK 128

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

// Reverse slots: rev[k][i] = slot of i in the list of its k-th neighbour.
// A block handles 16 atoms with one thread per (slot, atom): the 16 atoms of
// one slot row are contiguous, and each thread scans its neighbour's list
// eight entries at a time so the loads overlap.
#define REV_ATOMS 16
__global__ __launch_bounds__(FIB_BLOCK) void k_rev(int NA, int Wsw, const int *__restrict__ cnt_sw,
                                                   const int *__restrict__ nbr_sw,
                                                   uint8_t *__restrict__ rev) {
  const int i = blockIdx.x * REV_ATOMS + (threadIdx.x % REV_ATOMS);
  const int c = min(cnt_sw[i], Wsw);
  for (int k = threadIdx.x / REV_ATOMS; k < c; k += FIB_BLOCK / REV_ATOMS) {
    const int j = nbr_sw[(size_t)k * NA + i];
    const int cj = min(cnt_sw[j], Wsw);
    int r = 255;
    for (int q0 = 0; q0 < cj && r == 255; q0 += 8) {
      int e[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) e[u] = (q0 + u < cj) ? nbr_sw[(size_t)(q0 + u) * NA + j] : -1;
#pragma unroll
      for (int u = 7; u >= 0; --u)
        if (e[u] == i) r = q0 + u;
    }
    rev[(size_t)k * NA + i] = (uint8_t)r;
  }
}

// Longest SW / ion lists of the build (overflow check).
__global__ __launch_bounds__(FIB_BLOCK) void k_maxcnt(const int *__restrict__ cnt_sw,
                                                      const int *__restrict__ cnt_ion,
                                                      int pstride, double *__restrict__ partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {(double)cnt_sw[i], (double)cnt_ion[i]};
  const int op[2] = {1, 1};
  block_partials<2, FIB_BLOCK>(v, op, blockIdx.x, pstride, partials);
}

// ---------------------------------------------------------------------------
// force phase A: four lanes per atom
//
// A group of four lanes owns one atom.  Phase 1: the lanes split the atom's
// buffered SW list (slot k goes to lane k % 4), compute the reference pair
// force and flag the legs inside the SW range; the flags are OR-ed over the
// quad into the atom's active mask.  Phase 2: active leg L (in slot order) is
// owned by lane L % 4, register slot L / 4; the lane computes its geometry
// (unit vector, 1/r, E = exp(gamma/(r-a)), d ln E/dr).  Phase 3: every lane
// loops over the legs b of its atom, fetches them with wavefront shuffles and
// accumulates the three-body terms centred on the atom that act on its own
// legs a; the share acting on neighbour a goes to G[slot(a)], the reaction on
// the centre stays in registers.  Phase 4: quad reduction, lane 0 stores.
// No LDS, no atomics; legs beyond 8 per atom are recomputed on the fly.
template <bool MIXED>
struct Prec;
template <>
struct Prec<false> {
  typedef double T;
  typedef double4 T4;
};
template <>
struct Prec<true> {
  typedef float T;
  typedef float4 T4;
};

template <typename T>
struct Leg {
  T e0, e1, e2;  // unit vector from neighbour to centre
  T rinv;        // 1/r (FP64: m^-1; mixed: sigma^-1)
  T E;           // exp(gamma/(r/sigma - a))
  T g;           // d ln E / dr
  int t;         // species of the neighbour
};

// One ordered three-body term: centre c, legs a and b.  Returns the force on
// a (F_a = -dh/dx_a) and h.  Derivation: h = leps E(ra) E(rb) (cos+1/3)^2,
// d cos/dx_a = -(e_b - cos e_a)/r_a, d r_a/dx_a = -e_a.
template <typename T>
__device__ __forceinline__ void vterm(T leps, const Leg<T> &a, const Leg<T> &b, T *Fa, T &hh) {
  const T cs = a.e0 * b.e0 + a.e1 * b.e1 + a.e2 * b.e2;
  const T D = cs + T(1.0 / 3.0);
  const T Eab = leps * a.E * b.E;
  hh = Eab * D * D;
  const T c1 = hh * a.g;
  const T c2 = T(2) * Eab * D * a.rinv;
  Fa[0] = c1 * a.e0 + c2 * (b.e0 - cs * a.e0);
  Fa[1] = c1 * a.e1 + c2 * (b.e1 - cs * a.e1);
  Fa[2] = c1 * a.e2 + c2 * (b.e2 - cs * a.e2);
}

template <bool MIXED>
__device__ __forceinline__ void leg_geom(const double4 &xi, const double4 &xj, const DInst &in,
                                         const FConst &fc, int ti,
                                         Leg<typename Prec<MIXED>::T> &L) {
  const int tj = type_of(meta_of(xj));
  double d0, d1, d2;
  min_image(xi, xj, in, d0, d1, d2);
  const double isg = fc.isig[(ti == 2) ? 1 : 0][(tj == 2) ? 1 : 0];
  L.t = tj;
  if (!MIXED) {
    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double ri = 1.0 / r;
    const double rs = r * isg;
    const double inv_ra = 1.0 / (rs - fc.al);
    L.e0 = d0 * ri; L.e1 = d1 * ri; L.e2 = d2 * ri;
    L.rinv = ri;
    L.E = exp(fc.gamma * inv_ra);
    L.g = -fc.gamma * inv_ra * inv_ra * isg;
  } else {
    const float r0 = (float)(d0 * isg), r1 = (float)(d1 * isg), r2 = (float)(d2 * isg);
    const float rs = sqrtf(r0 * r0 + r1 * r1 + r2 * r2);
    const float ri = 1.0f / rs;
    const float inv_ra = 1.0f / (rs - (float)fc.al);
    L.e0 = r0 * ri; L.e1 = r1 * ri; L.e2 = r2 * ri;
    L.rinv = ri;
    L.E = __expf((float)fc.gamma * inv_ra);
    L.g = -(float)fc.gamma * inv_ra * inv_ra;
  }
}

// Slot of the n-th set bit of m (n is small).
__device__ __forceinline__ int nth_bit(uint64_t m, int n) {
  for (int r = 0; r < n; ++r) m &= m - 1;
  return __ffsll((unsigned long long)m) - 1;
}

template <typename T>
__device__ __forceinline__ T shfl_any(T v, int src) {
  return __shfl(v, src);
}

#define FA_ATOMS (FIB_BLOCK / 4)  // atoms per block

// Minimum image with v_rndne (ties to even).  Neighbour pairs are far from
// the half-box tie, so this equals the reference's NINT image there.
__device__ __forceinline__ void min_image_fast(const double4 &a, const double4 &b,
                                               const DInst &in, double &x0, double &x1,
                                               double &x2) {
  const double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
  x0 = d0 - rint(d0 * in.ilb[0]) * in.lb[0];
  x1 = d1 - rint(d1 * in.ilb[1]) * in.lb[1];
  x2 = d2 - rint(d2 * in.ilb[2]) * in.lb[2];
}

template <bool MIXED, bool EN, bool VIR>
__global__ __launch_bounds__(FIB_BLOCK) void k_force_a(
    const double4 *__restrict__ X, const DInst *__restrict__ I, const FConst fc,
    const TabEntry *__restrict__ tab, int NA, int Wsw, int Wion, const int *__restrict__ cnt_sw,
    const int *__restrict__ nbr_sw, const int *__restrict__ cnt_ion,
    const int *__restrict__ nbr_ion, typename Prec<MIXED>::T4 *__restrict__ G,
    uint64_t *__restrict__ amask, double *__restrict__ Fpart, double *__restrict__ upart,
    int pstride, double *__restrict__ partials, DCtl *__restrict__ ctl) {
  typedef typename Prec<MIXED>::T T;
  typedef typename Prec<MIXED>::T4 T4;

  const int nb = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nb);
  const int lane = threadIdx.x % FIB_WAVE;
  const int q = lane & 3;
  const int gbase = lane & ~3;
  const int i = lb * FA_ATOMS + (int)(threadIdx.x >> 2);
  const double4 xi = X[i];
  const uint32_t mi = meta_of(xi);
  // every block lies inside one instance: scalar loads of the instance record
  const int kin = __builtin_amdgcn_readfirstlane(inst_of(mi));
  const DInst &in = I[kin];
  const bool live = !(mi & (META_PAD | META_DEAD));
  const int ti = type_of(mi);
  const bool sw_atom = live && ti < 3;
  const int tr = (ti == 2) ? 1 : 0;
  const double isgA = fc.isig[tr][0], isgB = fc.isig[tr][1];
  const double esgA = fc.epssig[tr][0], esgB = fc.epssig[tr][1];
  const double epsA = fc.eps[tr][0], epsB = fc.eps[tr][1];
  // leps in the unit of the three-body pass (J, or eps_Si in mixed)
  const double lu = MIXED ? 1.0 / fc.eps[0][0] : 1.0;
  const T lp11 = (T)(fc.leps[tr][0][0] * lu), lp12 = (T)(fc.leps[tr][0][1] * lu),
          lp22 = (T)(fc.leps[tr][1][1] * lu);
  // unit factors: FP64 works in m and N; mixed in sigma and eps/sigma
  const double fscale = MIXED ? fc.eps[0][0] * fc.isig[0][0] : 1.0;
  const double escale = MIXED ? fc.eps[0][0] : 1.0;
  const double lscale = MIXED ? 1.0 / fc.isig[0][0] : 1.0;

  // force accumulators: FP64, or FP32 value + compensation (mixed)
  double fx = 0, fy = 0, fz = 0;
  float sx = 0, sy = 0, sz = 0, cx = 0, cy = 0, cz = 0;
  double ui = 0, etot = 0;
  T W[9];
#pragma unroll
  for (int w = 0; w < 9; ++w) W[w] = 0;
  auto addW = [&](T d0, T d1, T d2, T f0, T f1, T f2) {
    W[0] += d0 * f0; W[1] += d0 * f1; W[2] += d0 * f2;
    W[3] += d1 * f0; W[4] += d1 * f1; W[5] += d1 * f2;
    W[6] += d2 * f0; W[7] += d2 * f1; W[8] += d2 * f2;
  };

  // ---- phase 1: cheap distance test of every buffered SW slot -------------
  // lane q tests slots q, q+4, q+8, ...; pm collects the pairs that carry a
  // pair force (FP64: table range; mixed: SW range)
  uint64_t pm = 0;
  int cs = 0;
  if (sw_atom) {
    int jr[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) jr[u] = nbr_sw[(size_t)(q + 4 * u) * NA + i];
    cs = min(cnt_sw[i], Wsw);
    const double gate = MIXED ? fc.cut_sw2 : fc.r2gate;
    for (int k0 = q; k0 < cs; k0 += 16) {
      double4 xq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * u;
        int j = (k0 == q) ? jr[u] : ((k < cs) ? nbr_sw[(size_t)k * NA + i] : i);
        j = (k < cs) ? j : i;
        xq[u] = X[j];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * u;
        double d0, d1, d2;
        min_image_fast(xi, xq[u], in, d0, d1, d2);
        const int tj = type_of(meta_of(xq[u]));
        const double isg = (tj != 2) ? isgA : isgB;
        const double r2s = (d0 * d0 + d1 * d1 + d2 * d2) * (isg * isg);
        if (k < cs && r2s < gate) pm |= 1ull << k;
      }
    }
  }
  {
    uint32_t lo = (uint32_t)pm, hi = (uint32_t)(pm >> 32);
    lo |= __shfl_xor(lo, 1); hi |= __shfl_xor(hi, 1);
    lo |= __shfl_xor(lo, 2); hi |= __shfl_xor(hi, 2);
    pm = ((uint64_t)hi << 32) | lo;
  }
  const int np = __popcll(pm);

  // ---- phase 2: pair force and leg geometry, entry n of pm on lane n % 4 ----
  // register slot n / 4 (0 or 1); entries n >= 8 are recomputed in phase 3
  Leg<T> L0, L1;
  L0.e0 = L0.e1 = L0.e2 = L0.rinv = L0.E = L0.g = T(0); L0.t = 1;
  L1 = L0;
  uint64_t am = 0;  // legs inside the SW range found by this lane
  auto pair_leg = [&](int k, Leg<T> &L) -> bool {
    const int j = nbr_sw[(size_t)k * NA + i];
    const double4 xj = X[j];
    const int tj = type_of(meta_of(xj));
    const bool jA = tj != 2;
    const double isg = jA ? isgA : isgB;
    double d0, d1, d2;
    min_image_fast(xi, xj, in, d0, d1, d2);
    L.t = tj;
    L.E = T(0); L.g = T(0);
    bool leg = false;
    if (!MIXED) {
      const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      const double rinv = 1.0 / r;
      const double rs = r * isg;
      // reference pair force: 3-point Lagrange on the derivative table
      const int t1 = (int)(rs / fc.umax * fc.tabn1) + 1;
      if (t1 < fc.izero) {
        const TabEntry te = tab[t1];
        const double e1 = rs - te.u0, e2 = rs - te.u1, e3 = rs - te.u2;
        const double df = 0.5 * e2 * e3 * te.t0 - e1 * e3 * te.t1 + 0.5 * e1 * e2 * te.t2;
        const double fs = -df * (jA ? esgA : esgB) * rinv;
        fx += fs * d0; fy += fs * d1; fz += fs * d2;
        if (VIR) addW(d0, d1, d2, 0.5 * fs * d0, 0.5 * fs * d1, 0.5 * fs * d2);
      }
      L.e0 = d0 * rinv; L.e1 = d1 * rinv; L.e2 = d2 * rinv;
      L.rinv = rinv;
      if (rs < fc.cut_sw) {
        const double inv_ra = 1.0 / (rs - fc.al);
        L.E = exp(fc.gamma * inv_ra);
        L.g = -fc.gamma * inv_ra * inv_ra * isg;
        if (EN) {
          const double rs2 = rs * rs;
          const double ul = 0.5 * (jA ? epsA : epsB) * fc.A * (fc.B / (rs2 * rs2) - 1.0) *
                            exp(inv_ra);
          ui += ul;
          etot += ul;
        }
        leg = true;
      }
    } else {
      // FP64 difference in sigma units, everything after in FP32
      const float r0 = (float)(d0 * isg), r1 = (float)(d1 * isg), r2 = (float)(d2 * isg);
      const float rr = r0 * r0 + r1 * r1 + r2 * r2;
      const float rinv = __builtin_amdgcn_rsqf(rr);
      const float rs = rr * rinv;
      L.e0 = r0 * rinv; L.e1 = r1 * rinv; L.e2 = r2 * rinv;
      L.rinv = rinv;
      if (rs < (float)fc.cut_sw) {
        const float inv_ra = __builtin_amdgcn_rcpf(rs - (float)fc.al);
        const float ex = __expf(inv_ra);
        const float B = (float)fc.B, A = (float)fc.A;
        const float ir2 = rinv * rinv;
        const float ir4 = ir2 * ir2;
        // phi2(r) = A (B r^-4 - 1) e^{1/(r-a)};  force on i = -dphi/dr e
        const float dphi = A * ex * (-4.0f * B * ir4 * rinv - (B * ir4 - 1.0f) * inv_ra * inv_ra);
        const float fs = -dphi * rinv * (float)(jA ? esgA : esgB) * (float)(1.0 / fscale);
        // fs in eps/sigma units of the Si pair: scaled by fscale at the end
        const float f0 = fs * r0, f1 = fs * r1, f2 = fs * r2;
        kadd(sx, cx, f0); kadd(sy, cy, f1); kadd(sz, cz, f2);
        if (VIR) addW(r0, r1, r2, 0.5f * f0, 0.5f * f1, 0.5f * f2);
        L.E = __expf((float)fc.gamma * inv_ra);
        L.g = -(float)fc.gamma * inv_ra * inv_ra;
        if (EN) {
          const double ul = 0.5 * (jA ? epsA : epsB) * (double)(A * (B * ir4 - 1.0f) * ex);
          ui += ul;
          etot += ul;
        }
        leg = true;
      }
    }
    return leg;
  };
  if (q < np) {
    const int k = nth_bit(pm, q);
    if (pair_leg(k, L0)) am |= 1ull << k;
  }
  if (q + 4 < np) {
    const int k = nth_bit(pm, q + 4);
    if (pair_leg(k, L1)) am |= 1ull << k;
  }
  for (int n = q + 8; n < np; n += 4) {  // dense atoms: pair force only here
    Leg<T> Lx;
    const int k = nth_bit(pm, n);
    if (pair_leg(k, Lx)) am |= 1ull << k;
  }
  {
    uint32_t lo = (uint32_t)am, hi = (uint32_t)(am >> 32);
    lo |= __shfl_xor(lo, 1); hi |= __shfl_xor(hi, 1);
    lo |= __shfl_xor(lo, 2); hi |= __shfl_xor(hi, 2);
    am = ((uint64_t)hi << 32) | lo;
  }

  // ---- Moliere pairs (si.f90:426-446): both atoms must be fired ----------
  if (live) {
    const int ci = min(cnt_ion[i], Wion);
    const bool ifired = !(ti == 3 && orig_of(xi) >= in.nsg + in.impact);
    for (int k = q; ifired && k < ci; k += 4) {
      const int j = nbr_ion[(size_t)k * NA + i];
      const double4 xj = X[j];
      const uint32_t mj = meta_of(xj);
      const int tj = type_of(mj);
      if (tj == 3 && orig_of(xj) >= in.nsg + in.impact) continue;
      if (mj & META_DEAD) continue;
      double d0, d1, d2;
      min_image_fast(xi, xj, in, d0, d1, d2);
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
        if (MIXED) {
          kadd(sx, cx, (float)(f0 / fscale)); kadd(sy, cy, (float)(f1 / fscale));
          kadd(sz, cz, (float)(f2 / fscale));
          if (VIR) addW((T)(d0 / lscale), (T)(d1 / lscale), (T)(d2 / lscale),
                        (T)(0.5 * f0 / fscale), (T)(0.5 * f1 / fscale), (T)(0.5 * f2 / fscale));
        } else {
          fx += f0; fy += f1; fz += f2;
          if (VIR) addW(d0, d1, d2, 0.5 * f0, 0.5 * f1, 0.5 * f2);
        }
        if (EN) { const double ul = 0.5 * zz / r * sa; ui += ul; etot += ul; }
      } else {
        // FP32 with lengths in units of aa; F = (zz/aa^2) (1/x)(sa/x + sb) d/(aa x)
        const float x = (float)(r / aa);
        const float e1 = __expf(-(float)fc.molb[0] * x), e2 = __expf(-(float)fc.molb[1] * x),
                    e3 = __expf(-(float)fc.molb[2] * x);
        const float c0 = (float)fc.molc[0], c1 = (float)fc.molc[1], c2 = (float)fc.molc[2];
        const float sa = c0 * e1 + c1 * e2 + c2 * e3;
        const float sb = c0 * (float)fc.molb[0] * e1 + c1 * (float)fc.molb[1] * e2 +
                         c2 * (float)fc.molb[2] * e3;
        const float ix = __builtin_amdgcn_rcpf(x);
        const float fr = ix * (sa * ix + sb) * ix;
        const float scl = (float)(zz / (aa * aa) / fscale);
        const float f0 = fr * (float)(d0 / aa) * scl, f1 = fr * (float)(d1 / aa) * scl,
                    f2 = fr * (float)(d2 / aa) * scl;
        kadd(sx, cx, f0); kadd(sy, cy, f1); kadd(sz, cz, f2);
        if (VIR) addW((T)(d0 / lscale), (T)(d1 / lscale), (T)(d2 / lscale), 0.5f * f0,
                      0.5f * f1, 0.5f * f2);
        if (EN) { const double ul = 0.5 * zz / aa * (double)(sa * ix); ui += ul; etot += ul; }
      }
    }
  }

  // ---- phase 3: three-body terms centred on i ------------------------------
  // pm entry n lives on lane n % 4, slot n / 4; entries that are not legs
  // carry E = 0 and contribute exactly zero
  int npmax = np;
  for (int o = 32; o >= 4; o >>= 1) npmax = max(npmax, __shfl_xor(npmax, o));
  const int na = __popcll(am);
  if (npmax > 0 && __any(na >= 2)) {
    T G0[4] = {0, 0, 0, 0}, G1[4] = {0, 0, 0, 0};  // force on own legs + energy share
    T fcx = 0, fcy = 0, fcz = 0;                   // reaction on the centre
    T uc = 0;
    const int nreg = min(npmax, 8);
    for (int bi = 0; bi < nreg; ++bi) {
      const int src = gbase | (bi & 3);
      const bool s0 = bi < 4;
      Leg<T> Lb;
      Lb.e0 = __shfl(s0 ? L0.e0 : L1.e0, src);
      Lb.e1 = __shfl(s0 ? L0.e1 : L1.e1, src);
      Lb.e2 = __shfl(s0 ? L0.e2 : L1.e2, src);
      Lb.E = __shfl(s0 ? L0.E : L1.E, src);
      Lb.t = __shfl(s0 ? L0.t : L1.t, src);
      if (bi >= np || Lb.E == T(0)) continue;
      // own entries a = q (slot 0) and q + 4 (slot 1)
#pragma unroll
      for (int sa = 0; sa < 2; ++sa) {
        const int a = q + 4 * sa;
        const Leg<T> &La = sa == 0 ? L0 : L1;
        if (a >= np || a == bi || La.E == T(0)) continue;
        const T lep = (La.t == 1 && Lb.t == 1) ? lp11 : ((La.t == 2 && Lb.t == 2) ? lp22 : lp12);
        T Fa[3], hh;
        vterm<T>(lep, La, Lb, Fa, hh);
        if (sa == 0) {
          G0[0] += Fa[0]; G0[1] += Fa[1]; G0[2] += Fa[2];
          if (EN) G0[3] += hh * T(1.0 / 3.0);
        } else {
          G1[0] += Fa[0]; G1[1] += Fa[1]; G1[2] += Fa[2];
          if (EN) G1[3] += hh * T(1.0 / 3.0);
        }
        fcx -= Fa[0]; fcy -= Fa[1]; fcz -= Fa[2];
        if (EN) uc += hh;
      }
    }
    // entries beyond the register slots (dense atoms): recompute on the fly
    auto leg_at = [&](int n, Leg<T> &L) {
      const int k = nth_bit(pm, n);
      const int j = nbr_sw[(size_t)k * NA + i];
      leg_geom<MIXED>(xi, X[j], in, fc, ti, L);
      if (!((am >> k) & 1ull)) L.E = T(0);
    };
    if (npmax > 8) {
      // terms (a < 8, b >= 8)
      for (int bi = 8; bi < np; ++bi) {
        Leg<T> Lb;
        leg_at(bi, Lb);
        if (Lb.E == T(0)) continue;
#pragma unroll
        for (int sa = 0; sa < 2; ++sa) {
          const int a = q + 4 * sa;
          const Leg<T> &La = sa == 0 ? L0 : L1;
          if (a >= np || La.E == T(0)) continue;
          const T lep = (La.t == 1 && Lb.t == 1) ? lp11 : ((La.t == 2 && Lb.t == 2) ? lp22 : lp12);
          T Fa[3], hh;
          vterm<T>(lep, La, Lb, Fa, hh);
          if (sa == 0) {
            G0[0] += Fa[0]; G0[1] += Fa[1]; G0[2] += Fa[2];
            if (EN) G0[3] += hh * T(1.0 / 3.0);
          } else {
            G1[0] += Fa[0]; G1[1] += Fa[1]; G1[2] += Fa[2];
            if (EN) G1[3] += hh * T(1.0 / 3.0);
          }
          fcx -= Fa[0]; fcy -= Fa[1]; fcz -= Fa[2];
          if (EN) uc += hh;
        }
      }
      // own entries a >= 8: full sum over all b
      for (int a = q + 8; a < np; a += 4) {
        Leg<T> La;
        leg_at(a, La);
        T Ga[4] = {0, 0, 0, 0};
        // (a leg whose E underflowed still stores its zero share below)
        for (int bi = 0; bi < np && La.E != T(0); ++bi) {
          if (bi == a) continue;
          Leg<T> Lb;
          leg_at(bi, Lb);
          if (Lb.E == T(0)) continue;
          const T lep = (La.t == 1 && Lb.t == 1) ? lp11 : ((La.t == 2 && Lb.t == 2) ? lp22 : lp12);
          T Fa[3], hh;
          vterm<T>(lep, La, Lb, Fa, hh);
          Ga[0] += Fa[0]; Ga[1] += Fa[1]; Ga[2] += Fa[2];
          if (EN) Ga[3] += hh * T(1.0 / 3.0);
          fcx -= Fa[0]; fcy -= Fa[1]; fcz -= Fa[2];
          if (EN) uc += hh;
        }
        const int ka = nth_bit(pm, a);
        T4 gv;
        gv.x = Ga[0]; gv.y = Ga[1]; gv.z = Ga[2]; gv.w = Ga[3];
        G[(size_t)ka * NA + i] = gv;
        if (VIR) {
          const T rl = -T(1) / La.rinv;
          addW(rl * La.e0, rl * La.e1, rl * La.e2, Ga[0], Ga[1], Ga[2]);
        }
      }
    }
    // store the shares of the register-held legs
    auto store_leg = [&](int n, const Leg<T> &La, const T (&Gs)[4]) {
      const int ka = nth_bit(pm, n);
      T4 gv;
      gv.x = Gs[0]; gv.y = Gs[1]; gv.z = Gs[2]; gv.w = Gs[3];
      G[(size_t)ka * NA + i] = gv;
      if (VIR && La.E != T(0)) {
        const T rl = -T(1) / La.rinv;
        addW(rl * La.e0, rl * La.e1, rl * La.e2, Gs[0], Gs[1], Gs[2]);
      }
    };
    if (q < np) store_leg(q, L0, G0);
    if (q + 4 < np) store_leg(q + 4, L1, G1);
    if (MIXED) {
      kadd(sx, cx, (float)fcx); kadd(sy, cy, (float)fcy); kadd(sz, cz, (float)fcz);
    } else {
      fx += fcx; fy += fcy; fz += fcz;
    }
    if (EN) { ui += (double)uc * (1.0 / 6.0) * escale; etot += (double)uc * 0.5 * escale; }
  } else {
    // no three-body terms in this wave: every leg still owes its neighbour a
    // (zero) G entry -- including legs whose E underflowed to zero
    T4 gv;
    gv.x = gv.y = gv.z = gv.w = T(0);
    for (int n = q; n < np; n += 4) {
      const int k = nth_bit(pm, n);
      if ((am >> k) & 1ull) G[(size_t)k * NA + i] = gv;
    }
  }

  // ---- phase 4: quad reduction, lane 0 of the group stores ----------------
  if (MIXED) {
    fx = ((double)sx + (double)cx) * fscale;
    fy = ((double)sy + (double)cy) * fscale;
    fz = ((double)sz + (double)cz) * fscale;
  }
  fx += __shfl_xor(fx, 1); fy += __shfl_xor(fy, 1); fz += __shfl_xor(fz, 1);
  fx += __shfl_xor(fx, 2); fy += __shfl_xor(fy, 2); fz += __shfl_xor(fz, 2);
  if (EN) { ui += __shfl_xor(ui, 1); ui += __shfl_xor(ui, 2); }
  if (q == 0) {
    Fpart[i] = fx;
    Fpart[(size_t)NA + i] = fy;
    Fpart[2 * (size_t)NA + i] = fz;
    amask[i] = am;
    if (EN) upart[i] = ui;
  }

  if (EN || VIR) {
    double v[10];
    v[0] = etot;
    const double ws = MIXED ? fscale * lscale : 1.0;
#pragma unroll
    for (int w = 0; w < 9; ++w) v[1 + w] = (double)W[w] * ws;
    const int op[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    block_partials<10, FIB_BLOCK>(v, op, lb, pstride, partials);
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
  // every block lies inside one instance: make the index wave-uniform so the
  // instance record is read with scalar loads
  const int kin = __builtin_amdgcn_readfirstlane(inst_of(mi));
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
  // every block lies inside one instance: make the index wave-uniform so the
  // instance record is read with scalar loads
  const int kin = __builtin_amdgcn_readfirstlane(inst_of(mi));
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
  // every block lies inside one instance: make the index wave-uniform so the
  // instance record is read with scalar loads
  const int kin = __builtin_amdgcn_readfirstlane(inst_of(mi));
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
  switch (nv) {
    case 2:
      hipLaunchKernelGGL(k_finalize<2>, dim3(d.ninst), dim3(FIN_T), 0, s, d.I, d.partials,
                         d.pstride, rs, d.ctl);
      break;
    case 5:
      hipLaunchKernelGGL(k_finalize<5>, dim3(d.ninst), dim3(FIN_T), 0, s, d.I, d.partials,
                         d.pstride, rs, d.ctl);
      break;
    default:
      hipLaunchKernelGGL(k_finalize<10>, dim3(d.ninst), dim3(FIN_T), 0, s, d.I, d.partials,
                         d.pstride, rs, d.ctl);
      break;
  }
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
  hipLaunchKernelGGL(k_maxcnt, dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s, d.cnt_sw,
                     d.cnt_ion, d.pstride, d.partials);
  const int op[2] = {1, 1};
  const int dst[2] = {DCTL_IDX(maxcnt_sw), DCTL_IDX(maxcnt_ion)};
  finalize(d, 2, FIB_BLOCK, op, dst, s);
}

void launch_rev(const DevBufs &d, hipStream_t s) {
  hipLaunchKernelGGL(k_rev, dim3(d.NA / REV_ATOMS), dim3(FIB_BLOCK), 0, s, d.NA, d.Wsw, d.cnt_sw,
                     d.nbr_sw, d.rev);
}

void launch_cellid(const DevBufs &d, int inst, int32_t *dout, hipStream_t s) {
  hipLaunchKernelGGL(k_cellid, dim3(nblk(d.NA, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, d.X[d.cur],
                     d.NA, d.I, inst, dout);
}

template <bool MIXED, bool EN, bool VIR>
static void force_a(const DevBufs &d, const FConst &fc, hipStream_t s) {
  typedef typename Prec<MIXED>::T4 T4;
  hipLaunchKernelGGL((k_force_a<MIXED, EN, VIR>), dim3(d.NA / FA_ATOMS), dim3(FIB_BLOCK), 0, s,
                     d.X[d.cur], d.I, fc, d.tab, d.NA, d.Wsw, d.Wion, d.cnt_sw, d.nbr_sw,
                     d.cnt_ion, d.nbr_ion, (T4 *)d.G, d.amask, d.Fpart, d.upart, d.pstride,
                     d.partials, d.ctl);
  if (EN || VIR) {
    int op[10], dst[10];
    for (int k = 0; k < 10; ++k) { op[k] = 0; dst[k] = DCTL_IDX(utot) + k; }
    finalize(d, 10, FA_ATOMS, op, dst, s);
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
