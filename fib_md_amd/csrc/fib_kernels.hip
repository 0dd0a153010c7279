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
// Per force evaluation (replaces si_force, si.f90:384-552):
//   k_legs      one thread per atom: SW pair and Moliere forces in registers,
//               the atom's three-body legs to the leg arrays
//   k_tri       one lane per leg, legs staged in LDS: the vertex terms
//               centred on the leg's atom, force on the leg's neighbour into
//               G, and where the neighbour keeps its force on the atom (RA)
//   k_extra     vertex terms of the stale build-time triplet graph
//   k_force3    one thread per atom: own terms + the G entries its neighbours
//               wrote for it; the velocity update of either integrator
// Per step (replaces the inline Verlet of onestep, timeint.f90:165-236):
//   k_pos       thermostat scaling, position update, side freeze, rebox and
//               the sputter-layer rules in one pass
//   k_md_pos    the benchmark configurations' integrator (MD-step mode):
//               shells, half kick, drift, pins, removal, displacements
//
// Reductions (energy, virial, thermostat sums, max KE) use per-block
// partials; the last block of an instance sums them in block order, so every
// result is deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "fib_device.h"
#include "fib_kernels.h"

namespace fibk {

// Kernels of the list build take `gate`: null, or gate[0] = "some instance
// rebuilds" and gate[1 + k] = "instance k rebuilds".  With gate[0] zero they
// return at once (block-uniform), so a build can be enqueued -- or captured
// in a graph -- and skipped on the device; atoms, cells and lists of an
// instance whose own flag is zero are left untouched.  k_md_decide writes
// every flag at the start of each MD step, before any kernel reads them.
#define GATE_OR_RETURN(gate) \
  do {                       \
    if ((gate) && !(gate)[0]) return; \
  } while (0)
// MD-step mode: once a list build of a step overflowed its width, `halt` is
// set and every later kernel of the step (and of the steps after it, when
// they were already enqueued or captured) returns at once, so the host can
// grow the lists and finish the step from the state the overflow left.
#define HALT_OR_RETURN(halt) \
  do {                       \
    if ((halt) && *(halt)) return; \
  } while (0)
__device__ __forceinline__ bool gated_out(const int *gate, int inst) {
  return gate && !gate[1 + inst];
}
// Under a gate, the blocks of FIB_BLOCK atoms of instances that are not
// rebuilt are skipped whole, before any per-atom load: every such block lies
// inside one instance (binst), and the grid-stride loops that use this are
// launched with FIB_BLOCK threads and advance by whole blocks, so all lanes
// of a block take the same branch.
__device__ __forceinline__ bool block_gated_out(const int *gate, const int *__restrict__ binst,
                                                int p) {
  if (!gate) return false;
  const int b = __builtin_amdgcn_readfirstlane(p / FIB_BLOCK);
  return !gate[1 + binst[b]];
}

// Element loops of the list-build kernels run grid-stride over a capped grid
// (GRID_CAP workgroups): a build skipped by its gate then costs a small launch
// instead of one workgroup per 256 atoms.
#define GRID_CAP 2048
#define GRID_STRIDE(p, n) \
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < (n); p += gridDim.x * blockDim.x)

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

// LDS hand-off between the lanes of one wavefront (no workgroup barrier:
// the wavefronts of a workgroup may work independently).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
// The raw differences a - b of ref_r2, and its sum of squares when every
// NINT shift is zero (|d| < L/4 makes |d/L| < 1/2): then d - 0*L = d and the
// result equals ref_r2 bit for bit.
__device__ __forceinline__ void ref_diff(const double4 &a, const double4 &b, double &d0,
                                         double &d1, double &d2) {
  d0 = a.x - b.x; d1 = a.y - b.y; d2 = a.z - b.z;
}
__device__ __forceinline__ double ref_sq(double d0, double d1, double d2) {
#pragma clang fp contract(off)
  return d0 * d0 + d1 * d1 + d2 * d2;
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

// Streaming accesses: arrays that a force pass reads or writes once per step
// (its own rows, counts and per-atom results), loaded and stored
// non-temporally so that they do not push the gathered neighbour positions
// out of L2.  FIB_NT_LD (loads) and FIB_NT_ST (stores) hold one bit per
// pass: 1 k_legs, 2 k_tri, 4 k_force3, 8 k_md_pos (velocities, forces,
// build positions, displacements) and 16 (stores only) k_nbuild's rows.
// k_legs reads each list row as four or five 16-byte loads in turn, and
// non-temporal loads fetch the row again for each of them (k_legs 0.70 ->
// 1.04 ms on the 8.2 M-atom dose batch), so its loads stay cached; so do
// k_force3's reads of its own G entries, which the partners' lookups read
// again.  On that batch, streaming the force passes' other accesses makes
// the step 1.5 % faster (mixed), k_md_pos's a further 1.5 % (k_md_pos
// itself 13 %), and k_nbuild's row stores make the next k_md_pos faster too.
#ifndef FIB_NT_LD
#define FIB_NT_LD 14
#endif
#ifndef FIB_NT_ST
#define FIB_NT_ST 31
#endif
template <class V>
struct NtVec { typedef V type; };
template <>
struct NtVec<int4> { typedef int __attribute__((ext_vector_type(4))) type; };
template <>
struct NtVec<float4> { typedef float __attribute__((ext_vector_type(4))) type; };
template <>
struct NtVec<double4> { typedef double __attribute__((ext_vector_type(4))) type; };
template <bool NT, class V>
__device__ __forceinline__ V ld_nt(const V *p) {
  if constexpr (NT) {
    typedef typename NtVec<V>::type N;
    static_assert(sizeof(N) == sizeof(V), "vector size");
    const N t = __builtin_nontemporal_load(reinterpret_cast<const N *>(p));
    V v;
    __builtin_memcpy(&v, &t, sizeof(V));
    return v;
  } else {
    return *p;
  }
}
template <bool NT, class V>
__device__ __forceinline__ void st_nt(V *p, const V &v) {
  if constexpr (NT) {
    typedef typename NtVec<V>::type N;
    static_assert(sizeof(N) == sizeof(V), "vector size");
    N t;
    __builtin_memcpy(&t, &v, sizeof(V));
    __builtin_nontemporal_store(t, reinterpret_cast<N *>(p));
  } else {
    *p = v;
  }
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
// Reductions of NV doubles per slot.  Slot operations: RED_SUM, RED_MAX, and
// the exact top two of a set of values over three slots (RED_TOP2 at slot k:
// the largest; k + 1 RED_TOP2_CNT: how many values equal it; k + 2
// RED_TOP2_SND: the largest value below it, -1e300 when there is none).
// Every operation is associative and commutative and the combining order is
// fixed, so results are deterministic.
enum { RED_SUM = 0, RED_MAX = 1, RED_TOP2 = 2, RED_TOP2_CNT = 3, RED_TOP2_SND = 4 };

__device__ __forceinline__ double red_identity(int op) {
  return (op == RED_SUM || op == RED_TOP2_CNT) ? 0.0 : -1.0e300;
}

__device__ __forceinline__ void top2_merge(double &m, double &c, double &s, double m2, double c2,
                                           double s2) {
  const double M = fmax(m, m2);
  const double cn = (m == M ? c : 0.0) + (m2 == M ? c2 : 0.0);
  const double sn = fmax(m < M ? m : s, m2 < M ? m2 : s2);
  m = M;
  c = cn;
  s = sn;
}

// x <- x (op) y, slot by slot (a top-two triple at once).
template <int NV, typename OP>
__device__ __forceinline__ void red_merge(double (&x)[NV], const double (&y)[NV], const OP &op) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    if (op[k] == RED_TOP2) {
      if (k + 2 < NV) top2_merge(x[k], x[k + 1], x[k + 2], y[k], y[k + 1], y[k + 2]);
    } else if (op[k] == RED_SUM) {
      x[k] += y[k];
    } else if (op[k] == RED_MAX) {
      x[k] = fmax(x[k], y[k]);
    }
  }
}

// Block reduction into per-block partials, layout partials[k * pstride +
// block]; k_finalize combines them per instance.
template <int NV, int BS>
__device__ void block_partials(double (&v)[NV], const int (&op)[NV], int lblock, int pstride,
                               double *__restrict__ partials) {
  __shared__ double red[BS / FIB_WAVE][NV];
  const int lane = threadIdx.x % FIB_WAVE, wid = threadIdx.x / FIB_WAVE;
  for (int o = FIB_WAVE / 2; o > 0; o >>= 1) {
    double y[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) y[k] = __shfl_xor(v[k], o);
    red_merge<NV>(v, y, op);
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wid][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) x[k] = red[0][k];
    for (int w = 1; w < BS / FIB_WAVE; ++w) {
      double y[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k) y[k] = red[w][k];
      red_merge<NV>(x, y, op);
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)k * pstride + lblock] = x[k];
  }
}

#define FIN_T FIB_FIN_T
// Per instance, combine the partials of its blocks in a fixed order.  One
// workgroup per instance when the instance has at most FIN_T blocks (thread t
// takes block t, then a fixed tree); otherwise workgroup (inst, c) reduces
// blocks [c*FIN_T, (c+1)*FIN_T) into tmp and k_finalize2 combines the
// chunks in order -- a big instance then spreads over many CUs.  All NV
// slots at once, so the loads of every slot are in flight together.
template <int NV>
__device__ __forceinline__ void fin_block_tree(double (&x)[NV], const RedSpec &rs,
                                               double (*red)[NV]) {
  const int lane = threadIdx.x % FIB_WAVE, wid = threadIdx.x / FIB_WAVE;
  for (int o = FIB_WAVE / 2; o > 0; o >>= 1) {
    double y[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) y[k] = __shfl_xor(x[k], o);
    red_merge<NV>(x, y, rs.op);
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wid][k] = x[k];
  }
  __syncthreads();
}
template <int NV>
__global__ __launch_bounds__(FIN_T) void k_finalize(const DInst *__restrict__ I,
                                                    const double *__restrict__ partials,
                                                    int pstride, RedSpec rs,
                                                    DCtl *__restrict__ ctl,
                                                    const int *__restrict__ gate,
                                                    double *__restrict__ tmp,
                                                    const int *__restrict__ halt) {
  GATE_OR_RETURN(gate);
  HALT_OR_RETURN(halt);
  __shared__ double red[FIN_T / FIB_WAVE][NV];
  const int inst = blockIdx.x, ch = blockIdx.y, nch = gridDim.y;
  if (gated_out(gate, inst)) return;  // its producer skipped its blocks
  const int blk0 = I[inst].atom_off / rs.bs, nblk = I[inst].natm_pad / rs.bs;
  double x[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) x[k] = red_identity(rs.op[k]);
  const int b_lo = nch == 1 ? 0 : ch * FIN_T;
  const int b_hi = nch == 1 ? nblk : min(nblk, b_lo + FIN_T);
  for (int b = b_lo + (int)threadIdx.x; b < b_hi; b += FIN_T) {
    double y[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) y[k] = partials[(size_t)k * pstride + blk0 + b];
    for (int r = 1; r < rs.nreg; ++r) {
      double y2[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k) y2[k] = partials[(size_t)(r * FIB_NRED + k) * pstride + blk0 + b];
      red_merge<NV>(y, y2, rs.op);
    }
    red_merge<NV>(x, y, rs.op);
  }
  fin_block_tree<NV>(x, rs, red);
  if (threadIdx.x == 0) {
    double t[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) t[k] = red[0][k];
    for (int w = 1; w < FIN_T / FIB_WAVE; ++w) {
      double y[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k) y[k] = red[w][k];
      red_merge<NV>(t, y, rs.op);
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      if (nch == 1) {
        if (rs.dst[k] >= 0) reinterpret_cast<double *>(&ctl[inst])[rs.dst[k]] = t[k];
      } else {
        tmp[((size_t)inst * nch + ch) * NV + k] = t[k];
      }
    }
  }
}
template <int NV>
__global__ void k_finalize2(const DInst *__restrict__ I, int nch, RedSpec rs,
                            DCtl *__restrict__ ctl, const int *__restrict__ gate,
                            const double *__restrict__ tmp, const int *__restrict__ halt) {
  GATE_OR_RETURN(gate);
  HALT_OR_RETURN(halt);
  const int inst = blockIdx.x;
  if (threadIdx.x != 0 || gated_out(gate, inst)) return;
  const int nci = (I[inst].natm_pad / rs.bs + FIN_T - 1) / FIN_T;
  const double *t0 = tmp + (size_t)inst * nch * NV;
  double t[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) t[k] = red_identity(rs.op[k]);
  for (int c = 0; c < nci; ++c) {
    double y[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) y[k] = t0[(size_t)c * NV + k];
    red_merge<NV>(t, y, rs.op);
  }
#pragma unroll
  for (int k = 0; k < NV; ++k)
    if (rs.dst[k] >= 0) reinterpret_cast<double *>(&ctl[inst])[rs.dst[k]] = t[k];
}

// ---------------------------------------------------------------------------
// binning.  Atoms arrive mostly in cell order (the previous sort), so the
// lanes of a wavefront form runs of equal buckets: each run takes its ranks
// with one atomic add of its length.  Ranks only need to be distinct; k_fixup
// orders every bucket by previous position.
__global__ void k_bin(const double4 *__restrict__ X, int NA, const DInst *__restrict__ I,
                      int *__restrict__ bkey, int *__restrict__ brank, int *__restrict__ bcnt,
                      DCtl *__restrict__ ctl, const int *__restrict__ binst,
                      const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  const int lane = threadIdx.x % FIB_WAVE;
  const uint64_t le = (lane == FIB_WAVE - 1) ? ~0ull : ((2ull << lane) - 1ull);  // lanes <= me
  // NA and the grid stride are multiples of the wavefront: the lanes of a
  // wavefront run the loop together
  GRID_STRIDE(p, NA) {
    // atoms of instances that are not rebuilt keep no key (k_scatter skips
    // the same blocks)
    if (block_gated_out(gate, binst, p)) continue;
    const double4 x = X[p];
    const uint32_t m = meta_of(x);
    int b = -1;
    if (!(m & META_PAD)) {
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
        bool drop = (c0 == in.nc[0] + 1 && in.nc[0] > 2) ||
                    (c1 == in.nc[1] + 1 && in.nc[1] > 2) ||
                    (c2 == in.nc[2] + 1 && in.nc[2] > 2) || c0 < 1 || c1 < 1 || c2 < 1 ||
                    c0 > in.nc[0] + 1 || c1 > in.nc[1] + 1 || c2 > in.nc[2] + 1;
        if (drop) local = in.nbucket - 2;
        else local = ((c2 - 1) * in.next[1] + (c1 - 1)) * in.next[0] + (c0 - 1);
      }
      b = in.bucket_off + local;
    }
    // runs of equal buckets among the lanes; skipped lanes (-1) join none
    const int prev = __shfl_up(b, 1);
    const bool start = lane == 0 || prev != b || b < 0;
    const uint64_t starts = __ballot(start);
    const int lead = 63 - __clzll(starts & le);
    int base = 0;
    if (start && b >= 0) {
      const uint64_t later = starts & ~le;
      const int end = later ? __ffsll((unsigned long long)later) - 1 : FIB_WAVE;
      base = atomicAdd(&bcnt[b], end - lane);
    }
    base = __shfl(base, lead);
    bkey[p] = b;
    if (b >= 0) brank[p] = base + (lane - lead);
  }
}

// Zero fill (used instead of a memset node so the step can be captured).
__global__ void k_zero_i32(int *__restrict__ p, int n, const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  GRID_STRIDE(k, n) p[k] = 0;
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

__global__ void k_scan_reduce(const int *__restrict__ in, int n, int *__restrict__ bsum,
                              const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
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

__global__ void k_scan_top(int *__restrict__ bsum, int nb, const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
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

// out[b] = exclusive prefix over all buckets (made instance-local by
// k_scan_base / k_scan_fix)
__global__ void k_scan_down(const int *__restrict__ in, int n, const int *__restrict__ bsum,
                            const DInst *__restrict__ I, int ninst, int *__restrict__ out,
                            const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
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
    if (idx < n) out[idx] = ex;
    ex += vals[e];
  }
}

// Instance-local start positions: bucket b of instance k starts at
// atom_off(k) + (atoms of k in buckets before b), whatever other instances
// hold, so an instance can be sorted while the others stay in place.
__global__ void k_scan_base(const int *__restrict__ out, const DInst *__restrict__ I, int ninst,
                            int *__restrict__ ibase, const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  for (int k = threadIdx.x; k < ninst; k += blockDim.x)
    ibase[k] = out[I[k].bucket_off] - I[k].atom_off;
}
__global__ void k_scan_fix(int *__restrict__ out, int n, const DInst *__restrict__ I, int ninst,
                           const int *__restrict__ ibase, const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  GRID_STRIDE(b, n) out[b] -= ibase[find_inst_by_bucket(I, ninst, b)];
}

__global__ void k_scatter(const int *__restrict__ bkey, const int *__restrict__ brank, int NA,
                          const int *__restrict__ cstart, int *__restrict__ perm,
                          const int *__restrict__ binst, const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  GRID_STRIDE(p, NA) {
    if (block_gated_out(gate, binst, p)) continue;
    const int b = bkey[p];
    if (b >= 0) perm[cstart[b] + brank[p]] = p;
  }
}

// Deterministic order inside a bucket: ascending previous position.
__global__ void k_fixup(const int *__restrict__ cstart, const int *__restrict__ bcnt, int nb,
                        int *__restrict__ perm, const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  GRID_STRIDE(b, nb) {
    const int n = bcnt[b];
    if (n < 2) continue;
    int *a = perm + cstart[b];
    for (int i = 1; i < n; ++i) {
      int v = a[i], j = i - 1;
      while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; --j; }
      a[j + 1] = v;
    }
  }
}

__global__ void k_reorder(int NA, const int *__restrict__ perm, const DInst *__restrict__ I,
                          const double4 *__restrict__ X0, double4 *__restrict__ X1,
                          const double *__restrict__ V0, double *__restrict__ V1,
                          const double *__restrict__ F0, double *__restrict__ F1,
                          const double *__restrict__ Xi0, double *__restrict__ Xi1,
                          int *__restrict__ inv, const int *__restrict__ binst,
                          const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  GRID_STRIDE(p, NA) {
  if (block_gated_out(gate, binst, p)) continue;
  uint32_t mp = meta_of(X0[p]);
  if (mp & META_PAD) continue;
  int s = perm[p];
  double4 x = X0[s];
  X1[p] = x;
  for (int c = 0; c < 3; ++c) {
    size_t a = (size_t)c * NA;
    V1[a + p] = V0[a + s];
    if (F0) F1[a + p] = F0[a + s];
    Xi1[a + p] = Xi0[a + s];
  }
  const DInst &in = I[inst_of(mp)];
  inv[in.atom_off + orig_of(x)] = p;
  }
}

// In-place variant of the sort: copy the reordered arrays back, so the
// arrays stay where a captured step expects them.
__global__ void k_copyback(int NA, const double4 *__restrict__ X1, double4 *__restrict__ X0,
                           const double *__restrict__ V1, double *__restrict__ V0,
                           const double *__restrict__ F1, double *__restrict__ F0,
                           const double *__restrict__ Xi1, double *__restrict__ Xi0,
                           const int *__restrict__ binst, const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  GRID_STRIDE(p, NA) {
    if (block_gated_out(gate, binst, p)) continue;
    X0[p] = X1[p];
    for (int c = 0; c < 3; ++c) {
      const size_t a = (size_t)c * NA + p;
      V0[a] = V1[a];
      if (F0) F0[a] = F1[a];
      Xi0[a] = Xi1[a];
    }
  }
}

// ---------------------------------------------------------------------------
// neighbour build: one 64-lane block per real cell
#define NB_CHUNK 128
#define NB_ACC 384  // candidates whose acceptance masks are kept for the ranks
#ifndef NB_CELLS
#define NB_CELLS 8  // cells per k_nbuild wavefront
#endif
#ifndef NB_WAVES
#define NB_WAVES 4  // independent wavefronts per k_nbuild workgroup
#endif
#define OFF_W 32    // bytes per atom of the stencil-offset table (27 used)
// A row entry of rk_sw holds (stencil cell << 8 | rank) as the build wrote
// it, or, once a force pass has worked it out, RK_SLOT | the slot of the
// row's atom in the neighbour's row (values >= the list width, at most 127,
// mark a neighbour without a row or with the atom past its end).
#define RK_SLOT 0x8000
#define NB_JBITS 27  // atom index bits beside the stencil cell in a staged w (fib_create keeps NA below)

// Stencil bookkeeping for the force gather.  The list row of an atom is the
// concatenation, in the fixed order s = 0..26 of the 27-cell stencil, of the
// accepted atoms of each stencil cell in cell order.  For every atom the
// build stores off[s] = row position where stencil cell s starts (255 for an
// atom without a row), and for every row entry (i -> j) the stencil index s
// of j's cell and the rank of i among the atoms of i's cell that j accepts.
// Stencil offsets are symmetric (cell 26 - s seen from j) and so is the pair
// predicate, so the position of i in j's row is off_j[26 - s] + rank -- no
// search of j's row is needed.  The rank is counted from the acceptance
// masks of the cell's earlier atoms: j accepts a exactly when a accepted j.
//
// Neighbour cell index along one axis (d in {-1, 0, 1}): N > 2 wraps
// periodically ((c + d + N) mod N with one add or subtract); N <= 2 does not
// wrap and index N is the phantom layer holding atoms whose cell id came out
// as N+1; a stencil cell outside [0, N] does not exist.
__device__ __forceinline__ int stencil_axis(int c, int d, int n, bool &ok) {
  int v = c + d;
  if (n > 2) {
    v = v < 0 ? v + n : (v >= n ? v - n : v);
  } else if (v < 0 || v > n) {
    ok = false;
  }
  return v;
}

// One cell (cx, cy, cz) of instance `in`, built by one wave.
__device__ __forceinline__ void nbuild_cell(
    const double4 *__restrict__ X, const DInst &in, int cx, int cy, int cz, int sd0, int sd1,
    int sd2, const int *__restrict__ cstart, const int *__restrict__ bcnt, double rv2,
    double rv2lj, int Wsw, int Wion, int *__restrict__ cnt_sw, int *__restrict__ nbr_sw,
    int *__restrict__ cnt_ion, int *__restrict__ nbr_ion, uint16_t *__restrict__ rk_sw,
    uint8_t *__restrict__ off_sw, double4 *s_x, int *s_j, uint8_t *s_st, int *s_rs, int *s_off,
    uint32_t *s_acc) {
  const int lane = threadIdx.x % FIB_WAVE;
  constexpr bool NBNT = (FIB_NT_ST & 16) != 0;
  // the previous cell of this wave is done with the shared arrays
  wave_lds_sync();
  const int bi = in.bucket_off + (cz * in.next[1] + cy) * in.next[0] + cx;
  const int i0 = cstart[bi], ni = bcnt[bi];
  if (ni == 0) return;

  // lane s < 27 owns stencil cell s = (sd0, sd1, sd2) + 1 in base 3
  int rs = 0, rc = 0;
  if (lane < 27) {
    bool ok = true;
    const int e0 = stencil_axis(cx, sd0, in.nc[0], ok);
    const int e1 = stencil_axis(cy, sd1, in.nc[1], ok);
    const int e2 = stencil_axis(cz, sd2, in.nc[2], ok);
    if (ok) {
      const int b = in.bucket_off + (e2 * in.next[1] + e1) * in.next[0] + e0;
      rs = cstart[b];
      rc = bcnt[b];
    }
  }
  // inclusive scan of rc over lanes 0..31 (the stencil lanes)
  int sc = rc;
  for (int o = 1; o < 32; o <<= 1) {
    int y = __shfl_up(sc, o);
    if (lane >= o) sc += y;
  }
  const int total = __builtin_amdgcn_readfirstlane(__shfl(sc, 26));
  // stencil cell s covers candidates [my_lo, my_lo + rc)
  const int my_lo = sc - rc;
  // the candidates of a 64-batch starting at b that lie before stencil cell
  // s, as a lane mask (lane s); the one-chunk cell has batches 0 and 64 only
  auto before_mask = [&](int b) -> uint64_t {
    const int cut = min(max(my_lo - b, 0), 64);
    return cut >= 64 ? ~0ull : ((1ull << cut) - 1ull);
  };
  const bool masks = total <= NB_ACC && ni <= 32;
  if (masks)
    for (int g = lane; g < total; g += FIB_WAVE) s_acc[g] = 0;
  const bool one_chunk = total <= NB_CHUNK;
  const uint64_t lm0 = before_mask(0), lm1 = before_mask(FIB_WAVE);
  if (one_chunk) {
    // each stencil lane writes its own run of candidates
    for (int q = 0; __ballot(q < rc) != 0ull; ++q)
      if (q < rc) {
        s_j[my_lo + q] = rs + q;
        s_st[my_lo + q] = (uint8_t)lane;
      }
  } else {
    if (lane < 27) { s_rs[lane] = rs; s_off[lane + 1] = sc; }
    if (lane == 0) s_off[0] = 0;
  }
  wave_lds_sync();

  const double4 *Xc = X;
  // the cell's own atoms are stencil range 13: read them from the staged
  // chunk when it holds them (always, unless the halo is very dense)
  const int own0 = __builtin_amdgcn_readfirstlane(__shfl(my_lo, 13));
  const bool own_staged = one_chunk;
  const double qb0 = 0.25 * in.lb[0], qb1 = 0.25 * in.lb[1], qb2 = 0.25 * in.lb[2];
  for (int ii = 0; ii < ni; ++ii) {
    const int i = i0 + ii;
    int *const rowj = nbr_sw + (size_t)i * Wsw;
    uint16_t *const rowk = rk_sw + (size_t)i * Wsw;
    int *const rowi = nbr_ion + (size_t)i * Wion;
    int nsw = 0, nion = 0;
    int before = 0;  // lane s: accepted candidates before stencil cell s
    double4 xi;
    int ti = 0;
    for (int base = 0; base < total; base += NB_CHUNK) {
      const int n = min(NB_CHUNK, total - base);
      // (re)stage the chunk; a single chunk is loaded once per cell
      if (ii == 0 || !one_chunk) {
        if (!one_chunk) {
          wave_lds_sync();
          for (int q = lane; q < n; q += FIB_WAVE) {
            const int g = base + q;
            // stencil range holding g: binary search over the 28 offsets
            int lo = 0, hi = 26;
            while (lo < hi) {
              const int mid = (lo + hi + 1) >> 1;
              if (s_off[mid] <= g) lo = mid; else hi = mid - 1;
            }
            s_j[q] = s_rs[lo] + (g - s_off[lo]);
            s_st[q] = (uint8_t)lo;
          }
          wave_lds_sync();
        }
        // the staged copy carries the candidate's index and stencil cell in
        // the upper word of w (the original index, which the build does not
        // use): one 32-byte read per candidate in the batches below
        for (int q = lane; q < n; q += FIB_WAVE) {
          const int j = s_j[q];
          double4 x = Xc[j];
          const uint64_t up = ((uint64_t)s_st[q] << NB_JBITS) | (uint64_t)j;
          x.w = __longlong_as_double((long long)((up << 32) | (uint64_t)meta_of(x)));
          s_x[q] = x;
        }
        wave_lds_sync();
      }
      if (base == 0) {
        xi = own_staged ? s_x[own0 + ii] : Xc[i];
        ti = type_of(meta_of(xi));
      }
      for (int q0 = 0; q0 < n; q0 += FIB_WAVE) {
        const int q = q0 + lane;  // < NB_CHUNK: the stage arrays are read unguarded
        const int g = base + q;
        const double4 xj = s_x[q];
        const uint32_t up = (uint32_t)(wbits(xj) >> 32);
        const int j = (int)(up & ((1u << NB_JBITS) - 1u));
        const int st = (int)(up >> NB_JBITS);
        const bool cand = q < n && j != i;
        double d0, d1, d2;
        ref_diff(xi, xj, d0, d1, d2);
        const bool far = cand && !(fabs(d0) < qb0 && fabs(d1) < qb1 && fabs(d2) < qb2);
        // With every difference of the batch under a quarter box, the
        // reference's NINT shifts are all zero: r^2 is the plain sum
        // (bit-identical); otherwise the full minimum image
        double r2;
        if (__ballot(far) == 0ull) r2 = ref_sq(d0, d1, d2);
        else r2 = ref_r2(xi, xj, in);
        const bool swp = max(ti, type_of(meta_of(xj))) < 3;
        const bool sw = cand && swp && r2 < rv2;
        const bool io = cand && !swp && r2 < rv2lj;
        uint64_t msw = __ballot(sw), mio = __ballot(io);
        if (sw) {
          // rank of i among the atoms of its cell that j lists (in cell order)
          int rank = 0;
          if (masks) {
            const uint32_t acc = s_acc[g];
            rank = __popc(acc & ((ii == 0) ? 0u : (~0u >> (32 - ii))));
            s_acc[g] = acc | (1u << ii);
          } else {
            for (int a = 0; a < ii; ++a) {
              const int ia = i0 + a;
              if (ia == j) continue;
              const double4 xa = Xc[ia];
              if (type_of(meta_of(xa)) < 3 && ref_r2(xj, xa, in) < rv2) ++rank;
            }
          }
          const int pos = nsw + (int)__builtin_amdgcn_mbcnt_hi(
                                    (uint32_t)(msw >> 32),
                                    __builtin_amdgcn_mbcnt_lo((uint32_t)msw, 0u));
          if (pos < Wsw) {
            st_nt<NBNT>(rowj + pos, j);
            st_nt<NBNT>(rowk + pos, (uint16_t)((st << 8) | (rank & 0xff)));
          }
        }
        if (mio) {
          if (io) {
            const int pos = nion + (int)__builtin_amdgcn_mbcnt_hi(
                                       (uint32_t)(mio >> 32),
                                       __builtin_amdgcn_mbcnt_lo((uint32_t)mio, 0u));
            if (pos < Wion) st_nt<NBNT>(rowi + pos, j);
          }
        }
        // stencil cell s starts at candidate my_lo: count the accepted ones
        // of this batch of 64 candidates that lie before it
        {
          const int b = base + q0;
          const uint64_t lm = b == 0 ? lm0 : (b == FIB_WAVE ? lm1 : before_mask(b));
          before += __popcll(msw & lm);
        }
        nsw += __popcll(msw);
        nion += __popcll(mio);
      }
    }
    if (lane < 27) st_nt<NBNT>(off_sw + (size_t)i * OFF_W + lane, (uint8_t)min(before, 254));
    if (lane == 0) {
      st_nt<NBNT>(cnt_sw + i, nsw);
      st_nt<NBNT>(cnt_ion + i, nion);
    }
  }
}

// Instance k and its first cell number in the build's numbering of cells:
// all real cells in instance order, or under a gate the real cells of the
// gated instances only (prefix cplan, gated-out instances have width 0).
__device__ __forceinline__ int locate_cell(const DInst *__restrict__ I, int ninst,
                                           const int *__restrict__ cplan, int c, int &first) {
  int lo = 0, hi = ninst - 1;
  if (cplan) {
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (cplan[mid] <= c) lo = mid; else hi = mid - 1;
    }
    first = cplan[lo];
  } else {
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (I[mid].cell_real_off <= c) lo = mid; else hi = mid - 1;
    }
    first = I[lo].cell_real_off;
  }
  return lo;
}

// Each wavefront builds NBC consecutive cells (NB_CELLS by default); a workgroup holds
// NB_WAVES independent wavefronts (their own LDS, wave-level sync only), so
// that the workgroup-per-CU limit does not cap the resident waves.  The cells
// to build are all real cells, or under a gate the cells of the gated
// instances, numbered by the prefix that k_md_decide writes after the gates;
// the grid covers all real cells and the workgroups past the gated ones leave
// at once (gated-out instances cost a launch of empty workgroups, no more).
// The busy workgroups are remapped so that each XCD builds a contiguous run
// of cells, whose halos then stay in its L2.  A wave decodes its first cell
// once and steps through the rest in x-fastest order.  The occupancy hint
// asks for seven waves per SIMD; allocation still ends at six (73 VGPRs),
// but the schedule built for seven runs 3-4 % faster than the one for six
// (same-box timing of both builds).
template <int NBC>
__global__ __launch_bounds__(NB_WAVES * FIB_WAVE) __attribute__((amdgpu_waves_per_eu(7)))
void k_nbuild(
    const double4 *__restrict__ X, const DInst *__restrict__ I, int ninst, int total_cells,
    const int *__restrict__ cstart, const int *__restrict__ bcnt, double rv2, double rv2lj,
    int Wsw, int Wion, int *__restrict__ cnt_sw, int *__restrict__ nbr_sw,
    int *__restrict__ cnt_ion, int *__restrict__ nbr_ion, uint16_t *__restrict__ rk_sw,
    uint8_t *__restrict__ off_sw, const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  __shared__ double4 s_x[NB_WAVES][NB_CHUNK];
  __shared__ int s_j[NB_WAVES][NB_CHUNK];
  __shared__ uint8_t s_st[NB_WAVES][NB_CHUNK];
  __shared__ int s_rs[NB_WAVES][27], s_off[NB_WAVES][28];
  __shared__ uint32_t s_acc[NB_WAVES][NB_ACC];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / FIB_WAVE);  // wave-uniform
  const int *cplan = gate ? gate + 1 + ninst : nullptr;
  const int ncell = gate ? cplan[ninst] : total_cells;
  constexpr int CPB = NB_WAVES * NBC;  // cells per workgroup
  const int nbusy = (ncell + CPB - 1) / CPB;
  if ((int)blockIdx.x >= nbusy) return;
  const int c0 = xcd_remap(blockIdx.x, nbusy) * CPB + w * NBC;
  const int c1 = min(ncell, c0 + NBC);
  const int lane = threadIdx.x % FIB_WAVE;
  const int sd0 = lane % 3 - 1, sd1 = (lane / 3) % 3 - 1, sd2 = lane / 9 - 1;
  int cend = c0, k = 0, cx = 0, cy = 0, cz = 0;
  for (int c = c0; c < c1; ++c) {
    if (c >= cend) {
      // first cell, or the first of the next instance
      int first;
      k = locate_cell(I, ninst, cplan, c, first);
      const DInst &in = I[k];
      const int loc = c - first;
      cx = loc % in.nc[0];
      cy = (loc / in.nc[0]) % in.nc[1];
      cz = loc / (in.nc[0] * in.nc[1]);
      cend = first + in.nc[0] * in.nc[1] * in.nc[2];
    }
    const DInst &in = I[k];
    nbuild_cell(X, in, cx, cy, cz, sd0, sd1, sd2, cstart, bcnt, rv2, rv2lj, Wsw, Wion, cnt_sw,
                nbr_sw, cnt_ion, nbr_ion, rk_sw, off_sw, s_x[w], s_j[w], s_st[w], s_rs[w],
                s_off[w], s_acc[w]);
    if (++cx == in.nc[0]) {
      cx = 0;
      if (++cy == in.nc[1]) { cy = 0; ++cz; }
    }
  }
}

// Atoms outside every scanned cell (dropped, dead, or phantom i-atoms) get
// empty lists; k_nbuild only writes the atoms of real cells.
__global__ void k_clear_counts(int NA, int *__restrict__ cnt_sw, int *__restrict__ cnt_ion,
                               uint8_t *__restrict__ off_sw, const int *__restrict__ binst,
                               const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  GRID_STRIDE(p, NA) {
    if (block_gated_out(gate, binst, p)) continue;
    cnt_sw[p] = 0;
    cnt_ion[p] = 0;
    // no row until k_nbuild writes one: offsets past any row (see k_force3)
    uint4 *o = reinterpret_cast<uint4 *>(off_sw + (size_t)p * OFF_W);
    o[0] = make_uint4(~0u, ~0u, ~0u, ~0u);
    o[1] = make_uint4(~0u, ~0u, ~0u, ~0u);
  }
}

// Longest SW / ion lists of the build (overflow check).
// Under a gate only the rebuilt instances' blocks take part (their finalize
// skips the others, whose maxima stay those of their own last build).
__global__ __launch_bounds__(FIB_BLOCK) void k_maxcnt(const int *__restrict__ cnt_sw,
                                                      const int *__restrict__ cnt_ion,
                                                      int pstride, double *__restrict__ partials,
                                                      const int *__restrict__ binst,
                                                      const int *__restrict__ gate) {
  GATE_OR_RETURN(gate);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (block_gated_out(gate, binst, i)) return;
  double v[2] = {(double)cnt_sw[i], (double)cnt_ion[i]};
  const int op[2] = {1, 1};
  block_partials<2, FIB_BLOCK>(v, op, blockIdx.x, pstride, partials);
}

// Slot of the n-th set bit of m (n is small).
__device__ __forceinline__ int nth_bit(uint64_t m, int n) {
  for (int r = 0; r < n; ++r) m &= m - 1;
  return __ffsll((unsigned long long)m) - 1;
}

// ---------------------------------------------------------------------------
// Force evaluation in three passes, no atomics.
//
// k_legs (one thread per atom): the atom's buffered SW list is scanned once.
// Pairs inside the pair-force range add the SW pair force (and the ion list
// the Moliere pairs) to the atom in registers.  Pairs inside the three-body
// range are the atom's "legs": unit vector, E = exp(gamma/(r/sigma - a)),
// d ln E/dr, 1/r and the neighbour's index go to the leg arrays, slot-major
// by compact leg index n ([n * NA + i]), legs in list-slot order.
//
// k_tri (one lane per leg): every vertex term centred on atom i,
// h(a, i, b) = leps E(r_a) E(r_b) (cos + 1/3)^2, is a pair of i's legs.  The
// lane owning leg a sums the force on a's atom over all partners b into
// G[n(a)][i] (w: a's third of the energy).  A wavefront takes the legs of 64
// atoms in turn (a prefix sum over the atoms' leg counts in LDS maps lanes to
// legs), so atoms with many legs do not stall the others.
//
// k_force3 (one thread per atom): F_i = own pair/Moliere force - sum of its
// own G row (the reaction on the centre) + the G entries its neighbours wrote
// for it.  Neighbour c's entry for i is the leg of c in the list slot r that
// holds i; its compact index is the popcount of c's leg mask below r.  r
// follows from the stencil offsets the list build stores (see k_nbuild); k_tri
// works the index out per leg and leaves it in RA, so k_force3 makes one
// lookup per leg.  This turns the reference's scatter F(i) += fi, F(j) += fj,
// F(k) += fk (si.f90:544-546) into a gather.  MODE 1 applies the velocity
// update of onestep (timeint.f90:231-236), MODE 2 the MD-step second half kick.
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
  T e0, e1, e2;  // unit vector from the neighbour to the centre
  T rinv;        // 1/r (FP64: m^-1; mixed: sigma_Si^-1)
  T E;           // exp(gamma/(r/sigma - a))
  T g;           // d ln E / dr (same length unit as rinv)
  uint32_t ks;   // neighbour index | (species == Ge) << 31
  uint32_t slot; // list slot of the neighbour in the centre's row
  uint32_t elo;  // mixed: low parts of e0, e1, e2 (see pack_lo); FP64: 0
};

// Low part of an FP32 unit-vector component h of the FP64 value x, in
// units of ulp(h) / 1024 as a signed 10-bit field c of a word (|x - h| <=
// ulp(h) / 2, so the field carries x - h to 2^-34 of |h|).  The three
// fields ride in one word of the vertex pass's stage; with them the pass
// forms cos + 1/3 from (nearly) the FP64 unit vectors, where it cancels
// near the tetrahedral angle.
__device__ __forceinline__ float ulp_of(float h) {
  const uint32_t ex = __float_as_uint(h) & 0x7f800000u;
  return ex > (23u << 23) ? __uint_as_float(ex - (23u << 23)) : 0.0f;
}
__device__ __forceinline__ uint32_t pack_lo(double x, float h, int c) {
  const float u = ulp_of(h);
  if (!(u > 0.0f)) return 0u;
  const int q = min(max((int)rint((x - (double)h) * (1024.0 / (double)u)), -511), 511);
  return ((uint32_t)q & 0x3ffu) << (10 * c);
}
__device__ __forceinline__ int lo_field(uint32_t w, int c) {
  return (int)(w << (22 - 10 * c)) >> 22;
}
// the field times ulp(h) / 1024: for h = m 2^e with m in [0.5, 1), ulp(h)
// = 2^(e - 24), so the field times 2^(e - 34) (h = 0 carries a zero field)
__device__ __forceinline__ float lo_part(uint32_t w, float h, int c) {
  return __builtin_amdgcn_ldexpf((float)lo_field(w, c), __builtin_amdgcn_frexp_expf(h) - 34);
}
// the low parts of the reversed vector
__device__ __forceinline__ uint32_t neg_lo(uint32_t w) {
  uint32_t r = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) r |= ((uint32_t)(-lo_field(w, c)) & 0x3ffu) << (10 * c);
  return r;
}
// a + b - s exactly, for s = fl(a + b) (TwoSum)
__device__ __forceinline__ float two_sum_err(float a, float b, float s) {
  const float bb = s - a;
  return (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ double ks_pack(double, uint32_t ks) {
  return __longlong_as_double((long long)ks);
}
__device__ __forceinline__ float ks_pack(float, uint32_t ks) { return __uint_as_float(ks); }
__device__ __forceinline__ uint32_t ks_unpack(double v) {
  return (uint32_t)__double_as_longlong(v);
}
__device__ __forceinline__ uint32_t ks_unpack(float v) { return __float_as_uint(v); }
__device__ __forceinline__ int ks_index(uint32_t ks) { return (int)(ks & 0x7fffffffu); }
__device__ __forceinline__ int ks_ge(uint32_t ks) { return (int)(ks >> 31); }

// Force on leg a of one vertex term (centre c, legs a and b), and h.
// h = leps E(ra) E(rb) (cos+1/3)^2 with d cos/dx_a = -(e_b - cos e_a)/r_a and
// d r_a/dx_a = -e_a, so F_a = -dh/dx_a = h g_a e_a + 2 leps Ea Eb D (e_b - cos e_a)/r_a.
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
// Mixed: cos and D = cos + 1/3 from the FP32 unit vectors and their low
// parts, with the products' rounding errors (FMA) and the sums' (TwoSum)
// carried along, each rounded to FP32 once; the rest of the term in FP32.
// (The products and sums of cos_d stay unfused: an FMA formed from a
// product and the following sum would break the TwoSum error terms.)
__device__ __forceinline__ void cos_d(const Leg<float> &a, const Leg<float> &b, float &cs,
                                      float &D) {
#pragma clang fp contract(off)
  const float p0 = a.e0 * b.e0, p1 = a.e1 * b.e1, p2 = a.e2 * b.e2;
  const float rp = (__builtin_fmaf(a.e0, b.e0, -p0) + __builtin_fmaf(a.e1, b.e1, -p1)) +
                   __builtin_fmaf(a.e2, b.e2, -p2);
  const float x = ((a.e0 * lo_part(b.elo, b.e0, 0) + lo_part(a.elo, a.e0, 0) * b.e0) +
                   (a.e1 * lo_part(b.elo, b.e1, 1) + lo_part(a.elo, a.e1, 1) * b.e1)) +
                  (a.e2 * lo_part(b.elo, b.e2, 2) + lo_part(a.elo, a.e2, 2) * b.e2);
  const float s1 = p0 + p1, s2 = s1 + p2;
  const float lo = (two_sum_err(p0, p1, s1) + two_sum_err(s1, p2, s2)) + (rp + x);
  cs = s2 + lo;
  // 1/3 = THIRD_HI + THIRD_LO
  const float THIRD_HI = 0.333333343267440796f, THIRD_LO = -9.93410748e-09f;
  const float s3 = s2 + THIRD_HI;
  D = s3 + (two_sum_err(s2, THIRD_HI, s3) + (lo + THIRD_LO));
}
template <>
__device__ __forceinline__ void vterm<float>(float leps, const Leg<float> &a, const Leg<float> &b,
                                             float *Fa, float &hh) {
  float cs, D;
  cos_d(a, b, cs, D);
  const float Eab = leps * a.E * b.E;
  hh = Eab * D * D;
  const float c1 = hh * a.g;
  const float c2 = 2.0f * Eab * D * a.rinv;
  Fa[0] = c1 * a.e0 + c2 * (b.e0 - cs * a.e0);
  Fa[1] = c1 * a.e1 + c2 * (b.e1 - cs * a.e1);
  Fa[2] = c1 * a.e2 + c2 * (b.e2 - cs * a.e2);
}

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

// v[n] of a 4-entry register table by selects (a dynamic index would put
// the table in scratch memory).
template <typename T>
__device__ __forceinline__ T pick4(const T (&v)[4], int n) {
  return n == 0 ? v[0] : n == 1 ? v[1] : n == 2 ? v[2] : v[3];
}

// The leg from centre c to neighbour n (unit vector from n to c, 1/r, E, g),
// with k_legs' arithmetic: FP64 as the reference (si.f90:491-509), or the
// FP32 terms of the mixed mode.  Legs are not stored between passes: k_legs
// decides which pairs are legs (its leg mask), and the passes that need the
// values recompute them here from the positions.
template <bool MIXED>
__device__ __forceinline__ void make_leg(const double4 &xc, const double4 &xn, const DInst &in,
                                         const FConst &fc, int tc, int tn, uint32_t jn,
                                         uint32_t slot, Leg<typename Prec<MIXED>::T> &L) {
  double d0, d1, d2;
  min_image_fast(xc, xn, in, d0, d1, d2);
  const double isg = fc.isig[tc == 2 ? 1 : 0][tn == 2 ? 1 : 0];
  L.ks = jn | (tn == 2 ? 0x80000000u : 0u);
  L.slot = slot;
  L.elo = 0u;
  if (!MIXED) {
    const double rr = d0 * d0 + d1 * d1 + d2 * d2;
    const double r = sqrt(rr);
    const double rinv = 1.0 / r;
    const double rs = r * isg;
    const double inv_ra = 1.0 / (rs - fc.al);
    L.e0 = d0 * rinv; L.e1 = d1 * rinv; L.e2 = d2 * rinv;
    L.rinv = rinv;
    L.E = exp(fc.gamma * inv_ra);
    L.g = -fc.gamma * inv_ra * inv_ra * isg;
  } else {
    // every leg value from FP64 geometry, rounded to FP32 once: the FP32
    // vertex terms cannot recover accuracy their inputs do not carry
    const double s0 = d0 * isg, s1 = d1 * isg, s2 = d2 * isg;
    const double rs = sqrt(s0 * s0 + s1 * s1 + s2 * s2);
    const double rinv = 1.0 / rs;
    const double inv_ra = 1.0 / (rs - fc.al);
    const double rho = fc.rho[tc == 2 ? 1 : 0][tn == 2 ? 1 : 0];
    const double x0 = s0 * rinv, x1 = s1 * rinv, x2 = s2 * rinv;
    L.e0 = (float)x0; L.e1 = (float)x1; L.e2 = (float)x2;
    L.elo = pack_lo(x0, L.e0, 0) | pack_lo(x1, L.e1, 1) | pack_lo(x2, L.e2, 2);
    L.rinv = (float)(rinv * rho);
    L.E = expf((float)(fc.gamma * inv_ra));
    L.g = (float)(-fc.gamma * inv_ra * inv_ra * rho);
  }
}

// k_legs: list slots whose positions are loaded together (4 or 8).  Eight
// make the mixed pass 3.5 % faster on the 8.2 M-atom dose batch (110
// VGPRs, four waves per SIMD); in FP64 (130 VGPRs) they gain nothing.
#ifndef LEGS_BATCH_MIXED
#define LEGS_BATCH_MIXED 8
#endif
#ifndef LEGS_BATCH_FP64
#define LEGS_BATCH_FP64 4
#endif

// Register accumulator of one force vector: FP64, or FP32 with a TwoSum
// compensation term (mixed).
template <bool MIXED>
struct FAcc;
template <>
struct FAcc<false> {
  double x = 0, y = 0, z = 0;
  __device__ void add(double a, double b, double c) { x += a; y += b; z += c; }
  __device__ void get(double s, double &a, double &b, double &c) const {
    a = x * s; b = y * s; c = z * s;
  }
};
template <>
struct FAcc<true> {
  float x = 0, y = 0, z = 0, cx = 0, cy = 0, cz = 0;
  __device__ void add(float a, float b, float c) { kadd(x, cx, a); kadd(y, cy, b); kadd(z, cz, c); }
  __device__ void get(double s, double &a, double &b, double &c) const {
    a = ((double)x + (double)cx) * s;
    b = ((double)y + (double)cy) * s;
    c = ((double)z + (double)cz) * s;
  }
};

template <bool MIXED, bool EN, bool VIR>
__global__ __launch_bounds__(FIB_BLOCK) void k_legs(
    const double4 *__restrict__ X, const DInst *__restrict__ I, const FConst fc,
    const TabEntry *__restrict__ tab, int NA, int Wsw, int Wion, const int *__restrict__ cnt_sw,
    const int *__restrict__ nbr_sw, const int *__restrict__ cnt_ion,
    const int *__restrict__ nbr_ion, int *__restrict__ cnt_leg, uint64_t *__restrict__ amask,
    double *__restrict__ Fpart, double *__restrict__ upart, int pstride,
    double *__restrict__ partials, const int *__restrict__ halt) {
  HALT_OR_RETURN(halt);
  typedef typename Prec<MIXED>::T T;
  constexpr bool NT = (FIB_NT_LD & 1) != 0, NTS = (FIB_NT_ST & 1) != 0;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int i = lb * FIB_BLOCK + threadIdx.x;
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
  const double isgS = fc.isig[0][0];
  const double fscale = MIXED ? fc.eps[0][0] * isgS : 1.0;
  const double lscale = MIXED ? 1.0 / isgS : 1.0;
  // mixed: SW pair force factor eps_ij/sigma_ij in units of eps_Si/sigma_Si
  // and the pair's length unit over sigma_Si, divided on the host
  const float fsA = fc.fsw[tr][0], fsB = fc.fsw[tr][1];
  const float rhoA = fc.rho[tr][0], rhoB = fc.rho[tr][1];
  const size_t row = (size_t)i * Wsw;

  FAcc<MIXED> f;
  double ui = 0, etot = 0;
  T W[9];
#pragma unroll
  for (int w = 0; w < 9; ++w) W[w] = 0;
  auto addW = [&](T d0, T d1, T d2, T f0, T f1, T f2) {
    W[0] += d0 * f0; W[1] += d0 * f1; W[2] += d0 * f2;
    W[3] += d1 * f0; W[4] += d1 * f1; W[5] += d1 * f2;
    W[6] += d2 * f0; W[7] += d2 * f1; W[8] += d2 * f2;
  };

  int na = 0;
  uint64_t am = 0;
  // one buffered SW pair in list slot k: pair force, pair energy, leg
  auto pair = [&](int k, int jn, const double4 &xj) {
    const uint32_t mj = meta_of(xj);
    // an atom removed since the build is out of the simulation (MD-step mode)
    if (mj & META_DEAD) return;
    const int tj = type_of(mj);
    const bool jA = tj != 2;
    const double isg = jA ? isgA : isgB;
    double d0, d1, d2;
    min_image_fast(xi, xj, in, d0, d1, d2);
    if (!MIXED) {
      const double rr = d0 * d0 + d1 * d1 + d2 * d2;
      if (!(rr * (isg * isg) < fc.r2gate)) return;
      const double r = sqrt(rr);
      const double rinv = 1.0 / r;
      const double rs = r * isg;
      // reference pair force: 3-point Lagrange on the derivative table
      // (si.f90:416-424, 554-573)
      const int t1 = (int)(rs / fc.umax * fc.tabn1) + 1;
      if (t1 < fc.izero) {
        const TabEntry te = tab[t1];
        const double e1 = rs - te.u0, e2 = rs - te.u1, e3 = rs - te.u2;
        const double df = 0.5 * e2 * e3 * te.t0 - e1 * e3 * te.t1 + 0.5 * e1 * e2 * te.t2;
        const double fs = -df * (jA ? esgA : esgB) * rinv;
        f.add(fs * d0, fs * d1, fs * d2);
        if (VIR) addW(d0, d1, d2, 0.5 * fs * d0, 0.5 * fs * d1, 0.5 * fs * d2);
      }
      if (rs < fc.cut_sw) {
        const double inv_ra = 1.0 / (rs - fc.al);
        if (EN) {
          const double rs2 = rs * rs;
          const double ul =
              0.5 * (jA ? epsA : epsB) * fc.A * (fc.B / (rs2 * rs2) - 1.0) * exp(inv_ra);
          ui += ul;
          etot += ul;
        }
        am |= 1ull << k;
        ++na;
      }
    } else {
      // Pair geometry in FP64 (pair units: sigma_ij), the pair force factor
      // rounded to FP32 once.  Near the bond length the two parts of dphi/dr
      // cancel to a small remainder, and formed in FP32 their rounding was
      // the largest error of the mixed forces; the exponential and the
      // force vector stay FP32.
      const double s0 = d0 * isg, s1 = d1 * isg, s2 = d2 * isg;
      const double rr = s0 * s0 + s1 * s1 + s2 * s2;
      if (!(rr < fc.cut_sw2)) return;
      const double rs = sqrt(rr);
      if (!(rs < fc.cut_sw)) return;
      const double rinv = 1.0 / rs;
      const double inv_ra = 1.0 / (rs - fc.al);
      const double ex = (double)expf((float)inv_ra);
      const double ir2 = rinv * rinv;
      const double ir4 = ir2 * ir2;
      // phi2(r) = A (B r^-4 - 1) e^{1/(r-a)};  force on i = -dphi/dr e
      const double br = -4.0 * fc.B * ir4 * rinv - (fc.B * ir4 - 1.0) * inv_ra * inv_ra;
      const float fs = (float)(-fc.A * ex * br * rinv * (double)(jA ? fsA : fsB));
      const float r0 = (float)s0, r1 = (float)s1, r2 = (float)s2;
      const float f0 = fs * r0, f1 = fs * r1, f2 = fs * r2;
      f.add(f0, f1, f2);
      const float rho = jA ? rhoA : rhoB;  // 1 for Si-Si
      if (VIR) {
        const float q = 1.0f / rho;
        addW(r0 * q, r1 * q, r2 * q, 0.5f * f0, 0.5f * f1, 0.5f * f2);
      }
      if (EN) {
        const double ul = 0.5 * (jA ? epsA : epsB) * (fc.A * (fc.B * ir4 - 1.0) * ex);
        ui += ul;
        etot += ul;
      }
      // legs with E = exp(gamma/(r-a)) < e^-30 are left out of the
      // three-body sum: each of their terms is below 1e-9 eps/sigma
      if (rs < fc.mixed_xleg) {
        am |= 1ull << k;
        ++na;
      }
    }
  };

  if (sw_atom) {
    const int cs = min(ld_nt<NT>(cnt_sw + i), Wsw);
    const int4 *r4 = reinterpret_cast<const int4 *>(nbr_sw + row);
    // LEGS_BATCH slots at a time: their indices, then their positions; the
    // next indices are loaded while the positions arrive (every index read
    // lies below cs <= Wsw, inside the row)
    constexpr int LEGS_BATCH = MIXED ? LEGS_BATCH_MIXED : LEGS_BATCH_FP64;
    static_assert(LEGS_BATCH == 4 || LEGS_BATCH == 8, "LEGS_BATCH");
    constexpr int NQ = LEGS_BATCH / 4;
    int4 rn[NQ];
#pragma unroll
    for (int u = 0; u < NQ; ++u) rn[u] = (4 * u < cs) ? ld_nt<NT>(r4 + u) : make_int4(i, i, i, i);
    for (int k0 = 0; k0 < cs; k0 += LEGS_BATCH) {
      int jq[LEGS_BATCH];
#pragma unroll
      for (int u = 0; u < NQ; ++u) {
        const int b = k0 + 4 * u;
        jq[4 * u] = (b < cs) ? rn[u].x : i;
        jq[4 * u + 1] = (b + 1 < cs) ? rn[u].y : i;
        jq[4 * u + 2] = (b + 2 < cs) ? rn[u].z : i;
        jq[4 * u + 3] = (b + 3 < cs) ? rn[u].w : i;
      }
      double4 xq[LEGS_BATCH];
#pragma unroll
      for (int v = 0; v < LEGS_BATCH; ++v) xq[v] = X[jq[v]];
#pragma unroll
      for (int u = 0; u < NQ; ++u)
        if (k0 + LEGS_BATCH + 4 * u < cs) rn[u] = ld_nt<NT>(r4 + ((k0 + LEGS_BATCH) >> 2) + u);
#pragma unroll
      for (int v = 0; v < LEGS_BATCH; ++v)
        if (k0 + v < cs) pair(k0 + v, jq[v], xq[v]);
    }
  }
  st_nt<NTS>(cnt_leg + i, na);
  st_nt<NTS>(amask + i, am);

  // ---- Moliere pairs (si.f90:426-446): both atoms must be fired ----------
  if (live) {
    const int ci = min(cnt_ion[i], Wion);
    const bool ifired = !(ti == 3 && orig_of(xi) >= in.nsg + in.impact);
    const int *ri = nbr_ion + (size_t)i * Wion;
    for (int k = 0; ifired && k < ci; ++k) {
      const int j = ri[k];
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
        f.add((T)(f0 / fscale), (T)(f1 / fscale), (T)(f2 / fscale));
        if (VIR) addW((T)(d0 / lscale), (T)(d1 / lscale), (T)(d2 / lscale),
                      (T)(0.5 * f0 / fscale), (T)(0.5 * f1 / fscale), (T)(0.5 * f2 / fscale));
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
        f.add((T)f0, (T)f1, (T)f2);
        if (VIR) addW((T)(d0 / lscale), (T)(d1 / lscale), (T)(d2 / lscale), (T)(0.5f * f0),
                      (T)(0.5f * f1), (T)(0.5f * f2));
        if (EN) { const double ul = 0.5 * zz / aa * (double)(sa * ix); ui += ul; etot += ul; }
      }
    }
  }

  double fx, fy, fz;
  f.get(fscale, fx, fy, fz);
  st_nt<NTS>(Fpart + i, fx);
  st_nt<NTS>(Fpart + (size_t)NA + i, fy);
  st_nt<NTS>(Fpart + 2 * (size_t)NA + i, fz);
  if (EN) upart[i] = ui;
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


// Lane-per-leg three-body pass.  Block = 4 wavefronts, each owning 64
// consecutive atoms.  A wavefront computes the legs of a run of its atoms
// (the pairs of their leg masks, from the positions) into LDS, so the vertex
// terms, which read every leg of the centre once per leg, read LDS; no leg
// record goes through HBM.  Runs end on atom boundaries and hold at most
// TriStage::N legs.
#ifndef TRI_BATCH
#define TRI_BATCH 4  // legs per lane whose neighbour loads are in flight together
#endif
#ifndef TRI_NS_MIXED
#define TRI_NS_MIXED 256  // legs per stage run (mixed)
#endif
#ifndef TRI_NS_FP64
#define TRI_NS_FP64 256  // legs per stage run (FP64); 128 measured 10 % slower on the dose batch
#endif
template <bool MIXED>
struct TriStage {
  static constexpr int N = MIXED ? TRI_NS_MIXED : TRI_NS_FP64;  // legs per run
};
template <bool MIXED, bool EN, bool VIR, int NS, int TB>
__global__ __launch_bounds__(FIB_BLOCK) void k_tri(
    const double4 *__restrict__ X, const DInst *__restrict__ I, const FConst fc, int NA,
    int Wsw, const int *__restrict__ cnt_leg, const int *__restrict__ nbr_sw,
    uint16_t *__restrict__ rk_sw, const uint8_t *__restrict__ off_sw,
    const uint64_t *__restrict__ amask, typename Prec<MIXED>::T4 *__restrict__ G,
    uint32_t *__restrict__ RA, int pstride, double *__restrict__ partials,
    const int *__restrict__ halt) {
  HALT_OR_RETURN(halt);
  typedef typename Prec<MIXED>::T T;
  typedef typename Prec<MIXED>::T4 T4;
  constexpr int NW = FIB_BLOCK / FIB_WAVE;
  constexpr bool NT = (FIB_NT_LD & 2) != 0, NTS = (FIB_NT_ST & 2) != 0;
  // sA: e0, e1, e2 and E, negated when the leg's atom is Ge (E > 0);
  // sB: g, 1/r, neighbour bits, list slot bits
  __shared__ T4 sA[NW][NS];
  __shared__ T4 sB[NW][NS];
  __shared__ uint8_t owner[NW][NS];
  __shared__ int soff[NW][FIB_WAVE + 1];
  __shared__ uint8_t sge[NW][FIB_WAVE];
  __shared__ uint32_t sR[NW][NS];
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x % FIB_WAVE, wid = threadIdx.x / FIB_WAVE;
  const int i0 = lb * FIB_BLOCK + wid * FIB_WAVE;
  const double escale = MIXED ? fc.eps[0][0] : 1.0;
  const double fscale = MIXED ? fc.eps[0][0] * fc.isig[0][0] : 1.0;
  const double lscale = MIXED ? 1.0 / fc.isig[0][0] : 1.0;
  // lambda*eps by number of Ge atoms in the term (symmetric in the three
  // species, prms.f90:464-471), in J or eps_Si
  const double lu = MIXED ? 1.0 / fc.eps[0][0] : 1.0;
  const T lp[4] = {(T)(fc.leps[0][0][0] * lu), (T)(fc.leps[0][0][1] * lu),
                   (T)(fc.leps[0][1][1] * lu), (T)(fc.leps[1][1][1] * lu)};

  // legs of the wave's atoms: exclusive prefix sum in LDS
  const int nl = ld_nt<NT>(cnt_leg + i0 + lane);
  int inc = nl;
  for (int o = 1; o < FIB_WAVE; o <<= 1) {
    const int y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  soff[wid][lane] = inc - nl;
  if (lane == FIB_WAVE - 1) soff[wid][FIB_WAVE] = inc;
  // this lane's atom as a centre; every block lies inside one instance
  const int ic = i0 + lane;
  const double4 xc = X[ic];
  const int tci = type_of(meta_of(xc));
  const uint64_t amc = amask[ic];
  const DInst &in = I[__builtin_amdgcn_readfirstlane(inst_of(meta_of(xc)))];
  sge[wid][lane] = (tci == 2) ? 1 : 0;

  T uc = 0;
  T W[9];
#pragma unroll
  for (int w = 0; w < 9; ++w) W[w] = 0;

  for (int o0 = 0; o0 < FIB_WAVE;) {
    // run of atoms [o0, o1) whose legs [s0, s1) fit the stage (an atom has
    // at most Wsw <= 64 legs, so every run takes at least one atom)
    const int s0 = __shfl(inc - nl, o0);
    const uint64_t fit = __ballot(lane >= o0 && inc - s0 <= NS);
    const int o1 = o0 + __popcll(fit);
    const int s1 = __shfl(inc, o1 - 1);
    const bool mine = lane >= o0 && lane < o1;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // owner of every leg of the run (LDS only), so that the stage below can
    // give each lane legs rather than atoms
    if (mine)
      for (int a = 0; a < nl; ++a) owner[wid][inc - nl + a - s0] = (uint8_t)lane;
    wave_lds_sync();
    // stage: one lane per leg, in list order within each atom, computed from
    // the positions; TB legs per lane with their loads in flight together,
    // every lane busy whatever the atoms' leg counts
    {
      const int nq = s1 - s0;
      const int amlo = (int)(uint32_t)amc, amhi = (int)(uint32_t)(amc >> 32);
      for (int q0 = 0; q0 < nq; q0 += FIB_WAVE * TB) {
        int qv[TB], ov[TB], kq[TB], jq[TB], rq[TB];
#pragma unroll
        for (int u = 0; u < TB; ++u) {
          qv[u] = q0 + u * FIB_WAVE + lane;
          const bool ok = qv[u] < nq;
          ov[u] = ok ? owner[wid][min(qv[u], NS - 1)] : 0;
          // the owner's leg mask, from the owner's lane (every lane takes part)
          const uint64_t am = ((uint64_t)(uint32_t)__shfl(amhi, ov[u]) << 32) |
                              (uint64_t)(uint32_t)__shfl(amlo, ov[u]);
          kq[u] = ok ? nth_bit(am, qv[u] + s0 - soff[wid][ov[u]]) : 0;
          const size_t rw = (size_t)(i0 + ov[u]) * Wsw;
          jq[u] = ok ? ld_nt<NT>(nbr_sw + rw + kq[u]) : i0 + ov[u];
          rq[u] = ok ? (int)ld_nt<NT>(rk_sw + rw + kq[u]) : 0xffff;
        }
        // Where the partner j keeps the force of its own terms on the owner:
        // G[n * NA + j], n = j's leg index of the owner.  The slot r of the
        // owner in j's row follows from the stencil offsets (see k_nbuild), n
        // from j's leg mask (legs are symmetric); 255 offsets (r >= Wsw)
        // mark a j without a row.  The first time a pair is a leg after a
        // build, r is worked out from j's offsets and kept in the row entry
        // (RK_SLOT set); later steps read it there, without the gather of
        // j's offsets.
        double4 xq[TB];
        uint64_t aq[TB];
        int oq[TB];
#pragma unroll
        for (int u = 0; u < TB; ++u) {
          xq[u] = X[jq[u]];
          aq[u] = amask[jq[u]];
          oq[u] = (rq[u] & RK_SLOT) ? 0 : off_sw[(size_t)jq[u] * OFF_W + 26 - (rq[u] >> 8)];
        }
#pragma unroll
        for (int u = 0; u < TB; ++u) {
          // the owner's position and species, from the owner's lane
          double4 xo;
          xo.x = __shfl(xc.x, ov[u]);
          xo.y = __shfl(xc.y, ov[u]);
          xo.z = __shfl(xc.z, ov[u]);
          xo.w = 0.0;
          const int to = __shfl(tci, ov[u]);
          if (qv[u] < nq) {
            Leg<T> L;
            make_leg<MIXED>(xo, xq[u], in, fc, to, type_of(meta_of(xq[u])), (uint32_t)jq[u],
                            (uint32_t)kq[u], L);
            T4 x, y;
            x.x = L.e0; x.y = L.e1; x.z = L.e2; x.w = ks_ge(L.ks) ? -L.E : L.E;
            y.x = L.g; y.y = L.rinv; y.z = ks_pack(T(0), L.ks); y.w = ks_pack(T(0), L.elo);
            const int q = qv[u];
            sA[wid][q] = x;
            sB[wid][q] = y;
            const int r = (rq[u] & RK_SLOT) ? (rq[u] & ~RK_SLOT) : min(oq[u] + (rq[u] & 0xff), 127);
            if (!(rq[u] & RK_SLOT))
              rk_sw[(size_t)(i0 + ov[u]) * Wsw + kq[u]] = (uint16_t)(RK_SLOT | r);
            sR[wid][q] = r < Wsw ? (uint32_t)__popcll(aq[u] & ((1ull << r) - 1ull)) *
                                           (uint32_t)NA + (uint32_t)jq[u]
                                 : ~0u;
          }
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // mixed: each lane's legs' energy shares wait in registers until every
    // lane is done reading the stage's fourth words (the low parts of the
    // unit vectors)
    T gus[MIXED ? NS / FIB_WAVE : 1];
#pragma unroll
    for (int k = 0; k < (MIXED ? NS / FIB_WAVE : 1); ++k) gus[k] = 0;
    for (int g = s0 + lane, it = 0; g < s1; g += FIB_WAVE, ++it) {
      const int q = g - s0;
      const int o = owner[wid][q];
      const int c0 = soff[wid][o] - s0;  // the centre's first leg in the stage
      const int na = soff[wid][o + 1] - soff[wid][o];
      const int a = g - soff[wid][o];
      Leg<T> La;
      {
        const T4 x = sA[wid][q], y = sB[wid][q];
        La.e0 = x.x; La.e1 = x.y; La.e2 = x.z; La.E = fabs(x.w);
        La.g = y.x; La.rinv = y.y; La.ks = ks_unpack(y.z); La.elo = ks_unpack(y.w);
      }
      const int ga = sge[wid][o] + ks_ge(La.ks);
      T Ga0 = 0, Ga1 = 0, Ga2 = 0, Gu = 0;
      for (int b = 0; b < na; ++b) {
        if (b == a) continue;
        const T4 xb = sA[wid][c0 + b];
        Leg<T> Lb;
        Lb.e0 = xb.x; Lb.e1 = xb.y; Lb.e2 = xb.z; Lb.E = fabs(xb.w);
        Lb.elo = MIXED ? ks_unpack(sB[wid][c0 + b].w) : 0u;
        T Fa[3], hh;
        vterm<T>(pick4(lp, ga + (xb.w < T(0) ? 1 : 0)), La, Lb, Fa, hh);
        Ga0 += Fa[0]; Ga1 += Fa[1]; Ga2 += Fa[2];
        Gu += hh;
      }
      // results go back through the stage and leave plane by plane below,
      // with sR, so the stores coalesce: FP64, all four words at once (this
      // lane alone reads sB[q]); mixed, the forces into the first three
      // words now and the energy share after the loop
      if constexpr (MIXED) {
        sB[wid][q].x = Ga0; sB[wid][q].y = Ga1; sB[wid][q].z = Ga2;
#pragma unroll
        for (int k = 0; k < NS / FIB_WAVE; ++k)
          if (k == it) gus[k] = Gu * T(1.0 / 3.0);
      } else {
        T4 gv;
        gv.x = Ga0; gv.y = Ga1; gv.z = Ga2; gv.w = Gu * T(1.0 / 3.0);
        sB[wid][q] = gv;
      }
      uc += Gu;
      if (VIR) {
        // r_a = x_a - x_i = -e_a / rinv_a
        const T ra = -T(1) / La.rinv;
        const T d0 = ra * La.e0, d1 = ra * La.e1, d2 = ra * La.e2;
        W[0] += d0 * Ga0; W[1] += d0 * Ga1; W[2] += d0 * Ga2;
        W[3] += d1 * Ga0; W[4] += d1 * Ga1; W[5] += d1 * Ga2;
        W[6] += d2 * Ga0; W[7] += d2 * Ga1; W[8] += d2 * Ga2;
      }
    }
    if constexpr (MIXED) {
      // every lane is done with the fourth words: the energy shares go there
      wave_lds_sync();
#pragma unroll
      for (int k = 0; k < NS / FIB_WAVE; ++k) {
        const int g = s0 + lane + k * FIB_WAVE;
        if (g < s1) sB[wid][g - s0].w = gus[k];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int a = 0; __ballot(mine && a < nl) != 0ull; ++a) {
      if (mine && a < nl) {
        const size_t at = (size_t)a * NA + i0 + lane;
        const int q = inc - nl + a - s0;
        st_nt<NTS>(G + at, sB[wid][q]);
        st_nt<NTS>(RA + at, sR[wid][q]);
      }
    }
    o0 = o1;
  }
  if (EN || VIR) {
    double v[10];
    // every unordered term was summed twice, as (a, b) and (b, a)
    v[0] = (double)uc * 0.5 * escale;
    const double ws = MIXED ? fscale * lscale : 1.0;
#pragma unroll
    for (int w = 0; w < 9; ++w) v[1 + w] = (double)W[w] * ws;
    const int op[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    block_partials<10, FIB_BLOCK>(v, op, lb, pstride, partials);
  }
}

// ---------------------------------------------------------------------------
// Vertex terms of the build-time graph that the centre's own list misses.
//
// The reference stores its triplets once per list build (si.f90:191-232):
// every connected triple of the buffered graph.  Between builds the vertex
// term centred on c with legs a and b is evaluated whenever both legs are in
// range *now* (si.f90:491-509), so after atoms have moved more than the list
// skin a term can exist whose leg c-a is not a listed pair, the triple being
// connected through the pairs c-b and b-a instead.  k_legs/k_tri only pair
// legs from c's own list; this pass adds the rest, exactly and without
// atomics: the term (a, c, b) is found by the bridge atom b, which lists both
// a and c.  b keeps its own share and writes the forces on a and c into
// EX[slot][b] by its list slots of a and c; a and c gather them like G.
// Nothing can be missing unless two atoms moved together by more than the
// skin, so the pass exits at once while maxdisp + maxdisp2 <= skin.
template <bool MIXED, bool EN, bool VIR>
__global__ __launch_bounds__(FIB_BLOCK) void k_extra(
    const double4 *__restrict__ X, const DInst *__restrict__ I, const FConst fc, int NA, int Wsw,
    const int *__restrict__ cnt_sw, const int *__restrict__ nbr_sw,
    const int *__restrict__ cnt_leg, const uint64_t *__restrict__ amask,
    const float *__restrict__ disp, DCtl *__restrict__ ctl, typename Prec<MIXED>::T4 *__restrict__ EX,
    typename Prec<MIXED>::T4 *__restrict__ EXself, uint64_t *__restrict__ exmask, int pstride,
    double *__restrict__ partials, const int *__restrict__ halt) {
  HALT_OR_RETURN(halt);
  typedef typename Prec<MIXED>::T T;
  typedef typename Prec<MIXED>::T4 T4;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int b = lb * FIB_BLOCK + threadIdx.x;
  const double4 xb = X[b];
  const uint32_t mb = meta_of(xb);
  const int kin = __builtin_amdgcn_readfirstlane(inst_of(mb));
  const DInst &in = I[kin];
  const double isgS = fc.isig[0][0];
  const double fscale = MIXED ? fc.eps[0][0] * isgS : 1.0;
  const double escale = MIXED ? fc.eps[0][0] : 1.0;
  const double lscale = MIXED ? 1.0 / isgS : 1.0;
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int op[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  // block-uniform: one instance per block
  const bool active = ctl[kin].maxdisp + ctl[kin].maxdisp2 > fc.skin;
  if (!active) {
    // nothing to add: clear the instance's flag so the gather skips EX
    if (threadIdx.x == 0 && lb * FIB_BLOCK == in.atom_off) ctl[kin].extra_any = 0;
    if (EN || VIR) block_partials<10, FIB_BLOCK>(v, op, lb, pstride, partials);
    return;
  }
  const bool live = !(mb & (META_PAD | META_DEAD));
  const int tb = type_of(mb);
  const double lu = MIXED ? 1.0 / fc.eps[0][0] : 1.0;
  const T lp[4] = {(T)(fc.leps[0][0][0] * lu), (T)(fc.leps[0][0][1] * lu),
                   (T)(fc.leps[0][1][1] * lu), (T)(fc.leps[1][1][1] * lu)};
  const float skin = (float)fc.skin;
  uint64_t em = 0;
  FAcc<MIXED> fb;
  T ub = 0, et = 0;
  T W[9];
#pragma unroll
  for (int w = 0; w < 9; ++w) W[w] = 0;
  if (live && tb < 3) {
    const int na = cnt_leg[b];
    const int cs = min(cnt_sw[b], Wsw);
    const int *rowb = nbr_sw + (size_t)b * Wsw;
    const uint64_t amb = amask[b];
    for (int n = 0; n < na; ++n) {
      const int kc = nth_bit(amb, n);  // b's list slot of c
      const int c = rowb[kc];
      const float dc = disp[c];
      const double4 xc = X[c];
      const int tc = type_of(meta_of(xc));
      Leg<T> Lbc;  // b's leg to c: unit vector from c to b
      make_leg<MIXED>(xb, xc, in, fc, tb, tc, (uint32_t)c, (uint32_t)kc, Lbc);
      const int *rowc = nbr_sw + (size_t)c * Wsw;
      const int cc = min(cnt_sw[c], Wsw);
      for (int k = 0; k < cs; ++k) {
        const int a = rowb[k];
        if (a == c || !(disp[a] + dc > skin)) continue;
        const double4 xa = X[a];
        const int ta = type_of(meta_of(xa));
        // leg c -> a: d = x_c - x_a (unit vector from a to the centre c)
        double d0, d1, d2;
        min_image_fast(xc, xa, in, d0, d1, d2);
        const double isg = fc.isig[(tc == 2) ? 1 : 0][(ta == 2) ? 1 : 0];
        Leg<T> Lca;
        Lca.elo = 0u;
        if (!MIXED) {
          const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
          const double rs = r * isg;
          if (!(rs < fc.cut_sw)) continue;
          bool listed = false;
          for (int q = 0; q < cc; ++q) listed |= (rowc[q] == a);
          if (listed) continue;
          const double rinv = 1.0 / r, inv_ra = 1.0 / (rs - fc.al);
          Lca.e0 = d0 * rinv; Lca.e1 = d1 * rinv; Lca.e2 = d2 * rinv;
          Lca.rinv = rinv;
          Lca.E = exp(fc.gamma * inv_ra);
          Lca.g = -fc.gamma * inv_ra * inv_ra * isg;
        } else {
          // the leg values from FP64 geometry, as make_leg forms them
          const double s0 = d0 * isg, s1 = d1 * isg, s2 = d2 * isg;
          const double rr = s0 * s0 + s1 * s1 + s2 * s2;
          if (!(rr < fc.cut_sw2)) continue;
          const double rs = sqrt(rr);
          if (!(rs < fc.mixed_xleg)) continue;
          bool listed = false;
          for (int q = 0; q < cc; ++q) listed |= (rowc[q] == a);
          if (listed) continue;
          const double rho = fc.rho[(tc == 2) ? 1 : 0][(ta == 2) ? 1 : 0];
          const double rinv = 1.0 / rs;
          const double inv_ra = 1.0 / (rs - fc.al);
          const double x0 = s0 * rinv, x1 = s1 * rinv, x2 = s2 * rinv;
          Lca.e0 = (float)x0; Lca.e1 = (float)x1; Lca.e2 = (float)x2;
          Lca.elo = pack_lo(x0, Lca.e0, 0) | pack_lo(x1, Lca.e1, 1) | pack_lo(x2, Lca.e2, 2);
          Lca.rinv = (float)(rinv * rho);
          Lca.E = expf((float)(fc.gamma * inv_ra));
          Lca.g = (float)(-fc.gamma * inv_ra * inv_ra * rho);
        }
        // leg c -> b is b's leg to c, reversed
        Leg<T> Lcb = Lbc;
        Lcb.e0 = -Lbc.e0; Lcb.e1 = -Lbc.e1; Lcb.e2 = -Lbc.e2;
        Lcb.elo = neg_lo(Lbc.elo);
        const int nge = (tc == 2) + (ta == 2) + (tb == 2);
        T Fa[3], Fb[3], hh;
        // vterm gives the force on its first leg: once per end
        vterm<T>(pick4(lp, nge), Lca, Lcb, Fa, hh);
        vterm<T>(pick4(lp, nge), Lcb, Lca, Fb, hh);
        const T Fc0 = -(Fa[0] + Fb[0]), Fc1 = -(Fa[1] + Fb[1]), Fc2 = -(Fa[2] + Fb[2]);
        fb.add(Fb[0], Fb[1], Fb[2]);
        const T third = hh * T(1.0 / 3.0);
        // forces on a (slot k) and c (slot kc) for them to gather
        const uint64_t bits[2] = {1ull << k, 1ull << kc};
        const int slots[2] = {k, kc};
        const T fx[2] = {Fa[0], Fc0}, fy[2] = {Fa[1], Fc1}, fz[2] = {Fa[2], Fc2};
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const size_t at = (size_t)slots[e] * NA + b;
          T4 g;
          if (em & bits[e]) {
            g = EX[at];
          } else {
            g.x = g.y = g.z = g.w = T(0);
          }
          g.x += fx[e]; g.y += fy[e]; g.z += fz[e]; g.w += third;
          EX[at] = g;
          em |= bits[e];
        }
        ub += third;
        et += hh;
        if (VIR) {
          // relative to the centre: r_a = x_a - x_c = -e_ca / rinv_ca, same for b
          const T ra = -T(1) / Lca.rinv, rb = -T(1) / Lcb.rinv;
          const T a0 = ra * Lca.e0, a1 = ra * Lca.e1, a2 = ra * Lca.e2;
          const T b0 = rb * Lcb.e0, b1 = rb * Lcb.e1, b2 = rb * Lcb.e2;
          W[0] += a0 * Fa[0] + b0 * Fb[0]; W[1] += a0 * Fa[1] + b0 * Fb[1];
          W[2] += a0 * Fa[2] + b0 * Fb[2]; W[3] += a1 * Fa[0] + b1 * Fb[0];
          W[4] += a1 * Fa[1] + b1 * Fb[1]; W[5] += a1 * Fa[2] + b1 * Fb[2];
          W[6] += a2 * Fa[0] + b2 * Fb[0]; W[7] += a2 * Fa[1] + b2 * Fb[1];
          W[8] += a2 * Fa[2] + b2 * Fb[2];
        }
      }
    }
  }
  exmask[b] = em;
  {
    double f0, f1, f2;
    fb.get(1.0, f0, f1, f2);
    T4 g;
    g.x = (T)f0; g.y = (T)f1; g.z = (T)f2; g.w = ub;
    EXself[b] = g;
  }
  if (em) {
    // rare: only after atoms outran the list skin (global atomic, vector memory)
    __hip_atomic_fetch_or(&ctl[kin].extra_any, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (EN || VIR) {
    v[0] = (double)et * escale;
    const double ws = MIXED ? fscale * lscale : 1.0;
#pragma unroll
    for (int w = 0; w < 9; ++w) v[1 + w] = (double)W[w] * ws;
    block_partials<10, FIB_BLOCK>(v, op, lb, pstride, partials);
  }
}

// After a list build: positions of the build, displacement bookkeeping reset.
__global__ void k_mark_build(int NA, int ninst, int Wsw, int Wion, const double4 *__restrict__ X,
                             double4 *__restrict__ Xb, float *__restrict__ disp,
                             DCtl *__restrict__ ctl, const int *__restrict__ binst,
                             const int *__restrict__ gate, int *__restrict__ halt, double inc) {
  GATE_OR_RETURN(gate);
  GRID_STRIDE(i, max(NA, ninst)) {
  // NA is a multiple of FIB_BLOCK: the lanes of a block are all below it or all above
  if (i < NA && !block_gated_out(gate, binst, i)) {
    Xb[i] = X[i];
    disp[i] = 0.0f;
  }
  if (i < ninst && !gated_out(gate, i)) {
    ctl[i].maxdisp = 0.0;
    ctl[i].maxdisp2 = 0.0;
    ctl[i].nmaxdisp = 0.0;
    ctl[i].extra_any = 0;
    // a list longer than its width was cut: sticky, checked by the host;
    // in MD-step mode the rest of the step halts until the host has grown
    // the lists and rebuilt them (inc = 0 then: the same build once more)
    if (ctl[i].maxcnt_sw > Wsw || ctl[i].maxcnt_ion > Wion) {
      ctl[i].overflow = 1;
      if (halt) *halt = 1;
    }
    ctl[i].builds += inc;
  }
  }
}


// ---------------------------------------------------------------------------
// MD-step mode: velocity Verlet with a per-instance adaptive timestep
// dt = min(dt_max, dx_max / |v|max), all step control on the device, so one
// step is a fixed sequence of launches that can be captured in a graph:
//   k_md_pos     Berendsen shells, half kick, drift, pinned atoms, removal
//                above the surface, displacement since the list build and
//                the exact two largest per instance;
//                k_md_decide sets the rebuild gate (two largest displacements
//                add up to more than the skin)
//   gated build  bin / sort in place / k_nbuild (skipped on the
//                device when the gate is 0)
//   forces       k_legs, k_tri, k_extra, k_force3 MODE 2 (second half kick,
//                |v|max, shell sums)
//   k_md_ctl     time += dt, next dt, next shell factor
__device__ __forceinline__ bool in_shell(double x, double y, double z, const DInst &in, double w) {
  return x < w || x > in.lb[0] - w || y < w || y > in.lb[1] - w || z < in.bot_z + w;
}

__global__ __launch_bounds__(FIB_BLOCK) void k_md_pos(double4 *__restrict__ X,
                                                      const DInst *__restrict__ I, int NA,
                                                      double *__restrict__ V,
                                                      const double *__restrict__ F,
                                                      const double *__restrict__ Xi, StepArgs sa,
                                                      int pstride, double *__restrict__ partials,
                                                      const DCtl *__restrict__ ctl,
                                                      const double4 *__restrict__ Xb,
                                                      float *__restrict__ disp,
                                                      int *__restrict__ leave,
                                                      const int *__restrict__ halt) {
  HALT_OR_RETURN(halt);
  constexpr bool NT = (FIB_NT_LD & 8) != 0, NTS = (FIB_NT_ST & 8) != 0;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int i = lb * FIB_BLOCK + threadIdx.x;
  double4 x = X[i];
  uint32_t mi = meta_of(x);
  const int kin = __builtin_amdgcn_readfirstlane(inst_of(mi));
  const DInst &in = I[kin];
  const bool live = !(mi & (META_PAD | META_DEAD));
  const double dt = ctl[kin].cfg_dt, lam = ctl[kin].cfg_lambda;
  // the two largest displacements since the build (largest, how many share
  // it, largest below it), energy taken by the shells, energy of removed
  // atoms; an atom out of the simulation counts as displaced by 0
  double red[5] = {0.0, 1.0, -1.0e300, 0.0, 0.0};
  if (live) {
    const int t = type_of(mi);
    const double m = sa.mass[t - 1];
    double v[3], xx[3] = {x.x, x.y, x.z};
    for (int c = 0; c < 3; ++c) v[c] = ld_nt<NT>(V + (size_t)c * NA + i);
    if (sa.shell_w > 0.0 && t < 3 && !(mi & META_FROZEN) &&
        in_shell(xx[0], xx[1], xx[2], in, sa.shell_w)) {
      red[3] = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * (1.0 - lam * lam);
      for (int c = 0; c < 3; ++c) v[c] *= lam;
    }
    for (int c = 0; c < 3; ++c) {
      v[c] += 0.5 * ld_nt<NT>(F + (size_t)c * NA + i) * dt / m;
      xx[c] += v[c] * dt;
    }
    if (mi & META_FROZEN) {
      for (int c = 0; c < 3; ++c) {
        xx[c] = Xi[(size_t)c * NA + i];
        v[c] = 0.0;
      }
    }
    for (int c = 0; c < 3; ++c) xx[c] = fmod(xx[c] + 2. * in.lb[c], in.lb[c]);
    if (xx[2] - in.surf_z > sa.remove_dz && xx[2] < in.lb[2] - sa.remove_dz) {
      // sputtered (or backscattered): out of the simulation from now on,
      // with its kinetic energy here; its interaction energy with the atoms
      // that stay (an atom of a sputtered cluster is still bound) is booked
      // by k_md_decide from the removal list
      mi |= META_DEAD | META_LEAVING;
      red[4] = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int c = 0; c < 3; ++c) v[c] = 0.0;
      x.w = __longlong_as_double((long long)((wbits(x) & ~0xffffffffull) | mi));
      const int slot = atomicAdd(&leave[0], 1);
      if (slot < FIB_LEAVE_CAP) leave[1 + slot] = i;
    }
    x.x = xx[0]; x.y = xx[1]; x.z = xx[2];
    X[i] = x;
    for (int c = 0; c < 3; ++c) st_nt<NTS>(V + (size_t)c * NA + i, v[c]);
    double e0, e1, e2;
    min_image_fast(x, ld_nt<NT>(Xb + i), in, e0, e1, e2);
    const float dd = (float)sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    st_nt<NTS>(disp + i, dd);
    red[0] = (double)dd;
  } else {
    st_nt<NTS>(disp + i, 0.0f);
  }
  const int op[5] = {RED_TOP2, RED_TOP2_CNT, RED_TOP2_SND, RED_SUM, RED_SUM};
  block_partials<5, FIB_BLOCK>(red, op, lb, pstride, partials);
}

// Rebuild gates of the step: each instance whose two largest displacements
// since the build add up to more than the skin (only then can a pair the
// build left out have come within the cutoff), and whether there is any.
// Interaction energy (FP64, J) of atom i with the atoms that stay in the
// simulation: every SW pair and three-body term and every Moliere pair it
// takes part in (si.f90:278-368 forms), each term divided by the number of
// its atoms that leave in the current step, so that a term is booked once.
// The lists are those of the last build, which hold every pair in range.
// Used for atoms removed in a step and for ions where they are fired.
__device__ double interaction_energy(int i, const double4 *__restrict__ X,
                                 const DInst *__restrict__ I, const FConst &fc, int Wsw,
                                 int Wion, const int *__restrict__ cnt_sw,
                                 const int *__restrict__ nbr_sw, const int *__restrict__ cnt_ion,
                                 const int *__restrict__ nbr_ion) {
  // the parameters in scalars: indexing the parameter block by species (or
  // taking its address in a lambda) would copy it to scratch memory
  const double isg00 = fc.isig[0][0], isg01 = fc.isig[0][1], isg11 = fc.isig[1][1];
  const double eps00 = fc.eps[0][0], eps01 = fc.eps[0][1], eps11 = fc.eps[1][1];
  const double lp0 = fc.leps[0][0][0], lp1 = fc.leps[0][0][1], lp2 = fc.leps[0][1][1],
               lp3 = fc.leps[1][1][1];
  const double swA = fc.A, swB = fc.B, al = fc.al, gam = fc.gamma, cut = fc.cut_sw;
  const double z1 = fc.zmol[0], z2 = fc.zmol[1], z3 = fc.zmol[2];
  const double a13 = fc.aamol[0][2], a23 = fc.aamol[1][2], a33 = fc.aamol[2][2];
  const double c0 = fc.molc[0], c1 = fc.molc[1], c2 = fc.molc[2];
  const double b0 = fc.molb[0], b1 = fc.molb[1], b2 = fc.molb[2];
  const double ie0 = fc.ie0, ec = fc.ec, rlj2 = fc.rlj2;
  auto gone = [](uint32_t m) { return (m & META_DEAD) && !(m & META_LEAVING); };
  auto leaves = [](uint32_t m) { return (m & META_LEAVING) ? 1 : 0; };
  auto isg = [=](int ta, int tb) {
    const int g = (ta == 2) + (tb == 2);
    return g == 0 ? isg00 : g == 1 ? isg01 : isg11;
  };
  auto legE = [=](int ta, int tb, double r) {  // exp(gamma / (r/sigma_ab - a)), 0 out of range
    const double rs = r * isg(ta, tb);
    return rs < cut ? exp(gam / (rs - al)) : 0.0;
  };
  auto leps = [=](int ta, int tb, int tc) {  // by the number of Ge atoms of the term
    const int g = (ta == 2) + (tb == 2) + (tc == 2);
    return g == 0 ? lp0 : g == 1 ? lp1 : g == 2 ? lp2 : lp3;
  };
  const double4 xi = X[i];
  const uint32_t mi = meta_of(xi);
  const DInst &in = I[inst_of(mi)];
  const int ti = type_of(mi);
  double e = 0.0;
  if (ti < 3) {
    const int ci = min(cnt_sw[i], Wsw);
    const int *ri = nbr_sw + (size_t)i * Wsw;
    for (int a = 0; a < ci; ++a) {
      const int j = ri[a];
      const double4 xj = X[j];
      const uint32_t mj = meta_of(xj);
      if (gone(mj)) continue;
      const int tj = type_of(mj);
      double d0, d1, d2;  // x_i - x_j
      min_image_fast(xi, xj, in, d0, d1, d2);
      const double rij = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      const double rs = rij * isg(ti, tj);
      if (!(rs < cut)) continue;
      const double rs2 = rs * rs;
      const int g = (ti == 2) + (tj == 2);
      const double eij = g == 0 ? eps00 : g == 1 ? eps01 : eps11;
      const double phi = eij * swA * (swB / (rs2 * rs2) - 1.0) * exp(1.0 / (rs - al));
      e += phi / (1 + leaves(mj));
      const double Eij = legE(ti, tj, rij);
      // terms centred on j with legs to i and to another k of j's list
      const int cj = min(cnt_sw[j], Wsw);
      const int *rj = nbr_sw + (size_t)j * Wsw;
      for (int b = 0; b < cj; ++b) {
        const int k = rj[b];
        if (k == i) continue;
        const double4 xk = X[k];
        const uint32_t mk = meta_of(xk);
        if (gone(mk)) continue;
        const int tk = type_of(mk);
        double f0, f1, f2;  // x_k - x_j
        min_image_fast(xk, xj, in, f0, f1, f2);
        const double rjk = sqrt(f0 * f0 + f1 * f1 + f2 * f2);
        const double Ejk = legE(tj, tk, rjk);
        if (Ejk == 0.0) continue;
        const double cs = (d0 * f0 + d1 * f1 + d2 * f2) / (rij * rjk);
        const double D = cs + 1.0 / 3.0;
        e += leps(ti, tj, tk) * Eij * Ejk * D * D / (1 + leaves(mj) + leaves(mk));
      }
      // terms centred on i with legs to j and to a later k of i's list
      for (int b = a + 1; b < ci; ++b) {
        const int k = ri[b];
        const double4 xk = X[k];
        const uint32_t mk = meta_of(xk);
        if (gone(mk)) continue;
        const int tk = type_of(mk);
        double f0, f1, f2;  // x_i - x_k
        min_image_fast(xi, xk, in, f0, f1, f2);
        const double rik = sqrt(f0 * f0 + f1 * f1 + f2 * f2);
        const double Eik = legE(ti, tk, rik);
        if (Eik == 0.0) continue;
        const double cs = (d0 * f0 + d1 * f1 + d2 * f2) / (rij * rik);
        const double D = cs + 1.0 / 3.0;
        e += leps(ti, tj, tk) * Eij * Eik * D * D / (1 + leaves(mj) + leaves(mk));
      }
    }
  }
  // Moliere pairs (si.f90:426-446): both atoms fired, r < 2.5 sigma; one of
  // the two is Ga (the ion list holds no Si-Si or Si-Ge pair)
  const bool ifired = !(ti == 3 && orig_of(xi) >= in.nsg + in.impact);
  const int cn = min(cnt_ion[i], Wion);
  const int *rn = nbr_ion + (size_t)i * Wion;
  for (int a = 0; ifired && a < cn; ++a) {
    const int j = rn[a];
    const double4 xj = X[j];
    const uint32_t mj = meta_of(xj);
    if (gone(mj)) continue;
    const int tj = type_of(mj);
    if (tj == 3 && orig_of(xj) >= in.nsg + in.impact) continue;
    double d0, d1, d2;
    min_image_fast(xi, xj, in, d0, d1, d2);
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (!(r2 < rlj2)) continue;
    const double r = sqrt(r2);
    const int other = ti == 3 ? tj : ti;  // the partner of the Ga atom
    const double aa = other == 1 ? a13 : other == 2 ? a23 : a33;
    const double zi = ti == 1 ? z1 : ti == 2 ? z2 : z3, zj = tj == 1 ? z1 : tj == 2 ? z2 : z3;
    const double zz = ie0 * zi * zj * ec * ec;
    const double x = r / aa;
    const double sa = c0 * exp(-b0 * x) + c1 * exp(-b1 * x) + c2 * exp(-b2 * x);
    e += zz / r * sa / (1 + leaves(mj));
  }
  return e;
}

#define DECIDE_T 256  // threads of k_md_decide (a small block keeps the removal path in registers)
static_assert(FIB_LEAVE_CAP == DECIDE_T, "one thread per listed removal");
__global__ __launch_bounds__(DECIDE_T) void k_md_decide(
    int ninst, const DInst *__restrict__ I, DCtl *__restrict__ ctl, double skin,
    int *__restrict__ gate, int *__restrict__ leave, double4 *__restrict__ X, const FConst fc,
    int Wsw, int Wion, const int *__restrict__ cnt_sw, const int *__restrict__ nbr_sw,
    const int *__restrict__ cnt_ion, const int *__restrict__ nbr_ion,
    const int *__restrict__ halt) {
  // halted: no build in this step either (the host restores the flag)
  if (halt && *halt) {
    if (threadIdx.x == 0) gate[0] = 0;
    return;
  }
  __shared__ int s_w[DECIDE_T / FIB_WAVE];
  __shared__ int s_any;
  __shared__ int s_list[FIB_LEAVE_CAP];
  __shared__ double s_e[FIB_LEAVE_CAP];
  const int t = threadIdx.x, lane = t % FIB_WAVE, wv = t / FIB_WAVE;
  // atoms removed in this step: their interaction energy with the atoms that
  // stay goes to the energy carried out, summed in atom order (the list is
  // filled in arbitrary order and sorted here), so the ledger is the same
  // for an instance alone or in any batch
  const int nleave = leave[0];
  if (nleave > 0) {
    const int n = min(nleave, FIB_LEAVE_CAP);
    s_list[t] = t < n ? leave[1 + t] : 0x7fffffff;
    __syncthreads();
    for (int k = 2; k <= FIB_LEAVE_CAP; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        const int u = t ^ j;
        if (u > t) {
          const int a = s_list[t], b = s_list[u];
          if ((a > b) == ((t & k) == 0)) { s_list[t] = b; s_list[u] = a; }
        }
        __syncthreads();
      }
    if (t < n)
      s_e[t] = interaction_energy(s_list[t], X, I, fc, Wsw, Wion, cnt_sw, nbr_sw, cnt_ion,
                                  nbr_ion);
    __syncthreads();
    if (t == 0) {
      for (int q = 0; q < n; ++q) ctl[inst_of(meta_of(X[s_list[q]]))].e_out_step += s_e[q];
      if (nleave > FIB_LEAVE_CAP) ctl[0].leave_overflow = 1;
      leave[0] = 0;
    }
    __syncthreads();
    if (t < n) {
      double4 &x = X[s_list[t]];
      x.w = __longlong_as_double((long long)(wbits(x) & ~(uint64_t)META_LEAVING));
    }
  }
  // one workgroup; behind the gates, the prefix over instances of the real
  // cells of the gated ones (the cell numbering of a gated k_nbuild)
  int *cplan = gate + 1 + ninst;
  if (t == 0) s_any = 0;
  __syncthreads();
  int carry = 0;
  for (int k0 = 0; k0 < ninst; k0 += DECIDE_T) {
    const int k = k0 + t;
    int g = 0, w = 0;
    if (k < ninst) {
      const DCtl &c = ctl[k];
      const double m2 = c.nmaxdisp >= 2.0 ? c.maxdisp : fmax(c.maxdisp2, 0.0);
      g = (c.maxdisp + m2 > skin) ? 1 : 0;
      ctl[k].maxdisp2 = m2;  // the effective second maximum, for k_extra
      gate[1 + k] = g;
      if (g) w = I[k].nc[0] * I[k].nc[1] * I[k].nc[2];
    }
    // inclusive scan of w: in the wave, then over the waves' totals
    int sc = w;
    for (int o = 1; o < FIB_WAVE; o <<= 1) {
      const int y = __shfl_up(sc, o);
      if (lane >= o) sc += y;
    }
    if (lane == FIB_WAVE - 1) s_w[wv] = sc;
    if (__ballot(g) != 0ull && lane == 0) s_any = 1;
    __syncthreads();
    int before = carry;
    for (int v = 0; v < wv; ++v) before += s_w[v];
    if (k < ninst) cplan[k] = before + sc - w;
    for (int v = 0; v < DECIDE_T / FIB_WAVE; ++v) carry += s_w[v];
    __syncthreads();  // s_w is rewritten by the next chunk
  }
  if (t == 0) {
    cplan[ninst] = carry;
    gate[0] = s_any;
  }
}

// A fired ion brings its interaction energy with the atoms near its start
// point in with its kinetic energy (k_set_atom books the latter); q is its
// unsorted index atom_off + orig, evaluated after the lists were rebuilt.
__global__ void k_fire_energy(int q, const int *__restrict__ inv, const double4 *__restrict__ X,
                              const DInst *__restrict__ I, const FConst fc, int Wsw, int Wion,
                              const int *__restrict__ cnt_sw, const int *__restrict__ nbr_sw,
                              const int *__restrict__ cnt_ion, const int *__restrict__ nbr_ion,
                              DCtl *__restrict__ ctl) {
  if (threadIdx.x != 0) return;
  const int p = inv[q];
  const double e = interaction_energy(p, X, I, fc, Wsw, Wion, cnt_sw, nbr_sw, cnt_ion, nbr_ion);
  ctl[inst_of(meta_of(X[p]))].e_in += e;
}

// After the step's second half kick: advance time, choose the next dt and
// the next Berendsen factor of the shells.
__global__ void k_md_ctl(int ninst, DCtl *__restrict__ ctl, StepArgs sa,
                         const int *__restrict__ halt) {
  HALT_OR_RETURN(halt);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ninst) return;
  DCtl &c = ctl[k];
  c.cfg_time += c.cfg_dt;
  c.cfg_steps += 1.0;
  c.cfg_dt_min = fmin(c.cfg_dt_min, c.cfg_dt);
  c.e_bath += c.e_bath_step;
  c.e_out += c.e_out_step;
  const double dt = c.vmax > 0.0 ? fmin(sa.dt_max, sa.dx_max / c.vmax) : sa.dt_max;
  c.cfg_dt = dt;
  double lam = 1.0;
  if (sa.shell_w > 0.0 && c.n_shell > 0.0 && c.ke_shell > 0.0) {
    const double ts = c.ke_shell / (3.0 * c.n_shell * sa.kB);
    lam = sqrt(1.0 + dt / sa.tau * (sa.ttar / ts - 1.0));
  }
  c.cfg_lambda = lam;
}

// dt from the current |v|max (DCtl::vmax, from k_sums); reset also zeroes the
// clock and the shell factor.
__global__ void k_md_init(int ninst, DCtl *__restrict__ ctl, StepArgs sa, int reset) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ninst) return;
  DCtl &c = ctl[k];
  c.cfg_dt = c.vmax > 0.0 ? fmin(sa.dt_max, sa.dx_max / c.vmax) : sa.dt_max;
  if (reset) {
    c.cfg_time = 0.0;
    c.cfg_steps = 0.0;
    c.cfg_lambda = 1.0;
    c.cfg_dt_min = 1.0e300;
    c.e_bath = c.e_out = c.e_in = 0.0;
    c.leave_overflow = 0;
  }
}

// Cascade outputs: removed SW atoms (sputter yield), removed ions, displaced
// SW atoms (farther than disp_r from their initial site Xi), height of the
// instance's most recently fired ion.
__global__ __launch_bounds__(FIB_BLOCK) void k_md_report(const double4 *__restrict__ X,
                                                         const DInst *__restrict__ I, int NA,
                                                         const double *__restrict__ Xi,
                                                         StepArgs sa, int pstride,
                                                         double *__restrict__ partials) {
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int i = lb * FIB_BLOCK + threadIdx.x;
  const double4 x = X[i];
  const uint32_t mi = meta_of(x);
  const int kin = __builtin_amdgcn_readfirstlane(inst_of(mi));
  const DInst &in = I[kin];
  double v[4] = {0.0, 0.0, 0.0, -1.0e300};
  if (!(mi & META_PAD)) {
    const int t = type_of(mi);
    if (mi & META_DEAD) {
      if (t < 3) v[0] = 1.0; else v[1] = 1.0;
    } else if (t < 3) {
      double e0, e1, e2;
      double4 xi;
      xi.x = Xi[i]; xi.y = Xi[(size_t)NA + i]; xi.z = Xi[2 * (size_t)NA + i]; xi.w = 0.0;
      min_image_fast(x, xi, in, e0, e1, e2);
      if (e0 * e0 + e1 * e1 + e2 * e2 > sa.disp_r * sa.disp_r) v[2] = 1.0;
    }
    if (t == 3 && orig_of(x) == in.nsg + in.impact - 1) v[3] = (mi & META_DEAD) ? 1.0e300 : x.z;
  }
  const int op[4] = {0, 0, 0, 1};
  block_partials<4, FIB_BLOCK>(v, op, lb, pstride, partials);
}

// ---------------------------------------------------------------------------
// force pass 3: own terms + the G entries neighbours wrote for this atom;
// MODE 0 = store forces, MODE 1 = reference velocity update
// (timeint.f90:231-236) + thermostat sums; MODE 2 = MD-step mode.
// partner lookups in flight per atom in k_force3: FP64 8 (most atoms have at
// most 8 legs; 6 % faster on the dose batch than 4), mixed 4 (8 was 4 %
// slower: the registers cost occupancy)
#ifndef F3_BATCH_FP64
#define F3_BATCH_FP64 8
#endif
#ifndef F3_BATCH_MIXED
#define F3_BATCH_MIXED 4
#endif
template <bool MIXED, bool EN, int MODE>
__global__ __launch_bounds__(FIB_BLOCK) void k_force3(
    const double4 *__restrict__ X, const DInst *__restrict__ I, int NA, int Wsw,
    const int *__restrict__ nbr_sw, const uint16_t *__restrict__ rk_sw,
    const uint8_t *__restrict__ off_sw, const typename Prec<MIXED>::T4 *__restrict__ G,
    const uint32_t *__restrict__ RA,
    const int *__restrict__ cnt_leg, const int *__restrict__ cnt_sw,
    const double *__restrict__ Fpart, const double *__restrict__ upart, double fscale,
    double escale, double *__restrict__ Fout, double *__restrict__ uout, double *__restrict__ V,
    StepArgs sa, int pstride, double *__restrict__ partials, const DCtl *__restrict__ ctl,
    const typename Prec<MIXED>::T4 *__restrict__ EX,
    const typename Prec<MIXED>::T4 *__restrict__ EXself, const uint64_t *__restrict__ exmask,
    const int *__restrict__ halt) {
  HALT_OR_RETURN(halt);
  typedef typename Prec<MIXED>::T4 T4;
  constexpr bool NT = (FIB_NT_LD & 4) != 0, NTS = (FIB_NT_ST & 4) != 0;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int i = lb * FIB_BLOCK + threadIdx.x;
  const double4 xi = X[i];
  const uint32_t mi = meta_of(xi);
  const int kin = __builtin_amdgcn_readfirstlane(inst_of(mi));
  const DInst &in = I[kin];
  const bool live = !(mi & (META_PAD | META_DEAD));
  double f0 = 0, f1 = 0, f2 = 0, u = 0;
  if (live) {
    f0 = ld_nt<NT>(Fpart + i);
    f1 = ld_nt<NT>(Fpart + (size_t)NA + i);
    f2 = ld_nt<NT>(Fpart + 2 * (size_t)NA + i);
    if (EN) u = upart[i];
    const size_t row = (size_t)i * Wsw;
    double g0 = 0, g1 = 0, g2 = 0, gu = 0;
    // reaction on the centre: minus the forces its own terms put on its legs
    const int na = ld_nt<NT>(cnt_leg + i);
    for (int n = 0; n < na; ++n) {
      const T4 g = G[(size_t)n * NA + i];
      g0 -= (double)g.x; g1 -= (double)g.y; g2 -= (double)g.z;
      if (EN) gu += 0.5 * (double)g.w;
    }
    // forces the partners' terms put on this atom, at the places k_tri
    // noted for its legs: one lookup per leg, F3_BATCH in flight
    constexpr int F3_BATCH = MIXED ? F3_BATCH_MIXED : F3_BATCH_FP64;
    for (int n0 = 0; n0 < na; n0 += F3_BATCH) {
      uint32_t ra[F3_BATCH];
      T4 g4[F3_BATCH];
#pragma unroll
      for (int u = 0; u < F3_BATCH; ++u)
        ra[u] = (n0 + u < na) ? ld_nt<NT>(RA + (size_t)(n0 + u) * NA + i) : ~0u;
#pragma unroll
      for (int u = 0; u < F3_BATCH; ++u) {
        if (ra[u] != ~0u) {
          g4[u] = G[ra[u]];
        } else {
          g4[u].x = g4[u].y = g4[u].z = g4[u].w = 0;
        }
      }
#pragma unroll
      for (int u = 0; u < F3_BATCH; ++u)
        if (n0 + u < na) {
          g0 += (double)g4[u].x; g1 += (double)g4[u].y; g2 += (double)g4[u].z;
          if (EN) gu += (double)g4[u].w;
        }
    }
    // vertex terms outside the centre's list (k_extra), when there are any
    if (ctl[kin].extra_any) {
      const T4 gs = EXself[i];
      g0 += (double)gs.x; g1 += (double)gs.y; g2 += (double)gs.z;
      if (EN) gu += (double)gs.w;
      const int cs = min(cnt_sw[i], Wsw);
      for (int k = 0; k < cs; ++k) {
        const int j = nbr_sw[row + k];
        const uint64_t mj = exmask[j];
        if (!mj) continue;
        const int sr = rk_sw[row + k];
        const int r = (sr & RK_SLOT) ? (sr & ~RK_SLOT)
                                     : off_sw[(size_t)j * OFF_W + 26 - (sr >> 8)] + (sr & 0xff);
        if (r < Wsw && ((mj >> r) & 1ull)) {
          const T4 g = EX[(size_t)r * NA + j];
          g0 += (double)g.x; g1 += (double)g.y; g2 += (double)g.z;
          if (EN) gu += (double)g.w;
        }
      }
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
  if (MODE == 2) {
    // MD-step mode: second half kick with the instance's dt; F = F'
    // reductions: |v|max, kinetic energy, shell sum m v^2 and count
    double v[4] = {0, 0, 0, 0};
    if (live) {
      const int t = type_of(mi);
      const double m = sa.mass[t - 1];
      const double dt = ctl[kin].cfg_dt;
      const bool pinned = (mi & META_FROZEN) != 0;
      const double fn[3] = {f0, f1, f2};
      double vv[3];
      for (int c = 0; c < 3; ++c) {
        const size_t a = (size_t)c * NA + i;
        vv[c] = pinned ? 0.0 : ld_nt<NT>(V + a) + 0.5 * fn[c] * dt / m;
        st_nt<NTS>(V + a, vv[c]);
        st_nt<NTS>(Fout + a, fn[c]);
      }
      const double v2 = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2];
      v[0] = sqrt(v2);
      v[1] = 0.5 * m * v2;
      if (sa.shell_w > 0.0 && t < 3 && !pinned && in_shell(xi.x, xi.y, xi.z, in, sa.shell_w)) {
        v[2] = m * v2;
        v[3] = 1.0;
      }
    }
    const int op[4] = {1, 0, 0, 0};
    block_partials<4, FIB_BLOCK>(v, op, lb, pstride, partials);
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
                                                   DCtl *__restrict__ ctl,
                                                   const double4 *__restrict__ Xb,
                                                   float *__restrict__ disp) {
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
  // slots: thermostat sum, max KE, max displacement, and the displacement
  // again as the second maximum (k_extra tests maxdisp + maxdisp2)
  double red[4] = {0.0, -1.0, 0.0, 0.0};
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
    {
      // displacement since the list build (minimum image across the rebox)
      double e0, e1, e2;
      min_image_fast(x, Xb[i], in, e0, e1, e2);
      const double dd = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
      disp[i] = (float)dd;
      red[2] = dd;
      red[3] = dd;
    }
    for (int c = 0; c < 3; ++c) {
      V[(size_t)c * NA + i] = v[c];
      F[(size_t)c * NA + i] = f[c];
    }
  }
  const int op[4] = {0, 1, 1, 1};
  block_partials<4, FIB_BLOCK>(red, op, lb, pstride, partials);
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

// Places atom p (a fired ion) and books its change of kinetic energy as
// energy brought into its instance.
__global__ void k_set_atom(int NA, double4 *__restrict__ X, double *__restrict__ V, int p,
                           double x0, double x1, double x2, double v0, double v1, double v2,
                           StepArgs sa, DCtl *__restrict__ ctl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double4 x = X[p];
  const uint32_t m = meta_of(x);
  const double w0 = V[p], w1 = V[(size_t)NA + p], w2 = V[2 * (size_t)NA + p];
  ctl[inst_of(m)].e_in += 0.5 * sa.mass[type_of(m) - 1] *
                          ((v0 * v0 + v1 * v1 + v2 * v2) - (w0 * w0 + w1 * w1 + w2 * w2));
  x.x = x0; x.y = x1; x.z = x2;
  X[p] = x;
  V[p] = v0; V[(size_t)NA + p] = v1; V[2 * (size_t)NA + p] = v2;
}

// ---------------------------------------------------------------------------
// launchers
static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }
// grid of a GRID_STRIDE kernel over n elements
static inline int gcap(long n) { return std::max(1, std::min(nblk(n, FIB_BLOCK), GRID_CAP)); }

// Times the launches issued in its scope under one key.
struct KScope {
  const KTimer *t;
  int id, key;
  KScope(const DevBufs &d, int k) : t(d.kt), id(-1), key(k) {
    if (t && t->begin) id = t->begin(t->ctx, k);
  }
  ~KScope() {
    if (id >= 0) t->end(t->ctx, id, key);
  }
};

template <int NV>
static void fin_launch(const DevBufs &d, const RedSpec &rs, int nch, const int *gate,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_finalize<NV>, dim3(d.ninst, nch), dim3(FIN_T), 0, s, d.I, d.partials,
                     d.pstride, rs, d.ctl, gate, d.fin_tmp, d.halt);
  if (nch > 1)
    hipLaunchKernelGGL(k_finalize2<NV>, dim3(d.ninst), dim3(FIB_WAVE), 0, s, d.I, nch, rs, d.ctl,
                       gate, d.fin_tmp, d.halt);
}

static void finalize(const DevBufs &d, int nv, int bs, const int *op, const int *dst,
                     hipStream_t s, int nreg = 1, const int *gate = nullptr) {
  RedSpec rs;
  rs.nv = nv;
  rs.bs = bs;
  rs.nreg = nreg;
  for (int k = 0; k < FIB_NRED; ++k) {
    rs.op[k] = k < nv ? op[k] : 0;
    rs.dst[k] = k < nv ? dst[k] : -1;
  }
  // chunks per instance: one workgroup up to FIN_T blocks of the producer
  const int nch = (d.max_pad / bs + FIN_T - 1) / FIN_T;
  KScope ks_(d, KT_REDUCE);
  switch (nv) {
    case 1: fin_launch<1>(d, rs, nch, gate, s); break;
    case 2: fin_launch<2>(d, rs, nch, gate, s); break;
    case 3: fin_launch<3>(d, rs, nch, gate, s); break;
    case 4: fin_launch<4>(d, rs, nch, gate, s); break;
    case 5: fin_launch<5>(d, rs, nch, gate, s); break;
    default: fin_launch<10>(d, rs, nch, gate, s); break;
  }
}

static void finalize_thermo(const DevBufs &d, hipStream_t s) {
  const int op[5] = {0, 0, 0, 0, 1};
  const int dst[5] = {DCTL_IDX(glob_sum), DCTL_IDX(side_sum), DCTL_IDX(side_cnt),
                      DCTL_IDX(ke_sum), DCTL_IDX(vmax)};
  finalize(d, 5, FIB_BLOCK, op, dst, s);
}

void launch_bin(const DevBufs &d, hipStream_t s) {
  {
    KScope ks_(d, KT_BIN);
    hipLaunchKernelGGL(k_zero_i32, dim3(gcap(d.nbucket)), dim3(FIB_BLOCK), 0, s, d.bcnt,
                       d.nbucket, d.gate);
  }
  {
    KScope ks_(d, KT_BIN);
    hipLaunchKernelGGL(k_bin, dim3(gcap(d.NA)), dim3(FIB_BLOCK), 0, s, d.X[d.cur], d.NA,
                       d.I, d.bkey, d.brank, d.bcnt, d.ctl, d.binst, d.gate);
  }
}

void launch_scan(const DevBufs &d, hipStream_t s) {
  const int nb = nblk(d.nbucket, SCAN_CH);
  {
    KScope ks_(d, KT_SORT);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_T), 0, s, d.bcnt, d.nbucket, d.bsum,
                       d.gate);
  }
  {
    KScope ks_(d, KT_SORT);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_T), 0, s, d.bsum, nb, d.gate);
  }
  {
    KScope ks_(d, KT_SORT);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_T), 0, s, d.bcnt, d.nbucket, d.bsum, d.I,
                       d.ninst, d.cstart, d.gate);
  }
  {
    KScope ks_(d, KT_SORT);
    hipLaunchKernelGGL(k_scan_base, dim3(1), dim3(SCAN_T), 0, s, d.cstart, d.I, d.ninst, d.ibase,
                       d.gate);
  }
  {
    KScope ks_(d, KT_SORT);
    hipLaunchKernelGGL(k_scan_fix, dim3(gcap(d.nbucket)), dim3(FIB_BLOCK), 0, s,
                       d.cstart, d.nbucket, d.I, d.ninst, d.ibase, d.gate);
  }
}

void launch_sort(const DevBufs &d, hipStream_t s, bool inplace) {
  {
    KScope ks_(d, KT_SORT);
    hipLaunchKernelGGL(k_scatter, dim3(gcap(d.NA)), dim3(FIB_BLOCK), 0, s, d.bkey,
                       d.brank, d.NA, d.cstart, d.perm, d.binst, d.gate);
  }
  {
    KScope ks_(d, KT_SORT);
    hipLaunchKernelGGL(k_fixup, dim3(gcap(d.nbucket)), dim3(FIB_BLOCK), 0, s, d.cstart,
                       d.bcnt, d.nbucket, d.perm, d.gate);
  }
  const int c = d.cur, n = 1 - d.cur;
  // the in-place sort of MD-step mode leaves the forces behind: every caller
  // (an MD step's build, the fire-time build) computes them anew before
  // anything reads them
  double *const F0 = inplace ? nullptr : d.F[c];
  {
    KScope ks_(d, KT_SORT);
    hipLaunchKernelGGL(k_reorder, dim3(gcap(d.NA)), dim3(FIB_BLOCK), 0, s, d.NA, d.perm,
                       d.I, d.X[c], d.X[n], d.V[c], d.V[n], F0, d.F[n], d.Xi[c], d.Xi[n],
                       d.inv, d.binst, d.gate);
  }
  if (inplace) {
    KScope ks_(d, KT_SORT);
    hipLaunchKernelGGL(k_copyback, dim3(gcap(d.NA)), dim3(FIB_BLOCK), 0, s, d.NA, d.X[n],
                       d.X[c], d.V[n], d.V[c], (const double *)nullptr, (double *)nullptr,
                       d.Xi[n], d.Xi[c], d.binst, d.gate);
  }
}

void launch_nbuild(const DevBufs &d, const FConst &fc, hipStream_t s) {
  {
    KScope ks_(d, KT_LISTAUX);
    hipLaunchKernelGGL(k_clear_counts, dim3(gcap(d.NA)), dim3(FIB_BLOCK), 0, s, d.NA,
                       d.cnt_sw, d.cnt_ion, d.off_sw, d.binst, d.gate);
  }
  if (d.total_cells > 0) {
    KScope ks_(d, KT_NBUILD);
    const int nbc = d.nb_cells ? d.nb_cells : NB_CELLS;
#define NB_LAUNCH(C)                                                                          \
  hipLaunchKernelGGL(k_nbuild<C>, dim3(nblk(d.total_cells, C * NB_WAVES)),                    \
                     dim3(NB_WAVES * FIB_WAVE), 0, s, d.X[d.cur], d.I, d.ninst, d.total_cells, \
                     d.cstart, d.bcnt, fc.rv2, fc.rv2lj, d.Wsw, d.Wion, d.cnt_sw, d.nbr_sw,   \
                     d.cnt_ion, d.nbr_ion, d.rk_sw, d.off_sw, d.gate)
    if (nbc == NB_CELLS) NB_LAUNCH(NB_CELLS);
    else if (nbc == 4) NB_LAUNCH(4);
    else if (nbc == 16) NB_LAUNCH(16);
    else NB_LAUNCH(8);
#undef NB_LAUNCH
  }
  {
    KScope ks_(d, KT_LISTAUX);
    hipLaunchKernelGGL(k_maxcnt, dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s, d.cnt_sw,
                       d.cnt_ion, d.pstride, d.partials, d.binst, d.gate);
  }
  const int op[2] = {1, 1};
  const int dst[2] = {DCTL_IDX(maxcnt_sw), DCTL_IDX(maxcnt_ion)};
  finalize(d, 2, FIB_BLOCK, op, dst, s, 1, d.gate);
}

void launch_cellid(const DevBufs &d, int inst, int32_t *dout, hipStream_t s) {
  {
    KScope ks_(d, KT_IO);
    hipLaunchKernelGGL(k_cellid, dim3(nblk(d.NA, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, d.X[d.cur],
                       d.NA, d.I, inst, dout);
  }
}

template <bool MIXED, bool EN, bool VIR>
static void force_ab(const DevBufs &d, const FConst &fc, hipStream_t s, bool extra) {
  typedef typename Prec<MIXED>::T4 T4;
  {
    KScope ks_(d, KT_LEGS);
    hipLaunchKernelGGL((k_legs<MIXED, EN, VIR>), dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s,
                       d.X[d.cur], d.I, fc, d.tab, d.NA, d.Wsw, d.Wion, d.cnt_sw, d.nbr_sw,
                       d.cnt_ion, d.nbr_ion, d.cnt_leg, d.amask, d.Fpart, d.upart, d.pstride,
                       d.partials, d.halt);
  }
  {
    KScope ks_(d, KT_TRI);
    // the MD step's variant (no energy, no virial) in every tile; the others
    // in the default tile
    const int ns = (EN || VIR || !d.tri_ns) ? TriStage<MIXED>::N : d.tri_ns;
    const int tb = (EN || VIR || !d.tri_tb) ? TRI_BATCH : d.tri_tb;
#define TRI_LAUNCH(NS_, TB_)                                                                  \
  hipLaunchKernelGGL((k_tri<MIXED, EN, VIR, NS_, TB_>), dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), \
                     0, s, d.X[d.cur], d.I, fc, d.NA, d.Wsw, d.cnt_leg, d.nbr_sw, d.rk_sw,   \
                     d.off_sw, d.amask, (T4 *)d.G, d.RA, d.pstride,                          \
                     d.partials + (size_t)FIB_NRED * d.pstride, d.halt)
    if constexpr (EN || VIR) {
      TRI_LAUNCH(TriStage<MIXED>::N, TRI_BATCH);
    } else {
      if (ns == TriStage<MIXED>::N && tb == TRI_BATCH) TRI_LAUNCH(TriStage<MIXED>::N, TRI_BATCH);
      else if (ns == 128 && tb == 2) TRI_LAUNCH(128, 2);
      else if (ns == 128 && tb == 4) TRI_LAUNCH(128, 4);
      else if (ns == 256 && tb == 2) TRI_LAUNCH(256, 2);
      else TRI_LAUNCH(256, 4);
    }
#undef TRI_LAUNCH
  }
  // k_extra is launched whenever its partials are reduced (energy, virial)
  // or stale-list terms can exist; the MD step's gate rules the latter out
  if (extra || EN || VIR) {
    KScope ks_(d, KT_EXTRA);
    hipLaunchKernelGGL((k_extra<MIXED, EN, VIR>), dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s,
                       d.X[d.cur], d.I, fc, d.NA, d.Wsw, d.cnt_sw, d.nbr_sw, d.cnt_leg, d.amask,
                       d.disp, d.ctl, (T4 *)d.EX,
                       (T4 *)d.EXself, d.exmask, d.pstride,
                       d.partials + (size_t)2 * FIB_NRED * d.pstride, d.halt);
  }
  if (EN || VIR) {
    int op[10], dst[10];
    for (int k = 0; k < 10; ++k) { op[k] = 0; dst[k] = DCTL_IDX(utot) + k; }
    finalize(d, 10, FIB_BLOCK, op, dst, s, 3);
  }
}

template <bool MIXED, bool EN, int MODE>
static void force_c(const DevBufs &d, const FConst &fc, const StepArgs &sa, double *Fo,
                    hipStream_t s) {
  typedef typename Prec<MIXED>::T4 T4;
  const double fscale = MIXED ? fc.eps[0][0] * fc.isig[0][0] : 1.0;
  const double escale = MIXED ? fc.eps[0][0] : 1.0;
  {
    KScope ks_(d, KT_GATHER);
    hipLaunchKernelGGL((k_force3<MIXED, EN, MODE>), dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK),
                       0, s, d.X[d.cur], d.I, d.NA, d.Wsw, d.nbr_sw, d.rk_sw, d.off_sw,
                       (const T4 *)d.G, d.RA, d.cnt_leg, d.cnt_sw, d.Fpart,
                       d.upart, fscale, escale, Fo,
                       d.uout, d.V[d.cur], sa, d.pstride, d.partials, d.ctl, (const T4 *)d.EX,
                       (const T4 *)d.EXself, d.exmask, d.halt);
  }
  if (MODE == 1) finalize_thermo(d, s);
  if (MODE == 2) {
    const int op[4] = {1, 0, 0, 0};
    const int dst[4] = {DCTL_IDX(vmax), DCTL_IDX(ke_sum), DCTL_IDX(ke_shell), DCTL_IDX(n_shell)};
    finalize(d, 4, FIB_BLOCK, op, dst, s);
  }
}

void launch_force_a(const DevBufs &d, const FConst &fc, bool mixed, bool energy, bool virial,
                    hipStream_t s, bool extra) {
  if (mixed) {
    if (energy) { if (virial) force_ab<true, true, true>(d, fc, s, extra); else force_ab<true, true, false>(d, fc, s, extra); }
    else { if (virial) force_ab<true, false, true>(d, fc, s, extra); else force_ab<true, false, false>(d, fc, s, extra); }
  } else {
    if (energy) { if (virial) force_ab<false, true, true>(d, fc, s, extra); else force_ab<false, true, false>(d, fc, s, extra); }
    else { if (virial) force_ab<false, false, true>(d, fc, s, extra); else force_ab<false, false, false>(d, fc, s, extra); }
  }
}

template <bool MIXED>
static void force_c_mode(const DevBufs &d, const FConst &fc, bool energy, int mode,
                         const StepArgs &sa, double *Fout, hipStream_t s) {
  if (mode == 1) force_c<MIXED, false, 1>(d, fc, sa, Fout, s);
  else if (mode == 2) force_c<MIXED, false, 2>(d, fc, sa, Fout, s);
  else if (energy) force_c<MIXED, true, 0>(d, fc, sa, Fout, s);
  else force_c<MIXED, false, 0>(d, fc, sa, Fout, s);
}

void launch_force_b(const DevBufs &d, const FConst &fc, bool mixed, bool energy, int mode,
                    const StepArgs &sa, double *Fout, hipStream_t s) {
  if (mixed) force_c_mode<true>(d, fc, energy, mode, sa, Fout, s);
  else force_c_mode<false>(d, fc, energy, mode, sa, Fout, s);
}

void launch_pos(const DevBufs &d, const StepArgs &sa, hipStream_t s) {
  {
    KScope ks_(d, KT_POS);
    hipLaunchKernelGGL(k_pos, dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s, d.X[d.cur], d.I, d.NA,
                       d.V[d.cur], d.F[d.cur], d.Xi[d.cur], sa, d.pstride, d.partials, d.ctl,
                       d.Xb, d.disp);
  }
  const int op[4] = {0, 1, 1, 1};
  const int dst[4] = {DCTL_IDX(eq_sum), DCTL_IDX(kemax), DCTL_IDX(maxdisp), DCTL_IDX(maxdisp2)};
  finalize(d, 4, FIB_BLOCK, op, dst, s);
}

void launch_md_pos(const DevBufs &d, const FConst &fc, const StepArgs &sa, int *gate,
                   hipStream_t s) {
  int *leave = gate + 2 + 2 * d.ninst;
  {
    KScope ks_(d, KT_POS);
    hipLaunchKernelGGL(k_md_pos, dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s, d.X[d.cur], d.I,
                       d.NA, d.V[d.cur], d.F[d.cur], d.Xi[d.cur], sa, d.pstride, d.partials,
                       d.ctl, d.Xb, d.disp, leave, d.halt);
  }
  const int op[5] = {RED_TOP2, RED_TOP2_CNT, RED_TOP2_SND, RED_SUM, RED_SUM};
  const int dst[5] = {DCTL_IDX(maxdisp), DCTL_IDX(nmaxdisp), DCTL_IDX(maxdisp2),
                      DCTL_IDX(e_bath_step), DCTL_IDX(e_out_step)};
  finalize(d, 5, FIB_BLOCK, op, dst, s);
  hipLaunchKernelGGL(k_md_decide, dim3(1), dim3(DECIDE_T), 0, s, d.ninst, d.I, d.ctl, sa.skin,
                     gate, leave, d.X[d.cur], fc, d.Wsw, d.Wion, d.cnt_sw, d.nbr_sw, d.cnt_ion,
                     d.nbr_ion, d.halt);
}

void launch_fire_energy(const DevBufs &d, const FConst &fc, int q, hipStream_t s) {
  hipLaunchKernelGGL(k_fire_energy, dim3(1), dim3(FIB_WAVE), 0, s, q, d.inv, d.X[d.cur], d.I, fc,
                     d.Wsw, d.Wion, d.cnt_sw, d.nbr_sw, d.cnt_ion, d.nbr_ion, d.ctl);
}

void launch_md_ctl(const DevBufs &d, const StepArgs &sa, hipStream_t s) {
  hipLaunchKernelGGL(k_md_ctl, dim3(nblk(d.ninst, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, d.ninst,
                     d.ctl, sa, d.halt);
}

void launch_md_init(const DevBufs &d, const StepArgs &sa, bool reset, hipStream_t s) {
  hipLaunchKernelGGL(k_md_init, dim3(nblk(d.ninst, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, d.ninst,
                     d.ctl, sa, reset ? 1 : 0);
}

void launch_md_report(const DevBufs &d, const StepArgs &sa, hipStream_t s) {
  {
    KScope ks_(d, KT_REDUCE);
    hipLaunchKernelGGL(k_md_report, dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s, d.X[d.cur],
                       d.I, d.NA, d.Xi[d.cur], sa, d.pstride, d.partials);
  }
  const int op[4] = {0, 0, 0, 1};
  const int dst[4] = {DCTL_IDX(n_removed), DCTL_IDX(n_ga_removed), DCTL_IDX(n_displaced),
                      DCTL_IDX(ion_z)};
  finalize(d, 4, FIB_BLOCK, op, dst, s);
}

void launch_mark_build(const DevBufs &d, hipStream_t s, bool count) {
  KScope ks_(d, KT_LISTAUX);
  hipLaunchKernelGGL(k_mark_build, dim3(gcap(std::max(d.NA, d.ninst))), dim3(FIB_BLOCK), 0,
                     s, d.NA, d.ninst, d.Wsw, d.Wion, d.X[d.cur], d.Xb, d.disp, d.ctl, d.binst,
                     d.gate, d.halt, count ? 1.0 : 0.0);
}

// Rows of an ELL list copied into a wider layout (the list grew).
template <typename T>
__global__ void k_copy_rows(int NA, int wold, const T *__restrict__ src, int wnew,
                            T *__restrict__ dst) {
  GRID_STRIDE(p, NA) {
    for (int k = 0; k < wold; ++k) dst[(size_t)p * wnew + k] = src[(size_t)p * wold + k];
  }
}

void launch_copy_rows(int NA, int wold, const int *src, int wnew, int *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_rows<int>, dim3(gcap(NA)), dim3(FIB_BLOCK), 0, s, NA, wold, src,
                     wnew, dst);
}
void launch_copy_rows(int NA, int wold, const uint16_t *src, int wnew, uint16_t *dst,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_copy_rows<uint16_t>, dim3(gcap(NA)), dim3(FIB_BLOCK), 0, s, NA, wold, src,
                     wnew, dst);
}

// After a removal-list overflow: no atom stays marked as leaving.
__global__ void k_clear_leaving(int NA, double4 *__restrict__ X) {
  GRID_STRIDE(p, NA) {
    const double4 x = X[p];
    if (meta_of(x) & META_LEAVING) {
      double4 y = x;
      y.w = __longlong_as_double((long long)(wbits(x) & ~(uint64_t)META_LEAVING));
      X[p] = y;
    }
  }
}

// Clears the list-overflow flags and the halt flag (after the lists grew),
// or the removal-list overflow flags.
__global__ void k_clear_overflow(int ninst, DCtl *__restrict__ ctl, int *__restrict__ halt,
                                 int lists) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ninst) {
    if (lists) ctl[k].overflow = 0; else ctl[k].leave_overflow = 0;
  }
  if (k == 0 && lists && halt) *halt = 0;
}

void launch_clear_overflow(const DevBufs &d, hipStream_t s, bool lists) {
  hipLaunchKernelGGL(k_clear_overflow, dim3(nblk(d.ninst, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s,
                     d.ninst, d.ctl, d.halt, lists ? 1 : 0);
}

// Start state of a handle-free evaluation, written by kernels: the instance
// record (by value), zeroed bookkeeping, the energy total out.
__global__ void k_put_inst(DInst in, DInst *__restrict__ I) {
  if (blockIdx.x == 0 && threadIdx.x == 0) I[0] = in;
}
__global__ void k_fill_zero(uint32_t *__restrict__ p, size_t nwords) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < nwords;
       k += (size_t)gridDim.x * blockDim.x)
    p[k] = 0u;
}
__global__ void k_copy_utot(const DCtl *__restrict__ ctl, double *__restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = ctl[0].utot;
}

void launch_put_inst(const DevBufs &d, const DInst &in, hipStream_t s) {
  hipLaunchKernelGGL(k_put_inst, dim3(1), dim3(FIB_WAVE), 0, s, in, d.I);
}
void launch_fill_zero(void *p, size_t bytes, hipStream_t s) {
  const size_t nw = bytes / 4;  // every array of the arena is a multiple of 4 bytes
  hipLaunchKernelGGL(k_fill_zero, dim3(gcap((long)std::min<size_t>(nw, 1u << 30))),
                     dim3(FIB_BLOCK), 0, s, static_cast<uint32_t *>(p), nw);
}
void launch_copy_utot(const DevBufs &d, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_utot, dim3(1), dim3(FIB_WAVE), 0, s, d.ctl, out);
}

void launch_clear_leaving(const DevBufs &d, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_leaving, dim3(gcap(d.NA)), dim3(FIB_BLOCK), 0, s, d.NA, d.X[d.cur]);
}

void launch_sums(const DevBufs &d, const StepArgs &sa, hipStream_t s) {
  {
    KScope ks_(d, KT_REDUCE);
    hipLaunchKernelGGL(k_sums, dim3(d.NA / FIB_BLOCK), dim3(FIB_BLOCK), 0, s, d.X[d.cur], d.I, d.NA,
                       d.V[d.cur], sa, d.pstride, d.partials, d.ctl);
  }
  finalize_thermo(d, s);
}

void launch_import(const DevBufs &d, int inst, const DInst &in, const double *X,
                   const double *V, const int32_t *atype, const int32_t *xsides,
                   const double *Xi, hipStream_t s) {
  {
    KScope ks_(d, KT_IO);
    hipLaunchKernelGGL(k_import, dim3(nblk(in.natm_pad, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, inst, in,
                       d.NA, X, V, atype, xsides, Xi, d.X[0], d.X[1], d.V[0], d.V[1], d.F[0],
                       d.F[1], d.Xi[0], d.Xi[1], d.inv);
  }
}

void launch_export_x(const DevBufs &d, int inst, const DInst &in, double *out, hipStream_t s) {
  (void)inst;
  {
    KScope ks_(d, KT_IO);
    hipLaunchKernelGGL(k_export_x, dim3(nblk(in.natm_pad, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, in,
                       d.X[d.cur], out);
  }
}

void launch_export_vec(const DevBufs &d, int inst, const DInst &in, const double *src,
                       double *out, hipStream_t s) {
  (void)inst;
  {
    KScope ks_(d, KT_IO);
    hipLaunchKernelGGL(k_export_vec, dim3(nblk(in.natm_pad, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s, in,
                       d.NA, d.X[d.cur], src, out);
  }
}

void launch_export_scalar(const DevBufs &d, int inst, const DInst &in, const double *src,
                          double *out, hipStream_t s) {
  (void)inst;
  {
    KScope ks_(d, KT_IO);
    hipLaunchKernelGGL(k_export_scalar, dim3(nblk(in.natm_pad, FIB_BLOCK)), dim3(FIB_BLOCK), 0, s,
                       in, d.X[d.cur], src, out);
  }
}

void launch_set_atom(const DevBufs &d, int p, const double x[3], const double v[3],
                     const StepArgs &sa, hipStream_t s) {
  {
    KScope ks_(d, KT_IO);
    hipLaunchKernelGGL(k_set_atom, dim3(1), dim3(FIB_WAVE), 0, s, d.NA, d.X[d.cur], d.V[d.cur], p,
                       x[0], x[1], x[2], v[0], v[1], v[2], sa, d.ctl);
  }
}

}  // namespace fibk
