// Streamed fused variant: the batched SQP solve and adjoint for any horizon
// and nx <= 64, with the per-instance QP data in a caller-provided device
// workspace instead of registers and LDS.
//
// The register-resident kernels of dmpc_fused.hip keep each time block's
// A/B rows in registers and D_t in registers or LDS, which bounds the
// horizon (about 64 blocks of 13 states, 128 of 7) and excludes nx > 16.
// Here a workgroup still owns one instance and lane i of a G-lane group
// still owns row i of a block, but the groups walk over the blocks in
// rounds (block t = round * NG + group) and the per-block matrices live in
// global memory:
//   * [A_t | B_t] from the linearization (row-major), plus its transpose for
//     nx > 16 (compact_blocks below);
//   * D_t = theta_t^-1 in float64 (packed lower triangle for nx <= 16);
//   * the PCG vectors (each lane owns its entries; neighbours are read after
//     a workgroup barrier), in LDS when they fit, else in the workspace.
// The arithmetic and its order are those of the register-resident kernels
// (pcg.py:350-445 with the factored block-Jacobi-stair preconditioner), so
// both variants agree to rounding.  Assembly (theta_t, its Cholesky factor
// and D_t) runs in its own kernel with one workgroup per (instance, block).
#include <type_traits>

#include "dmpc_device.h"
#include "dmpc_internal.h"
#include "dmpc_linesearch.h"
#include "dmpc_models.h"

namespace dmpc {
namespace {

constexpr int kStreamVectors = 11;
enum { V_LAM = 0, V_P, V_R, V_RT, V_Y, V_YQ, V_YX, V_AP, V_YXO, V_GB, V_GAM };

__host__ __device__ inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Packed lower triangle of a symmetric nx x nx block: entry (r, c), r >= c.
__host__ __device__ constexpr int tri(int r, int c) { return r * (r + 1) / 2 + c; }

// Block layout by size.  Compact (nx <= 16): the kernels read A_t and B_t in
// both orientations from the caller's [A|B] rows (a lane reads a column for
// A^T v, its own row for A v: a row is a few cache lines that serve all its
// loads, and phase C's two A reads of a block share them) and D_t as a packed
// lower triangle, which cuts the bytes per PCG iteration by about a third.
// For larger blocks a lane's row spans many lines per load, so the workspace
// holds the transposed copy AT = [A|B]^T (every load coalesced) and D_t in
// full, transposed (DT[t][j][i] = D_t[i][j]).
__host__ __device__ constexpr bool compact_blocks(int nx) { return nx <= 16; }

// Per-instance workspace: 11 padded vectors (N+3 blocks, zero blocks at
// both ends), D_t = theta_t^-1 (N+1 blocks, packed or full), 1/L_kk (N+1,
// nx), the line-search step and scratch, and (large blocks) AT (N, nx+nu, nx)
// float.  Each array starts on 256 bytes.
constexpr size_t kOrderSlot = 256;

struct Work {
  double* v[kStreamVectors];
  double *DT, *ID, *DX, *DU, *LS;
  float* AT;
  const double* DG;  // the workspace's D blocks, when LDS holds D_1 only (d1_only)
};

// Block a knot's data is read from.  A time-invariant model (linear in x and
// u, kTimeInvariant) has the same [A_t | B_t] at every knot, and with the
// reference's time-invariant weights theta_t -- and so D_t -- is the same for
// 1 <= t < N (theta_0 = Qinv_0 and theta_N has the terminal weight).  Such
// instances store and read block 0 of [A|B] (and of its transposed copy)
// and D_0, D_1, D_N only: the blocks a PCG round reads then stay in cache
// instead of streaming from HBM, and the results are the same bits, since
// the shared copies are the values every knot would compute.  SH is a
// template flag of the streamed kernels (instantiated both ways for
// time-invariant models only), so the other kernels keep their plain
// per-knot addressing.  The SH kernels stage the shared copies in LDS
// (stage_shared_blocks): block 0 of [A|B] and of its transposed copy, and
// D_0, D_1, D_N in slots 0, 1, 2 -- for 64-lane groups D_1 only (d1_only).
// Their block index is zero through an
// empty asm: with the block addresses known loop-invariant the compiler
// hoists the block loads out of the PCG rounds and spills them (the
// 48-state chain's forward launch took 2.4x as long so).
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+s"(v));
  return v;
}
template <bool SH>
__device__ __forceinline__ int shared_ab(int tp) {
  if constexpr (SH) return opaque(0);
  return tp;
}
template <bool SH>
__device__ __forceinline__ int shared_d(int t, int N) {
  if constexpr (SH) return t == 0 ? 0 : (t < N ? 1 : 2);
  return t;
}

__host__ __device__ inline size_t carve_work(char* base, int nx, int nu, int N, int ls_doubles, Work* w) {
  size_t off = 0;
  const size_t vbytes = align256((size_t)(N + 3) * nx * 8);
  for (int k = 0; k < kStreamVectors; ++k) {
    if (w) w->v[k] = reinterpret_cast<double*>(base + off);
    off += vbytes;
  }
  if (w) w->DT = reinterpret_cast<double*>(base + off);
  off += align256((size_t)(N + 1) * (compact_blocks(nx) ? tri(nx, 0) : nx * nx) * 8);
  if (w) w->ID = reinterpret_cast<double*>(base + off);
  off += align256((size_t)(N + 1) * nx * 8);
  if (w) w->DX = reinterpret_cast<double*>(base + off);
  off += align256((size_t)(N + 1) * nx * 8);
  if (w) w->DU = reinterpret_cast<double*>(base + off);
  off += align256((size_t)(N + 1) * nu * 8);
  if (w) w->LS = reinterpret_cast<double*>(base + off);
  off += align256((size_t)ls_doubles * 8);
  if (w) w->AT = reinterpret_cast<float*>(base + off);
  if (!compact_blocks(nx)) off += align256((size_t)N * (nx + nu) * nx * 4);
  // last kOrderSlot bytes: one int of a launch order spread over the
  // instances' regions (entry i in region i), which no kernel of the
  // instance writes
  off += kOrderSlot;
  return off;
}

// Entry `lane` of block tb of a padded vector (slot tb + 1), 0 past nx.
template <int NX>
__device__ __forceinline__ double vget(const double* V, int tb, int lane) {
  return lane < NX ? V[(tb + 1) * NX + lane] : 0.0;
}

// Threads per workgroup of the streamed kernels: up to 64 lane groups of 16
// (1024 threads, 16 waves) for blocks of up to 16 states, whose PCG loop fits
// in the 128 registers per lane that 16 waves leave (the line search and the
// recovery spill a little, once per SQP iteration); 512 for 64-lane blocks.
// More lane groups mean fewer rounds per PCG phase and more waves to cover
// the latency of each round's block loads.
template <class Model>
__host__ __device__ constexpr int stream_max_threads() { return Model::NX <= 16 ? 1024 : 512; }

// 64-lane groups (more than 32 states) hold one knot per wavefront.  Their
// shared-block kernels stage only D_1 of the three D blocks in LDS (with the
// [A|B] copies 64.5 KB for 48 states instead of 101.4 KB) and read D_0 and
// D_N, which two knots of 257 use, from the workspace: the knot, and so the
// choice, is uniform over the wavefront.  Two 512-thread workgroups then fit
// in a CU's LDS, and the kernels without the line search are compiled for
// the 128 registers per lane that four waves per SIMD leave (the 48-state
// chain's forward launch: 0.669 -> 0.548 s; the line-search build spills at
// any budget and keeps the compiler's choice).
__host__ __device__ constexpr bool d1_only(int nx) { return nx > 32; }
template <class Model, bool SH, bool LS>
__host__ __device__ constexpr int stream_waves_per_simd() { return SH && !LS && d1_only(Model::NX) ? 4 : 1; }

// The first NV padded vectors (V_LAM, V_P, V_R, V_RT, V_Y, V_YQ, V_YX, then
// V_AP, V_YXO, V_GB, V_GAM) in LDS instead of the workspace: every PCG phase
// reads and writes them, and the neighbour exchanges go through them.  NV is
// a template constant so that each access compiles to LDS or global
// instructions, not generic ones.
template <int NV, int NX>
__device__ __forceinline__ void vectors_in_lds(Work* w, double* lds, int Np) {
  const size_t n = (size_t)(Np + 3) * NX;
#pragma unroll
  for (int k = 0; k < NV; ++k) w->v[k] = lds + k * n;
}

__host__ __device__ inline size_t vector_lds_bytes(int nv, int nx, int Np) { return (size_t)nv * (Np + 3) * nx * 8; }

// Diagonal cost model in closed form (sqp.py:160-197,
// optimal_control_problem.py:249-323) for one instance.
template <int NX, int NU>
struct CostView {
  const double *wx, *wu, *wf, *xref, *uref, *xf;  // instance rows / bases
  int penalize, N;                                 // N: the instance's horizon
  // the reciprocals a PCG round needs, computed once per instance
  // (cache_reciprocals; the same expressions, so the same bits): Qinv of
  // the lane's row before and at the terminal knot, Rinv in LDS
  double qi_run = 0.0, qi_end = 0.0;
  const double* rtab = nullptr;
  __device__ double qinv_lane(int t) const { return t < N ? qi_run : qi_end; }
  __device__ double qinv(int t, int i) const {
    const double w = t < N ? wx[i] : wx[i] + wf[i];
    return 1.0 / fmax(2.0 * w, kPsdMinEig);
  }
  __device__ double q(int t, int i) const {
    return t < N ? -2.0 * wx[i] * xref[t * NX + i] : -2.0 * (wx[i] + wf[i]) * xf[i];
  }
  __device__ double rinv(int k) const { return 1.0 / fmax(2.0 * wu[k], kPsdMinEig); }
  __device__ double r(int t, int k) const {
    return (t < N && penalize) ? -2.0 * wu[k] * uref[t * NU + k] : 0.0;
  }
};

// Fills c's cached reciprocals; rtab must be read only after a barrier.
template <int NX, int NU>
__device__ __forceinline__ void cache_reciprocals(CostView<NX, NU>& c, double* rtab, int lane) {
  if (threadIdx.x < NU) rtab[threadIdx.x] = c.rinv(threadIdx.x);
  c.rtab = rtab;
  if (lane < NX) {
    c.qi_run = c.qinv(0, lane);  // t = 0 < N
    c.qi_end = c.qinv(c.N, lane);
  }
}

// sum_j a[j * stride] * v_j (a in global memory), v_j in lane j of the group.
template <int G, int M, class AT>
__device__ __forceinline__ double gdot_mem(const AT* a, int stride, double v, int lane) {
  AT r[M];
#pragma unroll
  for (int j = 0; j < M; ++j) r[j] = a[j * stride];
  return gdot<G, M>(r, v, lane, 0);
}

// Group sum of b_k * v for the NU columns k of B_{t-1} (brow: the lane's
// entry of column 0, the next at stride `stride`), and the B_{t-1} Rinv
// B_{t-1}^T v product (see bgram in the register variant).
// 64-lane groups: the NU <= 16 column sums of B_{t-1}^T v by one
// reduce-scatter over each 16-lane row instead of NU wavefront sums.  Slot s
// of lane l holds column l ^ s (zero past NU); the exchanges with lanes
// l ^ 15, l ^ 7, l ^ 3 and l ^ 1 (row_mirror, row_half_mirror, quad_perm
// [3,2,1,0] and [1,0,3,2]) keep the slots whose column shares the partner
// bit with l, so slot 0 ends with column l's sum over the row; two row
// exchanges add the four rows.  Then y = Rinv s (lane k of every row holds
// y_k) and g = B_{t-1} y with one DPP FMA per column.  The quadratic form
// v^T B Rinv B^T v is left to the caller as the row sum of v g.
template <int NX, int NU, class T>
__device__ __forceinline__ double bgram64(const T* brow, int stride, const CostView<NX, NU>& c, double v) {
  static_assert(NU <= 16, "one 16-lane row holds the column sums");
  const int l = threadIdx.x & 15;
  double P[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int col = l ^ s;
    const double b = (double)brow[(col < NU ? col : 0) * stride];
    P[s] = col < NU ? b * v : 0.0;
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) P[s] += dpp_mov<kDppMirror>(P[15 - s]);
#pragma unroll
  for (int s = 0; s < 4; ++s) P[s] += dpp_mov<kDppHalfMirror>(P[7 - s]);
#pragma unroll
  for (int s = 0; s < 2; ++s) P[s] += dpp_mov<0x1B>(P[3 - s]);  // quad_perm [3,2,1,0]
  P[0] += dpp_mov<0xB1>(P[1]);                                     // quad_perm [1,0,3,2]
  double e, o, lo, hi;
  rows_swap16(P[0], e, o);  // rows (R0, R1) and (R2, R3)
  rows_swap32(e + o, lo, hi);
  const double y = (l < NU ? c.rtab[l < NU ? l : 0] : 0.0) * (lo + hi);
  double g = 0.0;
#pragma unroll
  for (int k = 0; k < NU; ++k) bcast_fmac_at<16>(g, y, (double)brow[k * stride], k, 0);  // lane k of the row
  return g;
}

template <int G, int NX, int NU, class T>
__device__ __forceinline__ double bgram_mem(const T* brow, int stride, const CostView<NX, NU>& c, double v,
                                            double* quad) {
  if constexpr (G == 64) {
    *quad = 0.0;  // the caller adds v g instead
    return bgram64<NX, NU>(brow, stride, c, v);
  }
  double g = 0.0, q = 0.0;
#pragma unroll
  for (int k = 0; k < NU; ++k) {
    const double bk = brow[k * stride];
    const double sk = gsum<G>(bk * v);
    const double yk = c.rtab[k] * sk;
    g += bk * yk;
    q += sk * yk;
  }
  *quad = q;
  return g;
}

// ---------------------------------------------------------------------------
// Assembly: one workgroup of 64 threads per (instance, block t).  Thread i
// owns row i.  theta_t (pcg.py:182-196 with As_next = -I), its Cholesky
// factor (right-looking, the operation order of the register variant's
// shuffle version) and D_t = theta_t^-1 by one column per thread, written to
// DT (packed: thread i writes rows j >= i of column i; full: transposed).
// Large blocks also get AT[t-1] = [A_{t-1} | B_{t-1}]^T.
struct AssembleArgs {
  int B, N;  // padded horizon
  const int32_t* horizons;
  const float* AB;
  const double *wx, *wu, *wf;
  // forward: skip instances that stopped iterating (null in the backward)
  const int32_t* num_iter;
  const double* conv_err;
  int maxit;
  double tol;
  char* work;
  size_t work_stride;
  int ls_doubles;
  int shared_blocks;  // assemble knots 0, 1 and N only (see shared_d)
};

template <int NX, int NU>
__global__ void __launch_bounds__(64) stream_assemble_kernel(AssembleArgs a) {
  constexpr int NZ = NX + NU;
  constexpr bool kColsInLds = NX > 16;     // large blocks: substitution columns in LDS
  __shared__ double AP[NX * NZ];          // [A_{t-1} | B_{t-1}] as float64
  __shared__ double TH[NX * (NX + 1)];    // theta_t, then its Cholesky factor
  __shared__ double IDs[NX];
  __shared__ double COL[kColsInLds ? NX * 64 : 1];  // COL[j * 64 + thread]
  const int Np = a.N, NB1 = Np + 1;
  const long long bt = blockIdx.x;
  const int b = (int)(bt / NB1), t = (int)(bt - (long long)b * NB1);
  if (a.num_iter) {
    const int it = a.num_iter[b];
    if (!((it < a.maxit && a.conv_err[b] > a.tol) || it < 1)) return;
  }
  const int N = horizon_of(a.horizons, Np, b);
  if (t > N) return;
  if (a.shared_blocks && t > 1 && t < N) return;  // the same as knot 1
  Work w;
  carve_work(a.work + (size_t)b * a.work_stride, NX, NU, Np, a.ls_doubles, &w);
  const int i = threadIdx.x;
  const double* wx = a.wx + (size_t)b * NX;
  const double* wu = a.wu + (size_t)b * NU;
  const double* wf = a.wf + (size_t)b * NX;
  auto qinv = [&](int tt, int k) {
    const double ww = tt < N ? wx[k] : wx[k] + wf[k];
    return 1.0 / fmax(2.0 * ww, kPsdMinEig);
  };
  if (t >= 1) {
    const float* Ap = a.AB + ((size_t)b * Np + (t - 1)) * NX * NZ;
    float* ATp = w.AT + (size_t)(t - 1) * NZ * NX;
    for (int e = i; e < NX * NZ; e += 64) {
      const float v = Ap[e];
      AP[e] = v;
      if constexpr (!compact_blocks(NX)) {
        const int r = e / NZ, k = e - r * NZ;
        ATp[k * NX + r] = v;
      }
    }
  }
  __syncthreads();
  if (i < NX) {
    // row i of theta_t; A_{t-1} Qinv_{t-1} row and B_{t-1} Rinv row first
    double aq[NX], bq[NU];
    if (t >= 1) {
#pragma unroll
      for (int k = 0; k < NX; ++k) aq[k] = AP[i * NZ + k] * qinv(t - 1, k);
#pragma unroll
      for (int k = 0; k < NU; ++k) bq[k] = AP[i * NZ + NX + k] * (1.0 / fmax(2.0 * wu[k], kPsdMinEig));
    }
    const double qi = qinv(t, i);
#pragma unroll 1
    for (int j = 0; j < NX; ++j) {
      double acc = (j == i) ? qi : 0.0;
      if (t >= 1) {
#pragma unroll
        for (int k = 0; k < NX; ++k) acc += aq[k] * AP[j * NZ + k];
#pragma unroll
        for (int k = 0; k < NU; ++k) acc += bq[k] * AP[j * NZ + NX + k];
      }
      TH[i * (NX + 1) + j] = acc;
    }
  }
  __syncthreads();
  // right-looking Cholesky: row i ends with L[i][k], k <= i
  for (int k = 0; k < NX; ++k) {
    const double lkk = sqrt(TH[k * (NX + 1) + k]);
    const double il = 1.0 / lkk;
    __syncthreads();  // every thread has read the pivot
    if (i == k) {
      TH[k * (NX + 1) + k] = lkk;
      IDs[k] = il;
    } else if (i > k && i < NX) {
      TH[i * (NX + 1) + k] *= il;
    }
    __syncthreads();
    if (i > k && i < NX) {
      const double lik = TH[i * (NX + 1) + k];
      for (int j = k + 1; j <= i; ++j) TH[i * (NX + 1) + j] -= lik * TH[j * (NX + 1) + k];
    }
    __syncthreads();
  }
  // column i of theta^-1 = L^-T L^-1 e_i (forward then back substitution;
  // the back substitution overwrites the forward result in place)
  if (i < NX) {
    constexpr bool kPacked = compact_blocks(NX);
    double* DTt = w.DT + (size_t)t * (kPacked ? tri(NX, 0) : NX * NX);
    if constexpr (kColsInLds) {
      double* v = COL + i;  // v[j * 64]
      for (int j = 0; j < NX; ++j) {
        double acc = (j == i) ? 1.0 : 0.0;
        for (int k = 0; k < j; ++k) acc -= TH[j * (NX + 1) + k] * v[k * 64];
        v[j * 64] = acc * IDs[j];
      }
      for (int j = NX - 1; j >= 0; --j) {
        double acc = v[j * 64];
        for (int k = j + 1; k < NX; ++k) acc -= TH[k * (NX + 1) + j] * v[k * 64];
        v[j * 64] = acc * IDs[j];
      }
      for (int j = 0; j < NX; ++j) {
        if (!kPacked) DTt[j * NX + i] = v[j * 64];
        else if (j >= i) DTt[tri(j, i)] = v[j * 64];
      }
    } else {
      double v[NX], d[NX];
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        double acc = (j == i) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < j; ++k) acc -= TH[j * (NX + 1) + k] * v[k];
        v[j] = acc * IDs[j];
      }
#pragma unroll
      for (int j = NX - 1; j >= 0; --j) {
        double acc = v[j];
#pragma unroll
        for (int k = j + 1; k < NX; ++k) acc -= TH[k * (NX + 1) + j] * d[k];
        d[j] = acc * IDs[j];
      }
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        if (!kPacked) DTt[j * NX + i] = d[j];
        else if (j >= i) DTt[tri(j, i)] = d[j];
      }
    }
  }
}

// D_t v: sum_j D_t[lane][j] v_j (v_j in lane j of the group), 0 when !ok.
template <int G, int NX, bool SH>
__device__ __forceinline__ double dmul_mem(const Work& w, int t, int N, double v, int lane, bool ok) {
  const int l = lane < NX ? lane : NX - 1;  // idle lanes read row NX-1 with lane NX-1: an LDS broadcast, not a bank conflict
  double acc;
  t = shared_d<SH>(t, N);
  if constexpr (SH && d1_only(NX)) {
    const int slot = __builtin_amdgcn_readfirstlane(t);
    if (slot == 1) acc = gdot_mem<G, NX>(w.DT + l, NX, v, lane);
    else acc = gdot_mem<G, NX>(w.DG + (size_t)(slot == 0 ? 0 : N) * NX * NX + l, NX, v, lane);
  } else if constexpr (compact_blocks(NX)) {
    const double* D = w.DT + (size_t)t * tri(NX, 0);
    const int lrow = tri(l, 0);
    double r[NX];
#pragma unroll
    for (int j = 0; j < NX; ++j) r[j] = D[l >= j ? lrow + j : tri(j, 0) + l];
    acc = gdot<G, NX>(r, v, lane, 0);
  } else {
    acc = gdot_mem<G, NX>(w.DT + (size_t)t * NX * NX + l, NX, v, lane);
  }
  return ok ? acc : 0.0;
}

// The lane's entry of column 0 of A_{t-1} (row `li`) and the stride to the
// next column, in whichever copy the block layout reads; B_{t-1} follows at
// NX strides.
template <int NX, int NU, bool SH, class T>
__device__ __forceinline__ const T* arow_mem(const T* ABb, const T* AT, int t, int li) {
  constexpr int NZ = NX + NU;
  const size_t tp = (size_t)shared_ab<SH>(t >= 1 ? t - 1 : 0);
  if constexpr (compact_blocks(NX)) return ABb + tp * NX * NZ + li * NZ;
  else return AT + tp * NZ * NX + li;
}
template <int NX>
__device__ __forceinline__ constexpr int arow_stride() { return compact_blocks(NX) ? 1 : NX; }

// Per-instance setup shared by the forward and the backward kernels.
template <int G, int NX, int NU>
struct Blocks {
  int NG, grp, lane, rounds, N;
  __device__ void init(int Nb) {
    NG = blockDim.x / G;
    grp = threadIdx.x / G;
    lane = threadIdx.x % G;
    N = Nb;
    rounds = (Nb + 1 + NG - 1) / NG;
  }
  // block of round r (clamped to N for masked lanes) and whether lane owns a row
  __device__ int t(int r) const {
    const int v = r * NG + grp;
    return v > N ? N : v;
  }
  __device__ bool ok(int r) const { return r * NG + grp <= N && lane < NX; }
};

// LDS bytes the SH kernels stage the shared blocks in, after the PCG
// vectors that live in LDS: D_0, D_1, D_N (D_1 only, d1_only), and block 0
// of [A|B] and, for
// large blocks, its transposed copy, converted to float64 (exact, so the
// products keep their bits and need no conversion per term).
template <int NX, int NU>
__host__ __device__ constexpr size_t shared_block_bytes() {
  return (d1_only(NX) ? 1 : 3) * (size_t)(compact_blocks(NX) ? tri(NX, 0) : NX * NX) * 8 +
         (size_t)NX * (NX + NU) * 8 * (compact_blocks(NX) ? 1 : 2);
}

// Copies an instance's shared blocks from the workspace into LDS at `lds`
// (slots of shared_d / shared_ab; [A|B] and its transpose as float64) and
// points ABd, ATd and w.DT there (w.DG at the workspace's D blocks): every
// PCG round then reads them from LDS instead of the cache hierarchy.  Called
// by every thread; one barrier.
template <int NX, int NU>
__device__ void stage_shared_blocks(double* lds, const float* ABb, Work& w, int N, const double*& ABd,
                                    const double*& ATd) {
  constexpr int NZ = NX + NU;
  constexpr int DB = compact_blocks(NX) ? tri(NX, 0) : NX * NX;
  double* D = lds;
  double* AB = lds + (d1_only(NX) ? 1 : 3) * DB;
  double* AT = AB + NX * NZ;
  for (int k = threadIdx.x; k < DB; k += blockDim.x) {
    if constexpr (d1_only(NX)) {
      D[k] = w.DT[DB + k];
    } else {
      D[k] = w.DT[k];
      D[DB + k] = w.DT[DB + k];
      D[2 * DB + k] = w.DT[(size_t)N * DB + k];
    }
  }
  w.DG = w.DT;
  for (int k = threadIdx.x; k < NX * NZ; k += blockDim.x) {
    AB[k] = (double)ABb[k];
    if constexpr (!compact_blocks(NX)) AT[k] = (double)w.AT[k];
  }
  __syncthreads();
  ABd = AB;
  ATd = AT;
  w.DT = D;
}

// Element type of the blocks a streamed PCG reads: the float32 linearization
// in the workspace, or the float64 copies the SH kernels stage in LDS.
template <bool SH>
using BlockT = typename std::conditional<SH, double, float>::type;

// PCG on S lam = gamma (pcg.py:350-445) with the vectors in the workspace.
// gamma in V_GAM, the guess in V_LAM; the solution is left in V_LAM.  Same
// phases, products and reductions as pcg() of the register variant.
template <int G, int NX, int NU, bool SH, class T>
__device__ int stream_pcg(const Blocks<G, NX, NU>& K, const Work& w, const CostView<NX, NU>& c, const T* ABb,
                          const T* ATb, double* red3, double eps, int max_iter, double* eta_out) {
  constexpr int NZ = NX + NU;
  double* LAM = w.v[V_LAM];
  double* P = w.v[V_P];
  double* R = w.v[V_R];
  double* RT = w.v[V_RT];
  double* Y = w.v[V_Y];
  double* YQ = w.v[V_YQ];
  double* YX = w.v[V_YX];
  double* AP = w.v[V_AP];
  double* YXO = w.v[V_YXO];
  double* GB = w.v[V_GB];
  const double* GAM = w.v[V_GAM];
  const int lane = K.lane, N = K.N;
  const int li = lane < NX ? lane : NX - 1;  // (see dmul_mem)
  auto acol = [&](int t) { return ABb + (size_t)shared_ab<SH>(t < N ? t : 0) * NX * NZ + li; };  // stride NZ
  auto arow = [&](int t) { return arow_mem<NX, NU, SH>(ABb, ATb, t, li); };
  constexpr int AS = arow_stride<NX>();
  auto has_a = [&](int t) { return t < N; };
  auto has_p = [&](int t) { return t >= 1; };
  // ---- r = gamma - S lam0 ----------------------------------------------
  for (int r = 0; r < K.rounds; ++r) {
    const int t = K.t(r);
    const bool ok = K.ok(r);
    const double qi = ok ? c.qinv_lane(t) : 0.0;
    const double pv = vget<NX>(LAM, t, lane);
    const double pn = vget<NX>(LAM, t + 1, lane);
    const double ax = gdot_mem<G, NX>(acol(t), NZ, pn, lane);
    double q;
    const double gb = bgram_mem<G, NX, NU>(arow(t) + NX * AS, AS, c, pv, &q);
    const double wx = ((t == 0) ? pv : -pv) + ((ok && has_a(t)) ? ax : 0.0);
    const double yxo = qi * wx;
    if (ok) {
      YX[(t + 1) * NX + lane] = yxo;
      YXO[(t + 1) * NX + lane] = yxo;
      GB[(t + 1) * NX + lane] = has_p(t) ? gb : 0.0;
      AP[(t + 1) * NX + lane] = 0.0;
      P[(t + 1) * NX + lane] = 0.0;
    }
  }
  __syncthreads();
  for (int r = 0; r < K.rounds; ++r) {
    const int t = K.t(r);
    const bool ok = K.ok(r);
    const double u = vget<NX>(YX, t - 1, lane);
    const double au = gdot_mem<G, NX>(arow(t), AS, u, lane);
    const double g = (t == 0 ? 1.0 : -1.0) * vget<NX>(YXO, t, lane) + ((ok && has_p(t)) ? au : 0.0) +
                     vget<NX>(GB, t, lane);
    const double rv = ok ? vget<NX>(GAM, t, lane) + g : 0.0;
    const double y = dmul_mem<G, NX, SH>(w, t, N, rv, lane, ok);
    if (ok) {
      R[(t + 1) * NX + lane] = rv;
      Y[(t + 1) * NX + lane] = y;
      YQ[(t + 1) * NX + lane] = c.qinv_lane(t) * y;
    }
  }
  __syncthreads();
  double part = 0.0;
  for (int r = 0; r < K.rounds; ++r) {
    const int t = K.t(r);
    const bool ok = K.ok(r);
    const double a1v = gdot_mem<G, NX>(arow(t), AS, vget<NX>(YQ, t - 1, lane), lane);
    const double a2v = gdot_mem<G, NX>(acol(t), NZ, vget<NX>(Y, t + 1, lane), lane);
    const double a1 = (ok && has_p(t)) ? a1v : 0.0, a2 = (ok && has_a(t)) ? a2v : 0.0;
    const double rv = vget<NX>(R, t, lane);
    const double qi = ok ? c.qinv_lane(t) : 0.0;
    const double z = ok ? rv - (t == 1 ? 1.0 : -1.0) * a1 - (t == 0 ? 1.0 : -1.0) * qi * a2 : 0.0;
    const double rt = dmul_mem<G, NX, SH>(w, t, N, z, lane, ok);
    if (ok) {
      part += rv * rt;
      RT[(t + 1) * NX + lane] = rt;
    }
  }
  double eta = block_sum(part, red3 + 32);
  double beta = 0.0;
  int it = 0;
  // Each phase loads the next round's vector entries before it works on
  // the current round, so that their latency (workspace vectors of large
  // instances come from the cache hierarchy) overlaps the products.  A round
  // stores only entries of its own knots, which the next round does not load
  // (a lane that would see its own knot again is masked there).
  while (true) {
    const bool cont = (fabs(eta) > eps && it <= max_iter - 1) || it < 1;
    if (!cont) break;
    // ---- phase A: p = rt + beta p; w = G^T p; y = H^-1 w; v = -w.y ------
    double vpart = 0.0;
    double nrt = vget<NX>(RT, K.t(0), lane), np = vget<NX>(P, K.t(0), lane);
    double nrn = vget<NX>(RT, K.t(0) + 1, lane), nap = vget<NX>(AP, K.t(0), lane);
    for (int r = 0; r < K.rounds; ++r) {
      const int t = K.t(r);
      const bool ok = K.ok(r);
      const double crt = nrt, cp = np, crn = nrn, cap = nap;
      {
        const int tn = K.t(r + 1);
        nrt = vget<NX>(RT, tn, lane);
        np = vget<NX>(P, tn, lane);
        nrn = vget<NX>(RT, tn + 1, lane);
        nap = vget<NX>(AP, tn, lane);
      }
      const double pv = ok ? crt + beta * cp : 0.0;
      const double rn = crn;
      const double arn = gdot_mem<G, NX>(acol(t), NZ, rn, lane);
      const double apv = ok ? ((has_a(t) ? arn : 0.0) + beta * cap) : 0.0;
      double q;
      const double gb = bgram_mem<G, NX, NU>(arow(t) + NX * AS, AS, c, pv, &q);
      const double qi = ok ? c.qinv_lane(t) : 0.0;
      const double wx = ((t == 0) ? pv : -pv) + apv;
      const double yxo = qi * wx;
      if (ok) {
        P[(t + 1) * NX + lane] = pv;
        AP[(t + 1) * NX + lane] = apv;
        YX[(t + 1) * NX + lane] = yxo;
        YXO[(t + 1) * NX + lane] = yxo;
        GB[(t + 1) * NX + lane] = has_p(t) ? gb : 0.0;
        vpart += wx * yxo;
        if constexpr (G == 64) {
          if (has_p(t)) vpart += pv * gb;  // v^T B Rinv B^T v, summed over the rows
        } else {
          if (lane == 0 && has_p(t)) vpart += q;
        }
      }
    }
    const double v = -block_sum(vpart, red3);
    const double alpha = eta / v;
    // ---- phase B: ups = S p; lam += alpha p; r -= alpha ups; y = D r -----
    double nu_ = vget<NX>(YX, K.t(0) - 1, lane), nyxo = vget<NX>(YXO, K.t(0), lane);
    double ngb = vget<NX>(GB, K.t(0), lane), nr = vget<NX>(R, K.t(0), lane);
    double nlam = vget<NX>(LAM, K.t(0), lane), npb = vget<NX>(P, K.t(0), lane);
    for (int r = 0; r < K.rounds; ++r) {
      const int t = K.t(r);
      const bool ok = K.ok(r);
      const double u = nu_, cyxo = nyxo, cgb = ngb, cr = nr, clam = nlam, cpb = npb;
      {
        const int tn = K.t(r + 1);
        nu_ = vget<NX>(YX, tn - 1, lane);
        nyxo = vget<NX>(YXO, tn, lane);
        ngb = vget<NX>(GB, tn, lane);
        nr = vget<NX>(R, tn, lane);
        nlam = vget<NX>(LAM, tn, lane);
        npb = vget<NX>(P, tn, lane);
      }
      const double au = gdot_mem<G, NX>(arow(t), AS, u, lane);
      const double g = (t == 0 ? 1.0 : -1.0) * cyxo + ((ok && has_p(t)) ? au : 0.0) + cgb;
      double rv = cr;
      if (ok) {
        LAM[(t + 1) * NX + lane] = clam + alpha * cpb;
        rv += alpha * g;
      }
      const double y = dmul_mem<G, NX, SH>(w, t, N, rv, lane, ok);
      if (ok) {
        R[(t + 1) * NX + lane] = rv;
        Y[(t + 1) * NX + lane] = y;
        YQ[(t + 1) * NX + lane] = c.qinv_lane(t) * y;
      }
    }
    __syncthreads();
    // ---- phase C: rt = Phi^-1 r; eta' = r.rt --------------------------------
    double epart = 0.0;
    double nyq = vget<NX>(YQ, K.t(0) - 1, lane), ny = vget<NX>(Y, K.t(0) + 1, lane), nrc = vget<NX>(R, K.t(0), lane);
    for (int r = 0; r < K.rounds; ++r) {
      const int t = K.t(r);
      const bool ok = K.ok(r);
      const double cyq = nyq, cy = ny, crc = nrc;
      {
        const int tn = K.t(r + 1);
        nyq = vget<NX>(YQ, tn - 1, lane);
        ny = vget<NX>(Y, tn + 1, lane);
        nrc = vget<NX>(R, tn, lane);
      }
      const double a1v = gdot_mem<G, NX>(arow(t), AS, cyq, lane);
      const double a2v = gdot_mem<G, NX>(acol(t), NZ, cy, lane);
      const double a1 = (ok && has_p(t)) ? a1v : 0.0, a2 = (ok && has_a(t)) ? a2v : 0.0;
      const double rv = crc;
      const double qi = ok ? c.qinv_lane(t) : 0.0;
      const double z = ok ? rv - (t == 1 ? 1.0 : -1.0) * a1 - (t == 0 ? 1.0 : -1.0) * qi * a2 : 0.0;
      const double rt = dmul_mem<G, NX, SH>(w, t, N, z, lane, ok);
      if (ok) {
        epart += rv * rt;
        RT[(t + 1) * NX + lane] = rt;
      }
    }
    const double eta_new = block_sum(epart, red3 + 16);
    beta = eta_new / eta;
    eta = eta_new;
    ++it;
  }
  __syncthreads();
  *eta_out = eta;
  return it;
}

// x_t and u_{t-1} of block t (pcg.py:316-348) from V_LAM, see recover_rows.
template <int G, int NX, int NU, bool SH, class T>
__device__ __forceinline__ void stream_recover(const Blocks<G, NX, NU>& K, int r, const Work& w,
                                               const CostView<NX, NU>& c, const T* ABb, const T* ATb, double* xn,
                                               double* un) {
  constexpr int NZ = NX + NU;
  const int t = K.t(r), lane = K.lane, N = K.N;
  const bool ok = K.ok(r);
  const int li = lane < NX ? lane : NX - 1;  // (see dmul_mem)
  const double* LAM = w.v[V_LAM];
  const double lt = vget<NX>(LAM, t, lane);
  const double ln = vget<NX>(LAM, t + 1, lane);
  const double ax = gdot_mem<G, NX>(ABb + (size_t)shared_ab<SH>(t < N ? t : 0) * NX * NZ + li, NZ, ln, lane);
  constexpr int AS = arow_stride<NX>();
  const T* brow = arow_mem<NX, NU, SH>(ABb, ATb, t, li) + NX * AS;
  double bl = 0.0;
#pragma unroll
  for (int k = 0; k < NU; ++k) {
    const double sk = gsum<G>((double)brow[k * AS] * lt);
    bl = (lane == k) ? sk : bl;
  }
  *xn = 0.0;
  *un = 0.0;
  if (ok) {
    *xn = -c.qinv(t, lane) * (c.q(t, lane) + (t == 0 ? lt : -lt) + (t < N ? ax : 0.0));
    if (lane < NU && t >= 1) *un = -c.rinv(lane) * (c.r(t - 1, lane) + bl);
  }
}

__device__ __forceinline__ int ctl_row(int t, int N) { return t == 0 ? N : t - 1; }

// Clears the padded vectors and loads the multiplier guess.
template <int NX>
__device__ void stream_init_vectors(const Work& w, int Np, int N, const double* lam_guess) {
  const int n = (Np + 3) * NX;
  for (int k = 0; k < kStreamVectors; ++k)
    for (int e = threadIdx.x; e < n; e += blockDim.x) w.v[k][e] = 0.0;
  __syncthreads();
  for (int e = threadIdx.x; e < (N + 1) * NX; e += blockDim.x) w.v[V_LAM][NX + e] = lam_guess[e];
  __syncthreads();
}

// ---------------------------------------------------------------------------
// One SQP iteration of one instance (streamed): gamma, PCG, recovery, line
// search, step test and outputs.  The block matrices were assembled by
// stream_assemble_kernel for this linearization.
template <class Model, int G, int NV, bool SH, bool LS>
__global__ void __launch_bounds__(stream_max_threads<Model>())
__attribute__((amdgpu_waves_per_eu(stream_waves_per_simd<Model, SH, LS>()))) stream_qp_kernel(SolveArgs a) {
  constexpr int NX = Model::NX, NU = Model::NU, NZ = NX + NU;
  __shared__ double red3[48];
  extern __shared__ double stream_vlds[];
  const int b = a.order ? a.order[blockIdx.x] : blockIdx.x;
  if (!sqp_active(a, b)) return;
  const int it = a.num_iter[b];
  const int Np = a.N, NBs = Np + 1;
  const int N = horizon_of(a.horizons, Np, b), NB = N + 1;
  const int lsd = linesearch_doubles(a.ls_eval, a.n_alphas, Np);
  Work w;
  carve_work(a.work + (size_t)b * a.work_stride, NX, NU, Np, lsd, &w);
  vectors_in_lds<NV, NX>(&w, stream_vlds, Np);
  Blocks<G, NX, NU> K;
  K.init(N);
  CostView<NX, NU> c{a.wx + (size_t)b * NX, a.wu + (size_t)b * NU, a.wf + (size_t)b * NX,
                     a.xref + (size_t)b * NBs * NX, a.uref + (size_t)b * NBs * NU, a.xf + (size_t)b * NX,
                     a.penalize, N};
  __shared__ double rinv_lds[NU];
  cache_reciprocals(c, rinv_lds, K.lane);
  const double* xb = a.x + (size_t)b * NBs * NX;
  const double* ub = a.u + (size_t)b * NBs * NU;
  const float* ABb = a.AB + (size_t)b * Np * NX * NZ;
  const double* fdb = a.fd + (size_t)b * Np * NX;
  const double* x0 = a.x0 + (size_t)b * NX;
  stream_init_vectors<NX>(w, Np, N, a.lam + (size_t)b * NBs * NX);
  // gamma = C + G H^-1 q (pcg.py:245-288), C = [x0; -f_d + A xbar + B ubar]
  for (int r = 0; r < K.rounds; ++r) {
    const int t = K.t(r), lane = K.lane;
    if (!K.ok(r)) continue;
    double g;
    if (t == 0) {
      g = c.qinv(0, lane) * c.q(0, lane) + x0[lane];
    } else {
      const int tp = t - 1;
      g = -c.qinv(t, lane) * c.q(t, lane) - fdb[tp * NX + lane];
      const float* Ar = ABb + (size_t)tp * NX * NZ + lane * NZ;
      for (int k = 0; k < NX; ++k) g += (double)Ar[k] * (c.qinv(tp, k) * c.q(tp, k) + xb[tp * NX + k]);
      for (int k = 0; k < NU; ++k) g += (double)Ar[NX + k] * (c.rinv(k) * c.r(tp, k) + ub[tp * NU + k]);
    }
    w.v[V_GAM][(t + 1) * NX + lane] = g;
  }
  __syncthreads();
  const BlockT<SH>* ABk;
  const BlockT<SH>* ATk;
  if constexpr (SH) {
    stage_shared_blocks<NX, NU>(stream_vlds + (size_t)NV * (Np + 3) * NX, ABb, w, N, ABk, ATk);
  } else {
    ABk = ABb;
    ATk = w.AT;
  }
  double eta;
  const int npcg = stream_pcg<G, NX, NU, SH>(K, w, c, ABk, ATk, red3, a.pcg_eps, a.pcg_max, &eta);
  // recovery: the QP solution into XN / UN (free PCG vectors) and the step
  // to it into DX / DU; with a line search the new iterate is
  // xbar + alpha * step, otherwise the QP solution itself
  double* XN = w.v[V_P];
  double* UN = w.v[V_AP];  // (N+3) nx >= (N+1) nu doubles
  for (int r = 0; r < K.rounds; ++r) {
    double xn, un;
    stream_recover<G, NX, NU, SH>(K, r, w, c, ABk, ATk, &xn, &un);
    if (K.ok(r)) {
      const int t = K.t(r), lane = K.lane, tu = ctl_row(t, N);
      XN[t * NX + lane] = xn;
      w.DX[t * NX + lane] = xn - xb[t * NX + lane];
      if (lane < NU) {
        UN[tu * NU + lane] = un;
        w.DU[tu * NU + lane] = un - ub[tu * NU + lane];
      }
    }
  }
  __syncthreads();
  double alpha = 1.0;
  // LS: the build that evaluates the merit line search (a.ls_eval); the
  // other one has none of its float64 model evaluations and their registers
  if constexpr (LS) {
    alpha = line_search<Model>(a, b, N, xb, ub, w.DX, w.DU, w.LS, red3, threadIdx.x, blockDim.x);
    for (int k = threadIdx.x; k < NB * NX; k += blockDim.x) XN[k] = xb[k] + alpha * w.DX[k];
    for (int k = threadIdx.x; k < NB * NU; k += blockDim.x) UN[k] = ub[k] + alpha * w.DU[k];
  }
  double e = 0.0;
  for (int k = threadIdx.x; k < NB * NX; k += blockDim.x) e = nanmax(e, fabs(XN[k] - xb[k]));
  for (int k = threadIdx.x; k < NB * NU; k += blockDim.x) e = nanmax(e, fabs(UN[k] - ub[k]));
  const double err = block_max(e, red3);  // every read of the old iterate precedes this barrier
  for (int k = threadIdx.x; k < NB * NX; k += blockDim.x) {
    a.x[(size_t)b * NBs * NX + k] = XN[k];
    a.lam[(size_t)b * NBs * NX + k] = w.v[V_LAM][NX + k];
  }
  for (int k = threadIdx.x; k < NB * NU; k += blockDim.x) a.u[(size_t)b * NBs * NU + k] = UN[k];
  for (int k = NB * NX + threadIdx.x; k < NBs * NX; k += blockDim.x) {
    a.x[(size_t)b * NBs * NX + k] = 0.0;
    a.lam[(size_t)b * NBs * NX + k] = 0.0;
  }
  for (int k = NB * NU + threadIdx.x; k < NBs * NU; k += blockDim.x) a.u[(size_t)b * NBs * NU + k] = 0.0;
  for (int k = threadIdx.x; k < NB * NX; k += blockDim.x) a.qinv[(size_t)b * NBs * NX + k] = c.qinv(k / NX, k % NX);
  for (int k = threadIdx.x; k < NU; k += blockDim.x) a.rinv[(size_t)b * NU + k] = c.rinv(k);
  if (threadIdx.x == 0) {
    a.conv_errs[(size_t)b * a.maxit + it] = err;
    a.pcg_iters[(size_t)b * a.maxit + it] = npcg;
    if (a.linesearch) a.ls_alphas[(size_t)b * a.maxit + it] = alpha;
    a.conv_err[b] = err;
    a.num_iter[b] = it + 1;
  }
}

// Implicit-differentiation backward (sqp.py:605-798), streamed: adjoint PCG
// with q = -dl/dx, r = -dl/du[:-1], C = 0 on the saved linearization, then
// the closed-form contraction of the register variant's backward.
template <class Model, int G, int NV, bool SH>
__global__ void __launch_bounds__(stream_max_threads<Model>())
__attribute__((amdgpu_waves_per_eu(stream_waves_per_simd<Model, SH, false>()))) stream_bwd_kernel(BackwardArgs a) {
  constexpr int NX = Model::NX, NU = Model::NU, NZ = NX + NU;
  __shared__ double red3[48];
  extern __shared__ double stream_vlds[];
  const int b = a.order ? a.order[blockIdx.x * a.order_stride] : blockIdx.x;
  const int Np = a.N, NBs = Np + 1;
  const int N = horizon_of(a.horizons, Np, b), NB = N + 1;
  Work w;
  carve_work(a.work + (size_t)b * a.work_stride, NX, NU, Np, 0, &w);
  vectors_in_lds<NV, NX>(&w, stream_vlds, Np);
  Blocks<G, NX, NU> K;
  K.init(N);
  // cost view of the adjoint system: q~ = -dl/dx and r~ = -dl/du replace
  // q and r; Qinv and Rinv are those of the solve (weights unchanged)
  CostView<NX, NU> c{a.wx + (size_t)b * NX, a.wu + (size_t)b * NU, a.wf + (size_t)b * NX,
                     a.xref + (size_t)b * NBs * NX, a.uref + (size_t)b * NBs * NU, a.xf + (size_t)b * NX,
                     a.penalize, N};
  __shared__ double rinv_lds[NU];
  cache_reciprocals(c, rinv_lds, K.lane);  // visible after the barrier that follows gamma
  const double* dlx = a.dl_dx + (size_t)b * NBs * NX;
  const double* dlu = a.dl_du + (size_t)b * NBs * NU;
  const float* ABb = a.AB + (size_t)b * Np * NX * NZ;
  stream_init_vectors<NX>(w, Np, N, a.lam_bwd + (size_t)b * NBs * NX);
  // gamma = G H^-1 q~ (C = 0)
  for (int r = 0; r < K.rounds; ++r) {
    const int t = K.t(r), lane = K.lane;
    if (!K.ok(r)) continue;
    double g;
    if (t == 0) {
      g = c.qinv(0, lane) * -dlx[lane];
    } else {
      const int tp = t - 1;
      g = -c.qinv(t, lane) * -dlx[t * NX + lane];
      const float* Ar = ABb + (size_t)tp * NX * NZ + lane * NZ;
      for (int k = 0; k < NX; ++k) g += (double)Ar[k] * (c.qinv(tp, k) * -dlx[tp * NX + k]);
      for (int k = 0; k < NU; ++k) g += (double)Ar[NX + k] * (c.rinv(k) * -dlu[tp * NU + k]);
    }
    w.v[V_GAM][(t + 1) * NX + lane] = g;
  }
  __syncthreads();
  const BlockT<SH>* ABk;
  const BlockT<SH>* ATk;
  if constexpr (SH) {
    stage_shared_blocks<NX, NU>(stream_vlds + (size_t)NV * (Np + 3) * NX, ABb, w, N, ABk, ATk);
  } else {
    ABk = ABb;
    ATk = w.AT;
  }
  double eta;
  const int npcg = stream_pcg<G, NX, NU, SH>(K, w, c, ABk, ATk, red3, a.pcg_eps, a.pcg_max, &eta);
  // recovery of x~, u~ with q~, r~; per-row contraction terms into DX / DU
  const double* xs = a.x + (size_t)b * NBs * NX;
  const double* us = a.u + (size_t)b * NBs * NU;
  for (int r = 0; r < K.rounds; ++r) {
    const int t = K.t(r), lane = K.lane;
    const bool ok = K.ok(r);
    const int li = lane < NX ? lane : NX - 1;  // (see dmul_mem)
    const double* LAM = w.v[V_LAM];
    const double lt = vget<NX>(LAM, t, lane);
    const double ln = vget<NX>(LAM, t + 1, lane);
    const double ax = gdot_mem<G, NX>(ABk + (size_t)shared_ab<SH>(t < N ? t : 0) * NX * NZ + li, NZ, ln, lane);
    constexpr int AS = arow_stride<NX>();
    const BlockT<SH>* brow = arow_mem<NX, NU, SH>(ABk, ATk, t, li) + NX * AS;
    double bl = 0.0;
#pragma unroll
    for (int k = 0; k < NU; ++k) {
      const double sk = gsum<G>((double)brow[k * AS] * lt);
      bl = (lane == k) ? sk : bl;
    }
    if (!ok) continue;
    const double xt = -c.qinv(t, lane) * (-dlx[t * NX + lane] + (t == 0 ? lt : -lt) + (t < N ? ax : 0.0));
    const int i = lane;
    const double ref = (t < N) ? c.xref[t * NX + i] : c.xf[i];
    const double ex = xs[t * NX + i] - ref;
    w.DX[t * NX + i] = 2.0 * ex * xt;
    const double wxi = c.wx[i];
    a.g_xref[((size_t)b * NBs + t) * NX + i] = (t < N) ? 2.0 * wxi * xt : 0.0;
    if (t == N) {
      a.g_wf[b * NX + i] = -2.0 * ex * xt;
      a.g_xf[b * NX + i] = 2.0 * (wxi + c.wf[i]) * xt;
    }
    if (i < NU) {
      const int tu = ctl_row(t, N);
      const double ut = (t >= 1) ? -c.rinv(i) * (-dlu[(t - 1) * NU + i] + bl) : 0.0;
      const double ur = a.penalize ? c.uref[tu * NU + i] : 0.0;
      w.DU[tu * NU + i] = (tu < N) ? 2.0 * (us[tu * NU + i] - ur) * ut : 0.0;
      a.g_uref[((size_t)b * NBs + tu) * NU + i] = (tu < N && a.penalize) ? 2.0 * c.wu[i] * ut : 0.0;
      if (a.adj_u) a.adj_u[((size_t)b * NBs + tu) * NU + i] = (tu < N) ? ut : 0.0;
    }
    a.lam_bwd[((size_t)b * NBs + t) * NX + i] = lt;
    if (t == 0) a.g_x0[b * NX + i] = lt;
    if (a.adj_x) a.adj_x[((size_t)b * NBs + t) * NX + i] = xt;
  }
  for (int k = NB * NX + threadIdx.x; k < NBs * NX; k += blockDim.x) {
    a.g_xref[(size_t)b * NBs * NX + k] = 0.0;
    a.lam_bwd[(size_t)b * NBs * NX + k] = 0.0;
    if (a.adj_x) a.adj_x[(size_t)b * NBs * NX + k] = 0.0;
  }
  for (int k = NB * NU + threadIdx.x; k < NBs * NU; k += blockDim.x) {
    a.g_uref[(size_t)b * NBs * NU + k] = 0.0;
    if (a.adj_u) a.adj_u[(size_t)b * NBs * NU + k] = 0.0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NX + NU; i += blockDim.x) {
    double acc = 0.0;
    if (i < NX) {
      for (int t = 0; t <= N; ++t) acc += w.DX[t * NX + i];
      a.g_wx[b * NX + i] = -acc;
    } else {
      const int k = i - NX;
      for (int t = 0; t < N; ++t) acc += w.DU[t * NU + k];
      a.g_wu[b * NU + k] = -acc;
    }
  }
  if (threadIdx.x == 0) a.pcg_iters[b] = npcg;
}

// ---------------------------------------------------------------------------
// host side

// Threads per workgroup: enough waves for one round over the N + 1 blocks,
// capped at maxt.  Above 512 the rounds get longer (a round of 16 waves per
// CU takes about 1.6 times as long as one of 8 waves, quadrotor N = 64..200),
// so 512 threads are kept unless the wider workgroup cuts the rounds by more
// than that: 512 up to N = 80 for 16-lane blocks, 1024 from N = 96.
inline int stream_threads(int N, int G, int maxt) {
  const int want = (N + 1) * G;
  int t = ((want + 63) / 64) * 64;
  t = t < 64 ? 64 : (t > maxt ? maxt : t);
  if (t > 512) {
    const int r512 = (N + 1 + 512 / G - 1) / (512 / G);
    const int rt = (N + 1 + t / G - 1) / (t / G);
    if (r512 * 10 <= rt * 16) t = 512;
  }
  return t;
}

template <class Model>
AssembleArgs assemble_args(const SolveArgs& a, int lsd) {
  AssembleArgs s;
  s.B = a.B;
  s.N = a.N;
  s.horizons = a.horizons;
  s.AB = a.AB;
  s.wx = a.wx;
  s.wu = a.wu;
  s.wf = a.wf;
  s.num_iter = a.num_iter;
  s.conv_err = a.conv_err;
  s.maxit = a.maxit;
  s.tol = a.tol;
  s.work = a.work;
  s.work_stride = a.work_stride;
  s.ls_doubles = lsd;
  s.shared_blocks = a.shared_blocks;
  return s;
}

// One copy of the blocks for time-invariant models, unless the caller asked
// for the streamed variant with a copy per knot (fused variant 3).
template <class Model>
int shares_blocks(int variant) {
  const int v = variant > 0 ? variant : g_fused_variant;
  return time_invariant<Model>::value && v != 3 ? 1 : 0;
}

}  // namespace

size_t stream_work_stride(int nx, int nu, int N, int ls_eval, int n_alphas) {
  return carve_work(nullptr, nx, nu, N, linesearch_doubles(ls_eval, n_alphas, N), nullptr);
}

// the linearization and reset/finish kernels are shared with the register
// variant (dmpc_fused.hip)
int stream_launch_linearize(int model, const SolveArgs& a, hipStream_t st);
int stream_launch_reset(const SolveArgs& a, hipStream_t st);
const int32_t* stream_launch_order(const SolveArgs& a, int it, hipStream_t st);
bool stream_backward_order(const BackwardArgs& a, int32_t* order, size_t stride, hipStream_t st);
int stream_launch_finish(int model, const SolveArgs& a, hipStream_t st);

// How many PCG vectors the streamed kernels keep in LDS for padded horizon N:
// all of them, or the seven of the PCG phases, when they fit next to the
// kernels' static LDS; else none (the workspace holds them).  The first call
// per kernel raises its dynamic-LDS limit.
constexpr int kHotVectors = 7;
constexpr size_t kStreamLdsLimit = 160 * 1024 - 1024;

// Dynamic LDS limit of the LDS-vector kernels of one sharing flag, once.
template <class Model, int G, bool SH>
bool set_lds_limit(bool forward) {
  static bool attr[2] = {false, false};
  if (attr[forward]) return true;
  const void* k[6] = {
      forward ? reinterpret_cast<const void*>(&stream_qp_kernel<Model, G, kStreamVectors, SH, false>)
              : reinterpret_cast<const void*>(&stream_bwd_kernel<Model, G, kStreamVectors, SH>),
      forward ? reinterpret_cast<const void*>(&stream_qp_kernel<Model, G, kHotVectors, SH, false>)
              : reinterpret_cast<const void*>(&stream_bwd_kernel<Model, G, kHotVectors, SH>),
      forward ? reinterpret_cast<const void*>(&stream_qp_kernel<Model, G, 0, SH, false>)
              : reinterpret_cast<const void*>(&stream_bwd_kernel<Model, G, 0, SH>),
      forward ? reinterpret_cast<const void*>(&stream_qp_kernel<Model, G, kStreamVectors, SH, true>) : nullptr,
      forward ? reinterpret_cast<const void*>(&stream_qp_kernel<Model, G, kHotVectors, SH, true>) : nullptr,
      forward ? reinterpret_cast<const void*>(&stream_qp_kernel<Model, G, 0, SH, true>) : nullptr};
  for (const void* f : k)
    if (f && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kStreamLdsLimit) != hipSuccess)
      return false;
  attr[forward] = true;
  return true;
}

// Dynamic LDS of a streamed launch: its LDS vectors, and the staged shared
// blocks with sharing.
template <class Model>
size_t stream_lds_bytes(int nv, int N, bool shared) {
  return vector_lds_bytes(nv, Model::NX, N) + (shared ? shared_block_bytes<Model::NX, Model::NU>() : 0);
}

template <class Model, int G>
int lds_vectors(int N, bool forward, bool shared) {
  constexpr int NX = Model::NX;
  int nv = 0;
  // the recovery writes the (N+1) x nu controls into V_AP, which may then
  // run into the following vectors: all eleven in LDS only when nu <= nx
  if (Model::NU <= NX && stream_lds_bytes<Model>(kStreamVectors, N, shared) <= kStreamLdsLimit) nv = kStreamVectors;
  else if (stream_lds_bytes<Model>(kHotVectors, N, shared) <= kStreamLdsLimit) nv = kHotVectors;
  if (!nv && !shared) return 0;
  bool ok;
  if constexpr (time_invariant<Model>::value) {
    ok = shared ? set_lds_limit<Model, G, true>(forward) : set_lds_limit<Model, G, false>(forward);
  } else {
    ok = set_lds_limit<Model, G, false>(forward);
  }
  return ok ? nv : -1;
}

template <class Model, int G, bool SH, bool LS>
void launch_stream_qp_ls(const SolveArgs& a, int nv, int threads, size_t lds, hipStream_t st) {
  if (nv == kStreamVectors)
    hipLaunchKernelGGL((stream_qp_kernel<Model, G, kStreamVectors, SH, LS>), dim3(a.B), dim3(threads), lds, st, a);
  else if (nv == kHotVectors)
    hipLaunchKernelGGL((stream_qp_kernel<Model, G, kHotVectors, SH, LS>), dim3(a.B), dim3(threads), lds, st, a);
  else
    hipLaunchKernelGGL((stream_qp_kernel<Model, G, 0, SH, LS>), dim3(a.B), dim3(threads), lds, st, a);
}
template <class Model, int G, bool SH>
void launch_stream_qp(const SolveArgs& a, int nv, int threads, size_t lds, hipStream_t st) {
  if (a.ls_eval) launch_stream_qp_ls<Model, G, SH, true>(a, nv, threads, lds, st);
  else launch_stream_qp_ls<Model, G, SH, false>(a, nv, threads, lds, st);
}

template <class Model, int G, bool SH>
void launch_stream_bwd(const BackwardArgs& a, int nv, int threads, size_t lds, hipStream_t st) {
  if (nv == kStreamVectors)
    hipLaunchKernelGGL((stream_bwd_kernel<Model, G, kStreamVectors, SH>), dim3(a.B), dim3(threads), lds, st, a);
  else if (nv == kHotVectors)
    hipLaunchKernelGGL((stream_bwd_kernel<Model, G, kHotVectors, SH>), dim3(a.B), dim3(threads), lds, st, a);
  else
    hipLaunchKernelGGL((stream_bwd_kernel<Model, G, 0, SH>), dim3(a.B), dim3(threads), lds, st, a);
}

template <class Model, int G>
int stream_forward_t(int model, const SolveArgs& a0, hipStream_t st) {
  constexpr int NX = Model::NX, NU = Model::NU;
  SolveArgs a = a0;
  a.shared_blocks = shares_blocks<Model>(a.variant);
  const int lsd = linesearch_doubles(a.ls_eval, a.n_alphas, a.N);
  const AssembleArgs aa = assemble_args<Model>(a, lsd);
  const int threads = stream_threads(a.N, G, stream_max_threads<Model>());
  const int nv = lds_vectors<Model, G>(a.N, true, a.shared_blocks);
  if (nv < 0) return DMPC_ERR_HIP;
  const size_t lds = stream_lds_bytes<Model>(nv, a.N, a.shared_blocks);
  int rc = stream_launch_reset(a, st);
  if (rc) return rc;
  for (int it = 0; it < a.maxit; ++it) {
    rc = stream_launch_linearize(model, a, st);
    if (rc) return rc;
    SolveArgs ai = a;
    ai.order = stream_launch_order(a, it, st);
    prof_begin(PROF_QP_ITERATION, st);
    hipLaunchKernelGGL((stream_assemble_kernel<NX, NU>), dim3((unsigned)((long long)a.B * (a.N + 1))), dim3(64), 0,
                       st, aa);
    if constexpr (time_invariant<Model>::value) {
      if (a.shared_blocks) launch_stream_qp<Model, G, true>(ai, nv, threads, lds, st);
      else launch_stream_qp<Model, G, false>(ai, nv, threads, lds, st);
    } else {
      launch_stream_qp<Model, G, false>(ai, nv, threads, lds, st);
    }
    prof_end(PROF_QP_ITERATION, st);
  }
  rc = stream_launch_finish(model, a, st);
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? DMPC_OK : DMPC_ERR_HIP;
}

template <class Model, int G>
int stream_backward_t(const BackwardArgs& a0, hipStream_t st) {
  constexpr int NX = Model::NX, NU = Model::NU;
  BackwardArgs a = a0;
  a.shared_blocks = shares_blocks<Model>(a.variant);
  AssembleArgs aa;
  aa.B = a.B;
  aa.N = a.N;
  aa.horizons = a.horizons;
  aa.AB = a.AB;
  aa.wx = a.wx;
  aa.wu = a.wu;
  aa.wf = a.wf;
  aa.num_iter = nullptr;
  aa.conv_err = nullptr;
  aa.maxit = 0;
  aa.tol = 0.0;
  aa.work = a.work;
  aa.work_stride = a.work_stride;
  aa.ls_doubles = 0;
  aa.shared_blocks = a.shared_blocks;
  // ragged batches: longest-first order in the instances' order slots
  BackwardArgs ab = a;
  int32_t* slots = reinterpret_cast<int32_t*>(a.work + a.work_stride - kOrderSlot);
  if (stream_backward_order(a, slots, a.work_stride / sizeof(int32_t), st)) {
    ab.order = slots;
    ab.order_stride = a.work_stride / sizeof(int32_t);
  }
  prof_begin(PROF_BACKWARD, st);
  hipLaunchKernelGGL((stream_assemble_kernel<NX, NU>), dim3((unsigned)((long long)a.B * (a.N + 1))), dim3(64), 0, st,
                     aa);
  const int nv = lds_vectors<Model, G>(a.N, false, a.shared_blocks);
  if (nv < 0) return DMPC_ERR_HIP;
  const size_t lds = stream_lds_bytes<Model>(nv, a.N, a.shared_blocks);
  const int threads = stream_threads(a.N, G, stream_max_threads<Model>());
  if constexpr (time_invariant<Model>::value) {
    if (a.shared_blocks) launch_stream_bwd<Model, G, true>(ab, nv, threads, lds, st);
    else launch_stream_bwd<Model, G, false>(ab, nv, threads, lds, st);
  } else {
    launch_stream_bwd<Model, G, false>(ab, nv, threads, lds, st);
  }
  prof_end(PROF_BACKWARD, st);
  return hipGetLastError() == hipSuccess ? DMPC_OK : DMPC_ERR_HIP;
}

bool stream_supported(int model, int nx, int nu) {
  switch (model) {
    case MODEL_LINEAR: return (nx == 4 && nu == 2) || (nx == 8 && nu == 4);
    case MODEL_QUADROTOR: return nx == 13 && nu == 4;
    case MODEL_SPACECRAFT3: return nx == 3 && nu == 3;
    case MODEL_SPACECRAFT7: return nx == 7 && nu == 3;
    case MODEL_CARTPOLE: return nx == 4 && nu == 1;
    case MODEL_CHAIN: return nx == 48 && nu == 12;
    DMPC_CUSTOM_CASE(return nx == CustomModel::NX && nu == CustomModel::NU && CustomModel::NX <= 64)
    default: return false;
  }
}

int stream_forward(int model, int nx, int nu, const SolveArgs& a, hipStream_t st) {
  if (!stream_supported(model, nx, nu)) return DMPC_ERR_UNSUPPORTED;
  if (!a.work) return DMPC_ERR_ARG;
  switch (model) {
    case MODEL_LINEAR:
      if (nx == 8) return stream_forward_t<LinearModel<8, 4>, 8>(model, a, st);
      return stream_forward_t<LinearModel<4, 2>, 4>(model, a, st);
    case MODEL_QUADROTOR: return stream_forward_t<QuadrotorModel, 16>(model, a, st);
    case MODEL_SPACECRAFT3: return stream_forward_t<Spacecraft3Model, 4>(model, a, st);
    case MODEL_SPACECRAFT7: return stream_forward_t<Spacecraft7Model, 8>(model, a, st);
    case MODEL_CARTPOLE: return stream_forward_t<CartpoleModel, 4>(model, a, st);
    case MODEL_CHAIN: return stream_forward_t<ChainModel, 64>(model, a, st);
    DMPC_CUSTOM_CASE(return stream_forward_t<CustomModel, kCustomG>(model, a, st))
    default: return DMPC_ERR_UNSUPPORTED;
  }
}

int stream_backward(int model, int nx, int nu, const BackwardArgs& a, hipStream_t st) {
  if (!stream_supported(model, nx, nu)) return DMPC_ERR_UNSUPPORTED;
  if (!a.work) return DMPC_ERR_ARG;
  switch (model) {
    case MODEL_LINEAR:
      if (nx == 8) return stream_backward_t<LinearModel<8, 4>, 8>(a, st);
      return stream_backward_t<LinearModel<4, 2>, 4>(a, st);
    case MODEL_QUADROTOR: return stream_backward_t<QuadrotorModel, 16>(a, st);
    case MODEL_SPACECRAFT3: return stream_backward_t<Spacecraft3Model, 4>(a, st);
    case MODEL_SPACECRAFT7: return stream_backward_t<Spacecraft7Model, 8>(a, st);
    case MODEL_CARTPOLE: return stream_backward_t<CartpoleModel, 4>(a, st);
    case MODEL_CHAIN: return stream_backward_t<ChainModel, 64>(a, st);
    DMPC_CUSTOM_CASE(return stream_backward_t<CustomModel, kCustomG>(a, st))
    default: return DMPC_ERR_UNSUPPORTED;
  }
}

}  // namespace dmpc
