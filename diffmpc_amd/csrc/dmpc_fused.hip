// Fused batched SQP solve and adjoint backward for CDNA4 (gfx950).
//
// One SQP iteration of the batch (diffmpc/solvers/sqp.py:262-299) is two
// launches: linearize_kernel (one thread per instance, step and tangent) and
// qp_iteration_kernel (one workgroup per instance: Schur/preconditioner
// assembly, the PCG loop of pcg.py:350-445 and primal recovery).  All
// num_scp_iteration_max iterations are enqueued back to back with no host
// round trip; an instance that has met the stop rule of sqp.py:460-466 skips
// its work in later launches.  fused_backward_kernel is the adjoint solve.
//
// Work mapping.  Time block t (0..N) of the Schur system is owned by a group
// of G lanes (G = next power of two >= nx, so a block never straddles a DPP
// row or a wavefront); lane i of the group owns row i of the block.  A thread
// group owns K blocks t = s * NG + group (s < K).  Products with a block's
// matrices keep one matrix row per lane in registers and broadcast the vector
// entries inside the group with DPP; each lane reads only its own entry of a
// neighbour block's vector from LDS, so one PCG iteration costs three
// workgroup barriers and about six LDS accesses per lane.
//
// Numerics.  Dynamics are integrated in float32 with forward-mode dual
// numbers (one tangent per lane).  Everything of the QP is float64: the
// Schur complement S = -G H^-1 G^T is applied matrix-free from the float32
// A_t, B_t held in registers, and the block-Jacobi-stair preconditioner
//   Phi^-1 r |_t = D_t (r_t - phi_t D_{t-1} r_{t-1} - phi_{t+1}^T D_{t+1} r_{t+1}),
//   D_t = theta_t^-1,
// which equals the explicit Phiinv rows of pcg.py:213-239, is applied in
// factored form with D_t (float64) in registers (nx <= 8) or LDS (nx > 8).
// As_next = -I and the diagonal cost Hessians of
// optimal_control_problem.py:301-322 are used in closed form.
#include <type_traits>

#include "dmpc_device.h"
#include "dmpc_internal.h"
#include "dmpc_linesearch.h"
#include "dmpc_models.h"

namespace dmpc {

int g_forced_K = 0;

// streamed variant (dmpc_stream.hip)
bool stream_supported(int model, int nx, int nu);
size_t stream_work_stride(int nx, int nu, int N, int ls_eval, int n_alphas);
int stream_forward(int model, int nx, int nu, const SolveArgs& a, hipStream_t st);
int stream_backward(int model, int nx, int nu, const BackwardArgs& a, hipStream_t st);

namespace {

// Where D_t = theta_t^-1 lives during PCG: in registers (row `lane` per lane)
// for small blocks, in LDS for nx > 8, where the registers are needed for the
// A/B rows.  The LDS copy is stored transposed, DT[t][j][i] = D_t[i][j] as
// computed by lane i, so the lanes of a group read consecutive words.
template <int NX>
__host__ __device__ constexpr bool d_in_lds() { return NX > 8; }

// Whether the block's triangular solves for D_t run in registers (4- and
// 16-lane groups) or against the factor parked in LDS (8-lane groups).
template <int NX>
__host__ __device__ constexpr bool solve_in_registers() { return !(NX > 4 && NX <= 8); }

// LDS carve-up.  Vectors over time blocks use stride NX per block; arrays
// marked "padded" have a zero block before t = 0 and after t = N, so
// neighbour reads need no branches: index (t + 1) * NX + i.
template <int NX, int NU>
struct LdsPlan {
  static __host__ __device__ size_t doubles(int N) {
    const size_t NB = N + 1, NBP = N + 3;
    return NB * NX /*xb*/ + NB * NU /*ub*/ + NBP * NX /*lamS*/ + NB * NX /*qinv*/ + NB * NX /*qv*/ +
           NB * NU /*rv*/ + NU /*rinv*/ + NX /*x0*/ + 3 * 64 /*red*/ + 4 * NBP * NX /*RT Y YQ YX*/ +
           (d_in_lds<NX>() || !solve_in_registers<NX>() ? NB * NX * NX : 0) /*LF*/ +
           (solve_in_registers<NX>() ? 0 : NB * NX) /*ID*/ + (size_t)N * NX /*fd*/;
  }
  static __host__ __device__ size_t floats(int N, int P) { return (size_t)P; }
  // doubles(N) plus `ls` doubles of line-search scratch
  static __host__ size_t bytes(int N, int P, int ls = 0) {
    return (doubles(N) + ls) * 8 + floats(N, P) * 4 + 16;
  }
};

struct Lds {
  double *xb, *ub, *lamS, *qinv, *qv, *rv, *rinv, *x0, *redA, *redB, *redC;
  double *RT, *Y, *YQ, *YX;
  double* LF;       // D_t = theta_t^-1, transposed (nx > 8); 8-lane groups: factor rows
  double* ID;       // 8-lane groups: 1 / L_kk
  double* fd;       // f_d(x_t, u_t), t < N
  double* LS;       // line-search scratch (see linesearch_doubles)
  float* prm;
};

template <int NX, int NU>
__device__ Lds carve(void* base, int N, int ls = 0) {
  const int NB = N + 1, NBP = N + 3;
  double* d = reinterpret_cast<double*>(base);
  Lds L;
  L.xb = d; d += NB * NX;
  L.ub = d; d += NB * NU;
  L.lamS = d; d += NBP * NX;
  L.qinv = d; d += NB * NX;
  L.qv = d; d += NB * NX;
  L.rv = d; d += NB * NU;
  L.rinv = d; d += NU;
  L.x0 = d; d += NX;
  L.redA = d; d += 64;  // 64 doubles each: block_sum_rows (one partial per 16-lane row)
  L.redB = d; d += 64;
  L.redC = d; d += 64;
  L.RT = d; d += NBP * NX;
  L.Y = d; d += NBP * NX;
  L.YQ = d; d += NBP * NX;
  L.YX = d; d += NBP * NX;
  L.LF = d; d += d_in_lds<NX>() || !solve_in_registers<NX>() ? NB * NX * NX : 0;
  L.ID = d; d += solve_in_registers<NX>() ? 0 : NB * NX;
  L.fd = d; d += N * NX;
  L.LS = d; d += ls;
  L.prm = reinterpret_cast<float*>(d);
  return L;
}


// Instances per workgroup.  The register-resident one-block-per-group
// kernels can hold `ipw` instances per workgroup as teams of `tt` threads, tt
// a multiple of 16 so that no DPP row holds two instances: team j of
// workgroup w solves launch slot w * ipw + j and owns the j-th LDS region.
// Without teams (TEAMS = false) the workgroup is the single team.  Inside a
// team, thread `tid` plays the role threadIdx.x plays in a one-instance
// workgroup: same lane groups, same reduction rows, same sums.
struct Team {
  int tid;   // thread index inside the team
  int nt;    // threads of the team
  int idx;   // team index inside the workgroup
  int slot;  // launch slot (instance order[slot], or slot)
};

// Waves per SIMD the kernels are compiled for: the launch bounds' own floor
// without teams; five (at most 96 VGPRs) with teams.  The per-thread team
// bases of the LDS arrays cost the team kernels about 20 VGPRs, which at
// their 128-VGPR bound gives four waves per SIMD, so that a CU held no more
// instances than without teams; at five, a CU holds 12 instances of 84 lanes
// (two per workgroup) instead of 10 (spills stay outside the PCG loop).
template <bool TEAMS>
__host__ __device__ constexpr int team_waves_per_simd() { return TEAMS ? 5 : 1; }

template <bool TEAMS>
__device__ __forceinline__ Team team_of(int ipw, int tt) {
  if constexpr (TEAMS) {
    const int j = threadIdx.x / tt;
    return Team{(int)threadIdx.x - j * tt, tt, j, (int)blockIdx.x * ipw + j};
  } else {
    return Team{(int)threadIdx.x, (int)blockDim.x, 0, (int)blockIdx.x};
  }
}

// 16-lane groups with four controls keep B's rows column-permuted per quad
// of lanes and form the B-terms by reduce-scatter (rs16/gather16); other
// shapes use one group all-reduce per column.
template <int G, int NU>
__host__ __device__ constexpr bool quad_bterm() { return G == 16 && NU == 4; }

// Per-thread register state of the K blocks it owns.
template <int NX, int NU, int K>
struct Regs {
  // A/B entries are float32 values; with one or two blocks per thread group
  // they are kept as float64 (no conversion in the PCG loop), with four as
  // float32
  using AT = typename std::conditional<K <= 2, double, float>::type;
  AT acol[K][NX];   // column `lane` of A_t            (zero for t = N)
  AT arow[K][NX];   // row `lane` of A_{t-1}           (zero for t = 0)
  AT brow[K][NU];   // row `lane` of B_{t-1}           (zero for t = 0;
                    // slot r = column quad16(lane) ^ r with quad_bterm)
                       // (B_{t-1} terms of S are formed by block t alone)
  double drow[K][d_in_lds<NX>() ? 1 : NX];  // row `lane` of D_t = theta_t^-1 (registers mode)
  double qi[K];        // Qinv_t[lane, lane]
  int t[K];
  bool ok[K];          // t <= N and lane < nx
};

// Load A/B rows and columns of the saved linearization into registers.
template <int G, int NX, int NU, int K>
__device__ __forceinline__ void load_ab(Regs<NX, NU, K>& R, const float* ABb, const Lds& L, int N, int lane) {
  constexpr int NZ = NX + NU;
  const int bq = quad_bterm<G, NU>() ? quad16(lane) : 0;  // column permutation of the B row
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int t = R.t[s];
    const bool ok = R.ok[s];
    // clamped to blocks and rows that exist, so that every load is issued
    // unconditionally in one batch; lanes without the block discard them
    const int li = lane < NX ? lane : 0;
    const float* At = ABb + (size_t)(t < N ? t : N - 1) * NX * NZ;
    const float* Ap = ABb + (size_t)(t >= 1 ? t - 1 : 0) * NX * NZ;
    float ac[NX], ar[NX], br[NU];
#pragma unroll
    for (int j = 0; j < NX; ++j) ac[j] = At[j * NZ + li];
#pragma unroll
    for (int k = 0; k < NX; ++k) ar[k] = Ap[li * NZ + k];
#pragma unroll
    for (int k = 0; k < NU; ++k) br[k] = Ap[li * NZ + NX + (k ^ bq)];
#pragma unroll
    for (int j = 0; j < NX; ++j) R.acol[s][j] = (ok && t < N) ? ac[j] : 0.0f;
#pragma unroll
    for (int k = 0; k < NX; ++k) R.arow[s][k] = (ok && t >= 1) ? ar[k] : 0.0f;
#pragma unroll
    for (int k = 0; k < NU; ++k) R.brow[s][k] = (ok && t >= 1) ? br[k] : 0.0f;
    R.qi[s] = ok ? L.qinv[t * NX + lane] : 0.0;
  }
}

// Row `lane` of theta_t = A_{t-1} Qinv_{t-1} A_{t-1}^T + B_{t-1} Rinv
// B_{t-1}^T + Qinv_t for 16-lane groups of 13 states, from the lanes' own
// rows of [A_{t-1} | B_{t-1}] (one load of the row per lane): term k of entry
// j is A[j][k] * (A[lane][k] q_k), with A[j][k] read from lane j by DPP
// row_newbcast:j inside the FMA.  k runs outermost, so each of the 13
// accumulators receives the same products in the same order (A terms k =
// 0..12, then the B terms) as the entry-by-entry form, which reads every
// A[j][k] from memory instead (221 loads per lane and block).  Lanes >= nx
// and blocks without a predecessor carry zero rows; masked lanes return the
// identity row.
template <int NU>
__device__ __forceinline__ void theta_row16x13(const float* Ap, const double* qprev, const double* rinv, double qi,
                                               bool has_prev, bool ok, int lane, double (&row)[13]) {
  constexpr int NX = 13, NZ = NX + NU;
  double acc[NX];
#pragma unroll
  for (int j = 0; j < NX; ++j) acc[j] = (j == lane) ? qi : 0.0;
  // the whole row in one batch of loads (Ap is a valid block for every lane;
  // lanes without a predecessor discard the values)
  const float* own = Ap + (size_t)(lane < NX ? lane : 0) * NZ;
  float ab[NZ];
#pragma unroll
  for (int k = 0; k < NZ; ++k) ab[k] = own[k];
#define DMPC_TH(j) "v_fmac_f64_dpp %" #j ", %13, %14 row_newbcast:" #j " row_mask:0xf bank_mask:0xf\n"
#define DMPC_TH_BLOCK(src, scale)                                                                            \
  asm("s_nop 4\n" DMPC_TH(0) DMPC_TH(1) DMPC_TH(2) DMPC_TH(3) DMPC_TH(4) DMPC_TH(5) DMPC_TH(6) DMPC_TH(7)    \
          DMPC_TH(8) DMPC_TH(9) DMPC_TH(10) DMPC_TH(11) DMPC_TH(12)                                          \
      : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]),     \
        "+v"(acc[7]), "+v"(acc[8]), "+v"(acc[9]), "+v"(acc[10]), "+v"(acc[11]), "+v"(acc[12])                 \
      : "v"(src), "v"(scale))
#pragma unroll
  for (int k = 0; k < NX; ++k) {
    const double a = has_prev ? (double)ab[k] : 0.0;
    const double aq = a * qprev[k];
    DMPC_TH_BLOCK(a, aq);
  }
#pragma unroll
  for (int k = 0; k < NU; ++k) {
    const double bk = has_prev ? (double)ab[NX + k] : 0.0;
    const double bq = bk * rinv[k];
    DMPC_TH_BLOCK(bk, bq);
  }
#undef DMPC_TH_BLOCK
#undef DMPC_TH
#pragma unroll
  for (int j = 0; j < NX; ++j) row[j] = ok ? acc[j] : ((j == lane) ? 1.0 : 0.0);
}

// theta_t row `lane`, its Cholesky factor and D_t = theta_t^-1 (row `lane`),
// float64, inside the lane group of block t:
//   theta_0 = Qinv_0,
//   theta_t = A_{t-1} Qinv_{t-1} A_{t-1}^T + B_{t-1} Rinv B_{t-1}^T + Qinv_t
// (pcg.py:182-196 with As_next = -I).  The factor is computed in registers
// with cross-lane broadcasts, and each lane then solves L L^T d = e_lane for
// its row of D_t, also in registers.
template <int G, int NX, int NU, int K>
__device__ __forceinline__ void assemble_dinv(Regs<NX, NU, K>& R, const float* ABb, const Lds& L, int lane,
                                              int gbase) {
  constexpr int NZ = NX + NU;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int t = R.t[s];
    const bool ok = R.ok[s];
    const bool has_prev = ok && t >= 1;
    const int tp = t >= 1 ? t - 1 : 0;
    const float* Ap = ABb + (size_t)tp * NX * NZ;  // [A_{t-1} | B_{t-1}], L1-cached broadcast reads
    double row[NX];
    if constexpr (G == 16 && NX == 13) {
      theta_row16x13<NU>(Ap, L.qinv + tp * NX, L.rinv, ok ? L.qinv[t * NX + lane] : 0.0, has_prev, ok, lane, row);
    } else {
      double aq[NX], bq[NU];
#pragma unroll
      for (int k = 0; k < NX; ++k) aq[k] = has_prev ? (double)Ap[lane * NZ + k] * L.qinv[tp * NX + k] : 0.0;
#pragma unroll
      for (int k = 0; k < NU; ++k) bq[k] = has_prev ? (double)Ap[lane * NZ + NX + k] * L.rinv[k] : 0.0;
      const double qi = ok ? L.qinv[t * NX + lane] : 0.0;
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        double acc = (j == lane) ? qi : 0.0;
        if (has_prev) {
#pragma unroll
          for (int k = 0; k < NX; ++k) acc += aq[k] * (double)Ap[j * NZ + k];
#pragma unroll
          for (int k = 0; k < NU; ++k) acc += bq[k] * (double)Ap[j * NZ + NX + k];
        }
        row[j] = ok ? acc : ((j == lane) ? 1.0 : 0.0);
      }
    }
    // right-looking Cholesky: lane i ends with row[k] = L[i][k], k <= i;
    // pivots and column entries reach the group's lanes by DPP broadcast;
    // id[k] = 1 / L_kk is the same in every lane of the group
    double id[NX];
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      const double dkk = gbcast_at<G>(row[k], k, lane);
      const double lkk = sqrt(dkk);
      const double il = 1.0 / lkk;
      id[k] = il;
      const double lik = (lane == k) ? lkk : row[k] * il;
      row[k] = lik;
#pragma unroll
      for (int j = k + 1; j < NX; ++j) {
        const double ljk = gbcast_at<G>(lik, j, lane);
        row[j] -= lik * ljk;
      }
    }
    if constexpr (solve_in_registers<NX>()) {
      // column `lane` of theta^-1 = L^-T L^-1 e_lane by forward and back
      // substitution in registers: L[j][k] is entry k of lane j's row,
      // broadcast by DPP, so no factor goes through LDS.  Forward column by
      // column, back substitution row by row; each unknown's terms are
      // subtracted in ascending index order, as the substitution through
      // LDS for 8-lane groups does, with the same roundings.
      double x[NX];
#pragma unroll
      for (int j = 0; j < NX; ++j) x[j] = (j == lane) ? 1.0 : 0.0;
#pragma unroll
      for (int k = 0; k < NX; ++k) {
        x[k] *= id[k];
        const double nx_k = -x[k];
#pragma unroll
        for (int j = k + 1; j < NX; ++j) bcast_fmac_at<G>(x[j], row[k], nx_k, j, lane);
      }
#pragma unroll
      for (int j = NX - 1; j >= 0; --j) {
        double acc = x[j];
#pragma unroll
        for (int k = j + 1; k < NX; ++k) bcast_fmac_at<G>(acc, row[j], -x[k], k, lane);
        x[j] = acc * id[j];
      }
      if constexpr (d_in_lds<NX>()) {
        // transposed, DT[t][j][i] = D_t[i][j], so a group reads consecutive
        // words; no other lane reads this block before the barrier below
        if (ok) {
#pragma unroll
          for (int j = 0; j < NX; ++j) L.LF[((size_t)t * NX + j) * NX + lane] = x[j];
        }
      } else {
#pragma unroll
        for (int j = 0; j < NX; ++j) R.drow[s][j] = ok ? x[j] : 0.0;
      }
    } else {
      // 8-lane groups: the factor rows and 1 / L_kk go through LDS and each
      // lane substitutes against them (their two-move DPP broadcasts would
      // hold more registers than the 8-lane kernels have to spare)
      if (ok) {
#pragma unroll
        for (int k = 0; k < NX; ++k) L.LF[(t * NX + lane) * NX + k] = row[k];
        if (lane == 0) {
#pragma unroll
          for (int k = 0; k < NX; ++k) L.ID[t * NX + k] = id[k];
        }
      }
    }
  }
  if constexpr (!solve_in_registers<NX>()) {
    __syncthreads();
#pragma unroll
    for (int s = 0; s < K; ++s) {
      const int t = R.t[s];
      const double* LF = L.LF + (size_t)t * NX * NX;
      const double* ID = L.ID + t * NX;
      // column `lane` of theta^-1 = L^-T L^-1 e_lane
      double w[NX], d[NX];
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        const int rj = j == 0 ? 0 : after(j * NX, w[j - 1]);  // row j of L
        double acc = (j == lane) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < j; ++k) acc -= LF[rj + k] * w[k];
        w[j] = acc * ID[j];
      }
#pragma unroll
      for (int j = NX - 1; j >= 0; --j) {
        const int cj = j == NX - 1 ? j : after(j, d[j + 1]);  // column j of L
        double acc = w[j];
#pragma unroll
        for (int k = j + 1; k < NX; ++k) acc -= LF[k * NX + cj] * d[k];
        d[j] = acc * ID[j];
      }
#pragma unroll
      for (int j = 0; j < NX; ++j) R.drow[s][j] = R.ok[s] ? d[j] : 0.0;
    }
  }
  if constexpr (d_in_lds<NX>()) __syncthreads();
}

// D_t v for block t: sum_j D_t[lane][j] v_j with v_j in lane j of the group.
// Zero on masked lanes.
template <int G, int NX, int NU, int K>
__device__ __forceinline__ double dmul(const Regs<NX, NU, K>& R, int s, const Lds& L, double v, int lane) {
  double acc;
  if constexpr (d_in_lds<NX>()) {
    const double* col = L.LF + (size_t)R.t[s] * NX * NX + (lane < NX ? lane : 0);
    if constexpr (G == 16 && NX == 13) {
      double r[NX];
#pragma unroll
      for (int j = 0; j < NX; ++j) r[j] = col[j * NX];
      acc = gdot16x13(r, v, 0.0);
    } else {
      acc = gdot_strided<G, NX, NX>(col, v, lane);
    }
  } else {
    acc = gdotd<G, NX>(R.drow[s], v, lane);
  }
  return R.ok[s] ? acc : 0.0;
}

// Two chains of a thread run as one interleaved pair (gdot16x13_pair) when
// both are 13-term float64 rows of 16-lane groups.
template <int G, int NX, class AT>
__host__ __device__ constexpr bool pairable() {
  return G == 16 && NX == 13 && std::is_same<AT, double>::value;
}

// (oa, ob) = (gdot(ma, va), gdot(mb, vb)) for two products that do not
// depend on each other; the same arithmetic either way.
template <int G, int NX, class AT>
__device__ __forceinline__ void gdot_two(const AT* ma, double va, const AT* mb, double vb, int lane, int salt,
                                         double& oa, double& ob) {
  if constexpr (pairable<G, NX, AT>()) {
    oa = 0.0;
    ob = 0.0;
    gdot16x13_pair(ma, va, mb, vb, oa, ob);
  } else {
    oa = gdot<G, NX>(ma, va, lane, salt);
    ob = gdot<G, NX>(mb, vb, lane, salt);
  }
}

// out[s] = gdot(M[s], v[s]) for the K blocks of the thread (two blocks as
// one interleaved pair).
template <int G, int NX, int K, class AT>
__device__ __forceinline__ void gdot_k(const AT (&M)[K][NX], const double (&v)[K], int lane, int salt,
                                       double (&out)[K]) {
  if constexpr (K == 2 && pairable<G, NX, AT>()) {
    gdot_two<G, NX>(M[0], v[0], M[1], v[1], lane, salt, out[0], out[1]);
  } else {
#pragma unroll
    for (int s = 0; s < K; ++s) out[s] = gdot<G, NX>(M[s], v[s], lane, salt);
  }
}

// y[s] = D_{t_s} v[s] for the K blocks of the thread, zero on masked lanes;
// two blocks with D in LDS read their columns in two halves and run the
// products as one interleaved pair (the same arithmetic as dmul).
template <int G, int NX, int NU, int K>
__device__ __forceinline__ void dmul_k(const Regs<NX, NU, K>& R, const Lds& L, const double (&v)[K], int lane,
                                       double (&y)[K]) {
  if constexpr (K == 2 && d_in_lds<NX>() && G == 16 && NX == 13) {
    const int li = lane < NX ? lane : 0;
    const double* c0 = L.LF + (size_t)R.t[0] * NX * NX + li;
    const double* c1 = L.LF + (size_t)R.t[1] * NX * NX + li;
    double a0 = 0.0, a1 = 0.0;
    {
      double r0[7], r1[7];
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        r0[j] = c0[j * NX];
        r1[j] = c1[j * NX];
      }
      gdot16x13_pair_lo(r0, v[0], r1, v[1], a0, a1);
    }
    {
      // the second half is addressed through the first half's result, so its
      // loads are not hoisted above the first half (register pressure)
      const int o = after(7 * NX, a0);
      double r0[6], r1[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        r0[j] = c0[o + j * NX];
        r1[j] = c1[o + j * NX];
      }
      gdot16x13_pair_hi(r0, v[0], r1, v[1], a0, a1);
    }
    y[0] = R.ok[0] ? a0 : 0.0;
    y[1] = R.ok[1] ? a1 : 0.0;
  } else {
#pragma unroll
    for (int s = 0; s < K; ++s) y[s] = dmul<G>(R, s, L, v[s], lane);
  }
}

// Entry `lane` of block tb of an LDS vector (padded layout).  Lanes >= nx
// read entries of the next block (or, past the last padding block, the first
// words of the next LDS array): every caller uses their values only as DPP
// broadcast sources, which read lanes < nx, or multiplied by the lane's zero
// matrix rows, so the read needs no EXEC mask and no address select.
template <int NX>
__device__ __forceinline__ double own(const double* V, int tb, int lane) {
  return V[(tb + 1) * NX + lane];
}

// B_{t-1} Rinv B_{t-1}^T v_t for block t from the lane's own entry v of v_t:
// the NU components of B_{t-1}^T v_t are group sums.  Also returns
// (B^T v) . Rinv (B^T v) (identical in every lane of the group).
template <int G, int NX, int NU, int K>
__device__ __forceinline__ double bgram(const Regs<NX, NU, K>& R, int s, const Lds& L, double v, int salt,
                                        double* quad) {
  if constexpr (quad_bterm<G, NU>()) {
    // the quadratic form is accumulated by the caller as v . g over the lanes
    const int lane = threadIdx.x & 15;
    const double y = L.rinv[quad16(lane)] * rs16(R.brow[s], v, salt);
    *quad = 0.0;
    return gather16(R.brow[s], y, salt);
  }
  double g = 0.0, q = 0.0;
#pragma unroll
  for (int k = 0; k < NU; ++k) {
    const double bk = cvt_in_loop(R.brow[s][k], salt);
    const double sk = gsum<G>(bk * v);
    const double yk = L.rinv[k] * sk;
    g += bk * yk;
    q += sk * yk;
  }
  *quad = q;
  return g;
}

// Block-Jacobi-preconditioned CG on S lam = gamma (pcg.py:350-445) with the
// reference's stop rule: at least one iteration, then continue while
// |eta| > eps and it <= max_iter - 1.  lam enters as the guess in L.lamS
// (padded), is updated there in place and is also returned in lam[] (the
// caller synchronises before other lanes read L.lamS).  Returns the
// iteration count.
//
// With several instances per workgroup (TEAMS) the loop runs until every
// team has stopped: a team that has met its stop rule (or is not `live`)
// keeps passing the barriers with its multipliers, residual, eta and count
// frozen, so each team's results and count are those of its own solve.
// Each iteration's team flags go through flags[] with the first workgroup
// sum's barrier.
//
// S = -G H^-1 G^T is applied matrix-free.  Row block t of G H^-1 G^T p is
//   (t == 0 ? 1 : -1) Qinv_t w_t + A_{t-1} Qinv_{t-1} w_{t-1} + B_{t-1} Rinv B_{t-1}^T p_t
// with w_t = (t == 0 ? p_t : -p_t) + A_t^T p_{t+1}.  Each lane loads only its
// own entry of a neighbour block's vector from LDS; products with a block's
// matrices broadcast the entries inside the lane group with DPP, and the B
// term is local to the block.  Three workgroup barriers per iteration.
template <int G, bool TEAMS = false, int NX, int NU, int K>
__device__ int pcg(Regs<NX, NU, K>& R, const double (&gam)[K], double (&lam)[K], const Lds& L, int N, int lane,
                   double eps, int max_iter, double* eta_out, const Team& T, bool live = true,
                   int* flags = nullptr, int ipw = 1) {
  const int row = T.tid >> 4;  // this row's workgroup-sum partial
  double r[K], p[K], rt[K], ap[K], yxo[K], gb[K];
  const bool lead = lane == 0;
  // ---- r = gamma - S lam0 -------------------------------------------------
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int t = R.t[s];
    const double pv = own<NX>(L.lamS, t, lane);
    const double pn = own<NX>(L.lamS, t + 1, lane);
    const double wx = ((t == 0) ? pv : -pv) + gdot<G, NX>(R.acol[s], pn, lane, 0);
    double q;
    gb[s] = bgram<G>(R, s, L, pv, 0, &q);
    lam[s] = R.ok[s] ? pv : 0.0;
    p[s] = 0.0;
    ap[s] = 0.0;
    yxo[s] = R.qi[s] * wx;
    if (R.ok[s]) L.YX[(t + 1) * NX + lane] = yxo[s];
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int t = R.t[s];
    const double u = own<NX>(L.YX, t - 1, lane);
    const double g = (t == 0 ? 1.0 : -1.0) * yxo[s] + gdot<G, NX>(R.arow[s], u, lane, 0) + gb[s];
    r[s] = R.ok[s] ? gam[s] + g : 0.0;  // gamma - S lam0, S lam0 = -g
    const double y = dmul<G>(R, s, L, r[s], lane);
    if (R.ok[s]) {
      L.Y[(t + 1) * NX + lane] = y;
      L.YQ[(t + 1) * NX + lane] = R.qi[s] * y;
    }
  }
  __syncthreads();
  double part = 0.0;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int t = R.t[s];
    const double a1 = gdot<G, NX>(R.arow[s], own<NX>(L.YQ, t - 1, lane), lane, 0);
    const double a2 = gdot<G, NX>(R.acol[s], own<NX>(L.Y, t + 1, lane), lane, 0);
    const double z = R.ok[s] ? r[s] - (t == 1 ? 1.0 : -1.0) * a1 - (t == 0 ? 1.0 : -1.0) * R.qi[s] * a2 : 0.0;
    rt[s] = dmul<G>(R, s, L, z, lane);
    if (R.ok[s]) {
      part += r[s] * rt[s];
      L.RT[(t + 1) * NX + lane] = rt[s];
    }
  }
  // Each workgroup sum is split around its barrier: the loads of neighbour
  // entries written before the barrier, and the products of those entries
  // that do not need the sum (A_t^T rt_{t+1} here and at the end of phase C,
  // A_{t-1} u_{t-1} in phase B), are issued between the barrier and the
  // reading of the partials, so their LDS latency overlaps the sum's.
  double rn[K], apn[K];
  block_sum_rows_publish(part, L.redC, row);
#pragma unroll
  for (int s = 0; s < K; ++s) rn[s] = own<NX>(L.RT, R.t[s] + 1, lane);
  gdot_k<G, NX>(R.acol, rn, lane, 0, apn);
  double eta = block_sum_rows_finish(L.redC);
  double beta = 0.0;
  int it = 0;
  bool go = live;  // TEAMS: this team still iterates
  while (true) {
    // (this form, not an if-constexpr block around the break, keeps the
    // one-instance loop's code as it was: the block changed its scheduling)
    const bool cont = (fabs(eta) > eps && it <= max_iter - 1) || it < 1;
    if (!TEAMS && !cont) break;
    if constexpr (TEAMS) go = go && cont;
    // ---- phase A: p = rt + beta p; w = G^T p; y = H^-1 w; v = -w.y --------
    // A_t^T p_{t+1} = A_t^T rt_{t+1} + beta A_t^T p_{t+1}(previous), so only
    // the neighbour's rt is read (apn, formed before beta was known); the
    // previous product stays in registers.
    double vpart = 0.0;
#pragma unroll
    for (int s = 0; s < K; ++s) {
      const int t = R.t[s];
      const double pv = rt[s] + beta * p[s];
      p[s] = pv;
      ap[s] = apn[s] + beta * ap[s];
      double q;
      gb[s] = bgram<G>(R, s, L, pv, it, &q);
      const double wx = ((t == 0) ? pv : -pv) + ap[s];
      yxo[s] = R.qi[s] * wx;
      if (R.ok[s]) {
        L.YX[(t + 1) * NX + lane] = yxo[s];
        vpart += wx * yxo[s];
        // u_{t-1} term p^T B Rinv B^T p: once per block, or spread over
        // the lanes as p_i (B Rinv B^T p)_i
        if constexpr (quad_bterm<G, NU>()) {
          vpart += pv * gb[s];
        } else {
          if (lead) vpart += q;
        }
      }
    }
    if constexpr (TEAMS) {
      if (T.tid == 0) flags[T.idx] = go;
    }
    block_sum_rows_publish(vpart, L.redA, row);
    if constexpr (TEAMS) {
      int any = 0;
      for (int j = 0; j < ipw; ++j) any |= flags[j];
      if (!any) break;  // the same flags in every thread: a uniform exit
    }
    // ---- phase B: ups = S p; lam += alpha p; r -= alpha ups; y = D r -----
    {
      double u[K], au[K], y[K];
#pragma unroll
      for (int s = 0; s < K; ++s) u[s] = own<NX>(L.YX, R.t[s] - 1, lane);
      gdot_k<G, NX>(R.arow, u, lane, it, au);
      const double v = -block_sum_rows_finish(L.redA);
      const double alpha = div_newton(eta, v);
#pragma unroll
      for (int s = 0; s < K; ++s) {
        const double g = (R.t[s] == 0 ? 1.0 : -1.0) * yxo[s] + au[s] + gb[s];
        if (R.ok[s] && (!TEAMS || go)) {
          lam[s] += alpha * p[s];
          r[s] += alpha * g;  // r -= alpha * ups, ups = -g
        }
      }
      dmul_k<G>(R, L, r, lane, y);
#pragma unroll
      for (int s = 0; s < K; ++s) {
        if (R.ok[s]) {
          const int t = R.t[s];
          L.Y[(t + 1) * NX + lane] = y[s];
          L.YQ[(t + 1) * NX + lane] = R.qi[s] * y[s];
        }
      }
    }
    __syncthreads();
    // ---- phase C: rt = Phi^-1 r; eta' = r.rt ------------------------------
    double epart = 0.0;
    {
      double z[K];
#pragma unroll
      for (int s = 0; s < K; ++s) {
        const int t = R.t[s];
        double a1, a2;
        gdot_two<G, NX>(R.arow[s], own<NX>(L.YQ, t - 1, lane), R.acol[s], own<NX>(L.Y, t + 1, lane), lane, it, a1,
                        a2);
        z[s] = R.ok[s] ? r[s] - (t == 1 ? 1.0 : -1.0) * a1 - (t == 0 ? 1.0 : -1.0) * R.qi[s] * a2 : 0.0;
      }
      dmul_k<G>(R, L, z, lane, rt);
    }
#pragma unroll
    for (int s = 0; s < K; ++s) {
      if (R.ok[s]) {
        epart += r[s] * rt[s];
        L.RT[(R.t[s] + 1) * NX + lane] = rt[s];
      }
    }
    block_sum_rows_publish(epart, L.redB, row);
    // next iteration's A_t^T rt_{t+1} (unused if the loop ends here)
#pragma unroll
    for (int s = 0; s < K; ++s) rn[s] = own<NX>(L.RT, R.t[s] + 1, lane);
    gdot_k<G, NX>(R.acol, rn, lane, it, apn);
    const double eta_new = block_sum_rows_finish(L.redB);
    if (!TEAMS || go) {
      beta = div_newton(eta_new, eta);
      eta = eta_new;
      ++it;
    }
  }
#pragma unroll
  for (int s = 0; s < K; ++s)
    if (R.ok[s]) L.lamS[(R.t[s] + 1) * NX + lane] = lam[s];
  *eta_out = eta;
  return it;
}

// gamma = C + G H^-1 q (pcg.py:245-288) for the rows this thread owns.
// WITH_C: C = [x0; -f_d + A xbar + B ubar] (forward); otherwise C = 0.
template <bool WITH_C, int G, int NX, int NU, int K>
__device__ __forceinline__ void gamma_rows(const Regs<NX, NU, K>& R, const Lds& L, int lane, double (&gam)[K]) {
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int t = R.t[s];
    double g = 0.0;
    if (R.ok[s]) {
      if (t == 0) {
        g = R.qi[s] * L.qv[lane] + (WITH_C ? L.x0[lane] : 0.0);
      } else {
        const int tp = t - 1;
        g = -R.qi[s] * L.qv[t * NX + lane];
        if (WITH_C) g -= L.fd[tp * NX + lane];
#pragma unroll
        for (int k = 0; k < NX; ++k) {
          double v = L.qinv[tp * NX + k] * L.qv[tp * NX + k];
          if (WITH_C) v += L.xb[tp * NX + k];
          g += (double)R.arow[s][k] * v;
        }
        const int bq = quad_bterm<G, NU>() ? quad16(lane) : 0;  // brow slot k is column k ^ bq
#pragma unroll
        for (int k = 0; k < NU; ++k) {
          const int c = k ^ bq;
          double v = L.rinv[c] * L.rv[tp * NU + c];
          if (WITH_C) v += L.ub[tp * NU + c];
          g += (double)R.brow[s][k] * v;
        }
      }
    }
    gam[s] = g;
  }
}

// Primal recovery from the multipliers in L.lamS (pcg.py:316-348).
// Block t yields state row t and control row t - 1 (lanes < nu):
//   x_t = -Qinv_t (q_t + (t == 0 ? lam_t : -lam_t) + A_t^T lam_{t+1}),
//   u_{t-1} = -Rinv (r_{t-1} + B_{t-1}^T lam_t);
// block 0 yields the zero last control row instead (control_row()).
template <int G, int NX, int NU, int K>
__device__ __forceinline__ void recover_rows(const Regs<NX, NU, K>& R, const Lds& L, int N, int lane,
                                             double (&xn)[K], double (&un)[K]) {
#pragma unroll
  for (int s = 0; s < K; ++s) {
    const int t = R.t[s];
    const double lt = own<NX>(L.lamS, t, lane);
    const double ln = own<NX>(L.lamS, t + 1, lane);
    const double ax = gdot<G, NX>(R.acol[s], ln, lane, 0);
    double bl = 0.0;  // component `lane` of B_{t-1}^T lam_t
    if constexpr (quad_bterm<G, NU>()) {
      // quad k holds component k; lanes 1, 2, 3 reach quads 1, 2, 3 with
      // row_half_mirror, row_ror:8 and row_mirror
      const double sq = rs16(R.brow[s], lt, 0);
      const double s1 = dpp_mov<kDppHalfMirror>(sq), s2 = dpp_mov<kDppRor8>(sq), s3 = dpp_mov<kDppMirror>(sq);
      bl = lane == 0 ? sq : lane == 1 ? s1 : lane == 2 ? s2 : lane == 3 ? s3 : 0.0;
    } else {
#pragma unroll
      for (int k = 0; k < NU; ++k) {
        const double sk = gsum<G>((double)R.brow[s][k] * lt);
        bl = (lane == k) ? sk : bl;
      }
    }
    xn[s] = 0.0;
    un[s] = 0.0;
    if (R.ok[s]) {
      xn[s] = -R.qi[s] * (L.qv[t * NX + lane] + (t == 0 ? lt : -lt) + ax);
      if (lane < NU && t >= 1) un[s] = -L.rinv[lane] * (L.rv[(t - 1) * NU + lane] + bl);
    }
  }
}

// Control row produced by block t (see recover_rows).
__device__ __forceinline__ int control_row(int t, int N) { return t == 0 ? N : t - 1; }

template <int NX, int NU, int K>
__device__ __forceinline__ void init_blocks(Regs<NX, NU, K>& R, int G, int N, const Team& T, int* lane_out,
                                            int* gbase_out) {
  const int NG = T.nt / G;
  const int grp = T.tid / G;
  const int lane = T.tid % G;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    R.t[s] = s * NG + grp;
    R.ok[s] = (R.t[s] <= N) && (lane < NX);
    if (R.t[s] > N) R.t[s] = N;  // keep indices in range for masked lanes
  }
  *lane_out = lane;
  *gbase_out = (threadIdx.x & 63) - lane;
}

template <int NX, int NU>
__device__ void zero_lds_vectors(const Lds& L, int N, const Team& T) {
  const int NBP = N + 3;
  for (int k = T.tid; k < NBP * NX; k += T.nt) {
    L.RT[k] = 0.0; L.Y[k] = 0.0; L.YQ[k] = 0.0; L.YX[k] = 0.0; L.lamS[k] = 0.0;
  }
  // block_sum_rows reads all 64 partials of each array
  for (int k = T.tid; k < 64; k += T.nt) {
    L.redA[k] = 0.0; L.redB[k] = 0.0; L.redC[k] = 0.0;
  }
}

// Staging of one instance for the forward QP: iterate, multipliers, f_d,
// x0 and the closed-form quadratic model of the diagonal cost (sqp.py:160-197,
// optimal_control_problem.py:249-323): Qinv = 1 / max(2 w, 1e-9),
// q = -2 w ref (terminal: w_x + w_f and final_state), r = -2 w_u u_ref.
// One pass over the knots in which each thread issues all its global loads
// at once (addresses clamped into the arrays, loads unconditional), then
// stores to LDS.  N: this instance's horizon; Np: the padded one.
template <int NX, int NU>
__device__ void stage_forward(const SolveArgs& a, const Lds& L, int b, int N, int Np, const Team& T) {
  const int NB = N + 1, NBs = Np + 1;
  const int nxe = NB * NX, nue = NB * NU, nfd = N * NX;
  for (int k = T.tid; k < nxe; k += T.nt) {
    const int t = k / NX, i = k % NX;
    const int ku = k < nue ? k : nue - 1, kf = k < nfd ? k : nfd - 1;
    const int tu = ku / NU, iu = ku % NU;
    const size_t xo = (size_t)b * NBs * NX + k;
    const double xv = a.x[xo], lv = a.lam[xo];
    const double uv = a.u[(size_t)b * NBs * NU + ku];
    const double fv = a.fd[(size_t)b * Np * NX + kf];
    const double w = a.wx[b * NX + i], wfi = a.wf[b * NX + i], xfi = a.xf[b * NX + i], x0i = a.x0[b * NX + i];
    const double xr = a.xref[((size_t)b * NBs + t) * NX + i];
    const double wu = a.wu[b * NU + iu];
    const double ur = a.uref[((size_t)b * NBs + tu) * NU + iu];
    L.xb[k] = xv;
    L.lamS[NX + k] = lv;
    if (t < N) {
      L.qinv[k] = 1.0 / fmax(2.0 * w, kPsdMinEig);
      L.qv[k] = -2.0 * w * xr;
    } else {
      const double wt = w + wfi;
      L.qinv[k] = 1.0 / fmax(2.0 * wt, kPsdMinEig);
      L.qv[k] = -2.0 * wt * xfi;
    }
    if (k < nue) {
      L.ub[k] = uv;
      L.rv[k] = (tu < N && a.penalize) ? -2.0 * wu * ur : 0.0;
    }
    if (k < nfd) L.fd[k] = fv;
    if (k < NX) L.x0[k] = x0i;
    if (k < NU) L.rinv[k] = 1.0 / fmax(2.0 * wu, kPsdMinEig);  // iu == k here
  }
}

// ---------------------------------------------------------------------------
// SQP driver kernels.  The SQP loop of sqp.py:460-533 runs as a fixed launch
// sequence (linearize, QP) x num_scp_iteration_max followed by a finish
// kernel, with no host synchronisation, so it can be captured in a hipGraph.
// Each instance decides on the device whether it is still iterating:
// (num_iter < maxit && err > tol) || num_iter < 1; finished instances return
// at once and keep their last linearization, like the reference's vmapped
// while_loop keeps frozen instances.

__global__ void sqp_reset_kernel(int B, int32_t* num_iter, int32_t* head) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < B) num_iter[b] = 0;
  if (head && b == 0) *head = 0;
}

// Longest-first launch order for the next QP launch.  The hardware hands
// workgroups to CUs in index order as CUs free up, so when several instances
// take turns on a CU the launch ends with whichever long instances started
// last.  The work of instance c is estimated as (N_c + 1) blocks times the
// PCG iterations of its previous SQP iteration (times 1 in the first, where
// only ragged horizons tell instances apart); finished instances get -1
// (they return at once).  order lists the instances by decreasing estimate,
// ties by index (a stable sort, so the result is deterministic).  Each
// workgroup stages every key in LDS and ranks kOrderPerGroup instances, 16
// threads per instance summing the comparisons of a sixteenth of the keys
// each.
constexpr int kOrderThreads = 1024;
constexpr int kOrderPerGroup = kOrderThreads / 16;
constexpr int kOrderMaxBatch = 32768;  // keys staged in LDS: 128 KB

// Work estimate of instance c before QP launch it of a solve.
struct SolveOrderKeys {
  SolveArgs a;
  __device__ int operator()(int c) const {
    if (!sqp_active(a, c)) return -1;
    const int ic = a.num_iter[c];
    const long long n = horizon_of(a.horizons, a.N, c) + 1;
    const long long k = ic < 1 ? n : (long long)a.pcg_iters[(size_t)c * a.maxit + ic - 1] * n;
    return k > 0x7fffffffLL ? 0x7fffffff : (int)k;
  }
};

// Work estimate of instance c of an adjoint solve: its blocks (the PCG
// iterations are not known yet).
struct HorizonOrderKeys {
  const int32_t* horizons;
  int N;
  __device__ int operator()(int c) const { return horizon_of(horizons, N, c) + 1; }
};

// order[i * stride] = instance of workgroup i.
template <class Keys>
__global__ void __launch_bounds__(kOrderThreads) order_kernel(Keys keys, int B, int32_t* order, size_t stride) {
  extern __shared__ int keys_[];
  __shared__ int rank[kOrderPerGroup];
  for (int c = threadIdx.x; c < B; c += blockDim.x) keys_[c] = keys(c);
  if (threadIdx.x < kOrderPerGroup) rank[threadIdx.x] = 0;
  __syncthreads();
  const int b = blockIdx.x * kOrderPerGroup + (threadIdx.x >> 4);
  const int part = threadIdx.x & 15;
  if (b < B) {
    const int kb = keys_[b];
    const int chunk = (B + 15) / 16;
    const int lo = part * chunk, hi = min(B, lo + chunk);
    int r = 0;
    for (int c = lo; c < hi; ++c) {
      const int kc = keys_[c];
      r += (kc > kb) || (kc == kb && c < b);
    }
    atomicAdd(&rank[threadIdx.x >> 4], r);
  }
  __syncthreads();
  if (part == 0 && b < B) order[(size_t)rank[threadIdx.x >> 4] * stride] = b;
}

// CUs of the current device (1 << 30 when unknown).
int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 1 << 30;
  }
  return cus;
}

// Whether ordering pays: more instances than CUs (then instances share CUs
// in turn), and few enough for the keys to fit in LDS.
int order_grid(int B) {
  if (B <= device_cus() || B > kOrderMaxBatch) return 0;
  return (B + kOrderPerGroup - 1) / kOrderPerGroup;
}

template <class Keys>
bool enqueue_order(const Keys& keys, int B, int32_t* order, size_t stride, hipStream_t st) {
  const int grid = order_grid(B);
  if (grid == 0 || !order) return false;
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&order_kernel<Keys>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, kOrderMaxBatch * (int)sizeof(int)) != hipSuccess)
      return false;
    attr = true;
  }
  hipLaunchKernelGGL(order_kernel<Keys>, dim3(grid), dim3(kOrderThreads), (size_t)B * sizeof(int), st, keys, B, order,
                     stride);
  return true;
}

// Enqueues the order kernel for QP launch `it` into `order`, or returns
// false when the identity order is kept (also for the first launch of a
// batch with one horizon, whose keys are all equal).
bool launch_order(const SolveArgs& a, int it, int32_t* order, hipStream_t st) {
  if (it < 1 && !a.horizons) return false;
  return enqueue_order(SolveOrderKeys{a}, a.B, order, 1, st);
}

// Column j of [A_t B_t], one thread per (b, t, j), then f_d(x_t, u_t), one
// thread per (b, t) after them (optimal_control_problem.py:360-383 with
// value_and_jacfwd).  The Jacobian is taken in float32 dual numbers; the
// value f_d, which fixes the constraint residual C_{t+1} = -f_d + A x_t + B u_t
// and so the point the SQP iteration converges to, is evaluated in float64 at
// the float64 iterate (a float32 value leaves the iteration hovering ~1e-8
// away from it).  The value threads fill waves of their own.
template <class Model>
__host__ __device__ constexpr long long linearize_threads(int B, int N) {
  return (long long)B * N * (Model::NX + Model::NU + 1);
}

template <class Model>
__global__ void __launch_bounds__(256) linearize_kernel(SolveArgs a) {
  constexpr int NX = Model::NX, NU = Model::NU, NZ = NX + NU;
  const int N = a.N;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long njac = (long long)a.B * N * NZ;
  if (idx >= njac) {
    const long long v = idx - njac;
    if (v >= (long long)a.B * N) return;
    const int t = v % N;
    const int b = v / N;
    if (!sqp_active(a, b) || t >= horizon_of(a.horizons, N, b)) return;
    float prm[Model::NPARAM];
#pragma unroll
    for (int k = 0; k < Model::NPARAM; ++k) prm[k] = a.dyn[(size_t)b * a.P + k];
    fd_value<Model>(a.x, a.u, prm, a.dt64, a.scheme, N, b, t, a.fd + ((size_t)b * N + t) * NX);
    return;
  }
  const int j = idx % NZ;
  const int t = (idx / NZ) % N;
  const int b = idx / ((long long)NZ * N);
  if (!sqp_active(a, b) || t >= horizon_of(a.horizons, N, b)) return;
  float xs[NX], us[NU], prm[Model::NPARAM];
#pragma unroll
  for (int i = 0; i < NX; ++i) xs[i] = (float)a.x[((size_t)b * (N + 1) + t) * NX + i];
#pragma unroll
  for (int k = 0; k < NU; ++k) us[k] = (float)a.u[((size_t)b * (N + 1) + t) * NU + k];
#pragma unroll
  for (int k = 0; k < Model::NPARAM; ++k) prm[k] = a.dyn[(size_t)b * a.P + k];
  float fv[NX], col[NX];
  step_jacobian_column<Model>(xs, us, prm, a.dt, a.scheme, j, fv, col);
  float* At = a.AB + ((size_t)b * N + t) * NX * NZ;
#pragma unroll
  for (int i = 0; i < NX; ++i) At[i * NZ + j] = col[i];
}

// One SQP iteration of one instance per workgroup: Schur/preconditioner
// assembly, gamma, the whole PCG loop, primal recovery and the step test
// (sqp.py:262-299, 506-528).
// LS: the build with the float64 merit line search (a.ls_eval).  Its model
// evaluations hold enough registers to make the whole kernel spill at the
// 128-VGPR cap of the one-block-per-group variant, so solves that do not
// evaluate the line search run a build without it, which spills nothing.
template <class Model, int G, int K, bool LS, bool TEAMS, bool PQ>
__device__ __forceinline__ void qp_iteration(const SolveArgs& a, int b, const Team& T, bool live, int* flags,
                                             char* lds) {
  constexpr int NX = Model::NX, NU = Model::NU, NZ = NX + NU;
  const int it = a.num_iter[b];
  const int Np = a.N, NBs = Np + 1;                          // padded horizon, knot stride
  const int N = horizon_of(a.horizons, Np, b), NB = N + 1;  // this instance
  const Lds L = carve<NX, NU>(lds, Np, linesearch_doubles(a.ls_eval, a.n_alphas, Np));
  Regs<NX, NU, K> R;
  int lane, gbase;
  init_blocks(R, G, N, T, &lane, &gbase);

  zero_lds_vectors<NX, NU>(L, Np, T);
  __syncthreads();
  stage_forward<NX, NU>(a, L, b, N, Np, T);
  __syncthreads();

  const float* ABb = a.AB + (size_t)b * Np * NX * NZ;
  assemble_dinv<G>(R, ABb, L, lane, gbase);
  load_ab<G>(R, ABb, L, N, lane);
  double gam[K], lam[K];
  gamma_rows<true, G>(R, L, lane, gam);
  double eta;
  const int npcg = pcg<G, TEAMS>(R, gam, lam, L, N, lane, a.pcg_eps, a.pcg_max, &eta, T, live, flags, a.ipw);
  __syncthreads();  // lam (in L.lamS) complete for the neighbour reads
  double xn[K], un[K];
  recover_rows<G>(R, L, N, lane, xn, un);
  double alpha = 1.0;
  if constexpr (LS) {
    double* DX = L.YX;  // (N+1, nx): QP step of the states
    double* DU = L.Y;   // (N+1, nu): QP step of the controls
#pragma unroll
    for (int s = 0; s < K; ++s) {
      if (R.ok[s]) {
        const int t = R.t[s], tu = control_row(t, N);
        DX[t * NX + lane] = xn[s] - L.xb[t * NX + lane];
        if (lane < NU) DU[tu * NU + lane] = un[s] - L.ub[tu * NU + lane];
      }
    }
    __syncthreads();
    alpha = line_search<Model, TEAMS>(a, b, N, L.xb, L.ub, DX, DU, L.LS, L.redA, T.tid, T.nt);
#pragma unroll
    for (int s = 0; s < K; ++s) {
      if (R.ok[s]) {
        const int t = R.t[s], tu = control_row(t, N);
        xn[s] = L.xb[t * NX + lane] + alpha * DX[t * NX + lane];
        if (lane < NU) un[s] = L.ub[tu * NU + lane] + alpha * DU[tu * NU + lane];
      }
    }
  }
  double e = 0.0;
#pragma unroll
  for (int s = 0; s < K; ++s) {
    if (R.ok[s]) {
      const int t = R.t[s];
      e = nanmax(e, fabs(xn[s] - L.xb[t * NX + lane]));
      if (lane < NU) e = nanmax(e, fabs(un[s] - L.ub[control_row(t, N) * NU + lane]));
    }
  }
  double err;
  if constexpr (TEAMS) {
    err = block_max_rows(e, L.redA, T.tid >> 4, T.nt >> 4);
  } else {
    err = block_max(e, L.redA);
  }
  if (!live) return;  // no barrier follows
#pragma unroll
  for (int s = 0; s < K; ++s) {
    if (R.ok[s]) {
      const int t = R.t[s];
      a.x[((size_t)b * NBs + t) * NX + lane] = xn[s];
      a.lam[((size_t)b * NBs + t) * NX + lane] = lam[s];
      if (lane < NU) a.u[((size_t)b * NBs + control_row(t, N)) * NU + lane] = un[s];
    }
  }
  // knots past this instance's horizon (ragged batches) are zero
  for (int k = NB * NX + T.tid; k < NBs * NX; k += T.nt) {
    a.x[(size_t)b * NBs * NX + k] = 0.0;
    a.lam[(size_t)b * NBs * NX + k] = 0.0;
  }
  for (int k = NB * NU + T.tid; k < NBs * NU; k += T.nt) a.u[(size_t)b * NBs * NU + k] = 0.0;
  for (int k = T.tid; k < NB * NX; k += T.nt) a.qinv[(size_t)b * NBs * NX + k] = L.qinv[k];
  for (int k = T.tid; k < NU; k += T.nt) a.rinv[(size_t)b * NU + k] = L.rinv[k];
  if (T.tid == 0) {
    a.conv_errs[(size_t)b * a.maxit + it] = err;
    a.pcg_iters[(size_t)b * a.maxit + it] = npcg;
    if (a.linesearch) a.ls_alphas[(size_t)b * a.maxit + it] = alpha;
    a.conv_err[b] = err;
    if constexpr (!PQ) a.num_iter[b] = it + 1;
  }
  if constexpr (PQ) {
    // publish the iteration to the workgroup that takes the instance's next
    // one (maybe on another XCD): every wave's stores complete at L2, then
    // one release at agent scope writes them back ahead of the count
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (T.tid == 0) __hip_atomic_store(a.num_iter + b, it + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// One launch for a whole solve instead of one linearization and one QP
// launch per SQP iteration: B * maxit workgroups, each of which takes the
// next position of a schedule from a counter (the solution's status[0], free
// until sqp_finish_kernel).  Position p is SQP iteration p / B of instance
// p % B; its workgroup waits until the instance has finished the previous
// iteration, linearizes it and runs the iteration.  So an instance's next
// iteration starts as soon as its own previous one is done, instead of after
// the whole batch's, and the tails of the per-iteration launches disappear.
// No deadlock: a position's predecessor was taken earlier, by a resident
// workgroup that is running it or has finished it (a waiting workgroup holds
// no unfinished position), so the lowest waiting position always has a
// running predecessor.  An instance that stopped (converged, or not finite)
// skips its remaining positions.  The wait is bounded: past kPqWaitTicks of
// the 100 MHz wall clock the workgroup ends the schedule (the counter is
// pushed past the end, so the remaining workgroups exit at once); the
// instances left unfinished then report their convergence error.  Results
// are the per-launch schedule's bits (the same per-instance computations).
// Returns the instance to solve, or -1.
constexpr long long kPqWaitTicks = 1LL << 29;  // ~5 s

__device__ int pq_take(const SolveArgs& a, int32_t* head) {
  const int total = a.B * a.maxit;
  const int p = atomicAdd(head, 1);
  if (p < 0 || p >= total) return -1;
  const int s = p / a.B, c = p - s * a.B;
  const long long t0 = wall_clock64();
  for (;;) {
    // iterations done so far, published by the release store that ends each
    const int n = __hip_atomic_load(a.num_iter + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (n >= s) break;
    if (n >= 1) {
      // stopped after n iterations: converged, or its error is not finite
      const long long e = __hip_atomic_load(reinterpret_cast<const long long*>(a.conv_err + c), __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
      if (!(__longlong_as_double(e) > a.tol)) return -1;
    }
    if (wall_clock64() - t0 > kPqWaitTicks) {
      atomicMax(head, total);
      return -1;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return sqp_active(a, c) ? c : -1;
}

// The linearization of instance b (linearize_kernel's work for one
// instance, the same per-item computations), by the threads of a workgroup:
// Jacobian columns first, then the values, so that few waves mix the two.
template <class Model>
__device__ void linearize_instance(const SolveArgs& a, int b) {
  constexpr int NX = Model::NX, NU = Model::NU, NZ = NX + NU;
  const int N = a.N, Nb = horizon_of(a.horizons, N, b);
  const int njac = Nb * NZ;
  for (int k = threadIdx.x; k < njac + Nb; k += blockDim.x) {
    float prm[Model::NPARAM];
#pragma unroll
    for (int i = 0; i < Model::NPARAM; ++i) prm[i] = a.dyn[(size_t)b * a.P + i];
    if (k >= njac) {
      const int t = k - njac;
      fd_value<Model>(a.x, a.u, prm, a.dt64, a.scheme, N, b, t, a.fd + ((size_t)b * N + t) * NX);
      continue;
    }
    const int j = k % NZ, t = k / NZ;
    float xs[NX], us[NU];
#pragma unroll
    for (int i = 0; i < NX; ++i) xs[i] = (float)a.x[((size_t)b * (N + 1) + t) * NX + i];
#pragma unroll
    for (int i = 0; i < NU; ++i) us[i] = (float)a.u[((size_t)b * (N + 1) + t) * NU + i];
    float fv[NX], col[NX];
    step_jacobian_column<Model>(xs, us, prm, a.dt, a.scheme, j, fv, col);
    float* At = a.AB + ((size_t)b * N + t) * NX * NZ;
#pragma unroll
    for (int i = 0; i < NX; ++i) At[i * NZ + j] = col[i];
  }
}

template <class Model, int G, int K, int MAXT, bool LS, bool TEAMS, bool PQ = false>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(team_waves_per_simd<TEAMS>())))
qp_iteration_kernel(SolveArgs a) {
  extern __shared__ double smem_[];
  const Team T = team_of<TEAMS>(a.ipw, a.team_threads);
  char* lds = reinterpret_cast<char*>(smem_);
  if constexpr (PQ) {
    static_assert(!TEAMS, "the schedule runs one instance per workgroup");
    __shared__ int task;
    if (threadIdx.x == 0) task = pq_take(a, a.status);
    __syncthreads();
    const int b = __builtin_amdgcn_readfirstlane(task);  // uniform, as blockIdx.x is
    if (b < 0) return;
    linearize_instance<Model>(a, b);
    __syncthreads();
    qp_iteration<Model, G, K, LS, false, true>(a, b, T, true, nullptr, lds);
  } else {
    int b;
    bool live = true;  // this team's instance is solved (else its team only keeps the barriers)
    int* flags = nullptr;
    if constexpr (TEAMS) {
      // a workgroup whose instances have all stopped returns at once
      bool any = false;
      for (int j = 0; j < a.ipw; ++j) {
        const int sl = (int)blockIdx.x * a.ipw + j;
        if (sl < a.B) any = any || sqp_active(a, a.order ? a.order[sl] : sl);
      }
      if (!any) return;
      const int sl = T.slot < a.B ? T.slot : a.B - 1;  // slots past the batch read the last one
      b = a.order ? a.order[sl] : sl;
      live = T.slot < a.B && sqp_active(a, b);
      flags = reinterpret_cast<int*>(lds + (size_t)a.ipw * a.team_lds);
      lds += (size_t)T.idx * a.team_lds;
    } else {
      b = a.order ? a.order[blockIdx.x] : blockIdx.x;
      if (!sqp_active(a, b)) return;
    }
    qp_iteration<Model, G, K, LS, TEAMS, false>(a, b, T, live, flags, lds);
  }
}

// status and nan_to_num (sqp.py:535-559)
__global__ void __launch_bounds__(256) sqp_finish_kernel(SolveArgs a, int nx, int nu) {
  __shared__ double red[16];
  const int b = blockIdx.x;
  const int NB = a.N + 1;
  double bad = 0.0;
  for (int k = threadIdx.x; k < NB * nx; k += blockDim.x) {
    const double v = a.x[(size_t)b * NB * nx + k];
    if (v != v) bad = 1.0;
  }
  for (int k = threadIdx.x; k < NB * nu; k += blockDim.x) {
    const double v = a.u[(size_t)b * NB * nu + k];
    if (v != v) bad = 1.0;
  }
  bad = block_max(bad, red);
  for (int k = threadIdx.x; k < NB * nx; k += blockDim.x) {
    a.x[(size_t)b * NB * nx + k] = nan_to_num(a.x[(size_t)b * NB * nx + k]);
    a.lam[(size_t)b * NB * nx + k] = nan_to_num(a.lam[(size_t)b * NB * nx + k]);
  }
  for (int k = threadIdx.x; k < NB * nu; k += blockDim.x)
    a.u[(size_t)b * NB * nu + k] = nan_to_num(a.u[(size_t)b * NB * nu + k]);
  if (threadIdx.x == 0) a.status[b] = (bad > 0.0 || a.conv_err[b] > a.tol) ? -100 : 0;
}

// ---------------------------------------------------------------------------
// Adjoint backward (sqp.py:668-740): PCG on the saved linearization with
// q~ = -dl/dx, r~ = -dl/du[:-1], C = 0, recovery of (x~, u~), then the
// closed-form mixed-derivative contraction of the diagonal tracking cost.
template <class Model, int G, int K, int MAXT, bool TEAMS>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(team_waves_per_simd<TEAMS>())))
fused_backward_kernel(BackwardArgs a) {
  constexpr int NX = Model::NX, NU = Model::NU, NZ = NX + NU;
  extern __shared__ double smem_[];
  const Team T = team_of<TEAMS>(a.ipw, a.team_threads);
  int b = blockIdx.x;
  bool live = true;  // as in qp_iteration_kernel
  int* flags = nullptr;
  char* lds = reinterpret_cast<char*>(smem_);
  if constexpr (TEAMS) {
    b = T.slot < a.B ? T.slot : a.B - 1;  // slots past the batch read the last one
    live = T.slot < a.B;
    flags = reinterpret_cast<int*>(lds + (size_t)a.ipw * a.team_lds);
    lds += (size_t)T.idx * a.team_lds;
  }
  const int Np = a.N, NBs = Np + 1;                          // padded horizon, knot stride
  const int N = horizon_of(a.horizons, Np, b), NB = N + 1;  // this instance
  const Lds L = carve<NX, NU>(lds, Np);
  Regs<NX, NU, K> R;
  int lane, gbase;
  init_blocks(R, G, N, T, &lane, &gbase);

  zero_lds_vectors<NX, NU>(L, Np, T);
  __syncthreads();
  // one pass, all global loads of a thread issued at once (clamped
  // addresses), as stage_forward
  for (int k = T.tid; k < NB * NX; k += T.nt) {
    const int ku = k < NB * NU ? k : NB * NU - 1;
    const size_t xo = (size_t)b * NBs * NX + k, uo = (size_t)b * NBs * NU + ku;
    const double qi = a.qinv[xo], dx = a.dl_dx[xo], lb = a.lam_bwd[xo], xv = a.x[xo];
    const double du = a.dl_du[uo], uv = a.u[uo];
    const double ri = a.rinv[(size_t)b * NU + (ku < NU ? ku : 0)];
    L.qinv[k] = qi;
    L.qv[k] = -dx;
    L.lamS[NX + k] = lb;
    L.xb[k] = xv;
    if (k < NB * NU) {
      L.rv[k] = ku / NU < N ? -du : 0.0;
      L.ub[k] = uv;
    }
    if (k < NU) L.rinv[k] = ri;
  }
  __syncthreads();
  const float* ABb = a.AB + (size_t)b * Np * NX * NZ;
  assemble_dinv<G>(R, ABb, L, lane, gbase);
  load_ab<G>(R, ABb, L, N, lane);
  double gam[K], lam[K];
  gamma_rows<false, G>(R, L, lane, gam);
  double eta;
  const int npcg = pcg<G, TEAMS>(R, gam, lam, L, N, lane, a.pcg_eps, a.pcg_max, &eta, T, live, flags, a.ipw);
  __syncthreads();  // lam (in L.lamS) complete for the neighbour reads
  double xt[K], ut[K];
  recover_rows<G>(R, L, N, lane, xt, ut);
  // per-row contraction terms -> LDS (reusing the PCG vectors), then one
  // thread per component sums over time in a fixed order
  double* TX = L.YX;  // (NB, NX): 2 (x - ref) x~
  double* TU = L.Y;   // (NB, NU)
  double* XT = L.RT;  // (NB, NX): x~, staged for the optional adjoint output
  double* UT = L.YQ;  // (NB, NU): u~
#pragma unroll
  for (int s = 0; s < K; ++s) {
    if (R.ok[s] && live) {
      const int t = R.t[s];
      const int i = lane;
      const double ref = (t < N) ? a.xref[((size_t)b * NBs + t) * NX + i] : a.xf[b * NX + i];
      const double ex = L.xb[t * NX + i] - ref;
      TX[t * NX + i] = 2.0 * ex * xt[s];
      const double wxi = a.wx[b * NX + i];
      a.g_xref[((size_t)b * NBs + t) * NX + i] = (t < N) ? 2.0 * wxi * xt[s] : 0.0;
      if (t == N) {
        a.g_wf[b * NX + i] = -2.0 * ex * xt[s];
        a.g_xf[b * NX + i] = 2.0 * (wxi + a.wf[b * NX + i]) * xt[s];
      }
      if (i < NU) {
        const int tu = control_row(t, N);
        const double ur = a.penalize ? a.uref[((size_t)b * NBs + tu) * NU + i] : 0.0;
        TU[tu * NU + i] = (tu < N) ? 2.0 * (L.ub[tu * NU + i] - ur) * ut[s] : 0.0;
        a.g_uref[((size_t)b * NBs + tu) * NU + i] =
            (tu < N && a.penalize) ? 2.0 * a.wu[b * NU + i] * ut[s] : 0.0;
      }
      a.lam_bwd[((size_t)b * NBs + t) * NX + i] = lam[s];
      if (t == 0) a.g_x0[b * NX + i] = lam[s];
      XT[t * NX + i] = xt[s];
      if (i < NU) {
        const int tu = control_row(t, N);
        UT[tu * NU + i] = (tu < N) ? ut[s] : 0.0;
      }
    }
  }
  // knots past this instance's horizon (ragged batches) are zero
  for (int k = NB * NX + T.tid; live && k < NBs * NX; k += T.nt) {
    a.g_xref[(size_t)b * NBs * NX + k] = 0.0;
    a.lam_bwd[(size_t)b * NBs * NX + k] = 0.0;
  }
  for (int k = NB * NU + T.tid; live && k < NBs * NU; k += T.nt) a.g_uref[(size_t)b * NBs * NU + k] = 0.0;
  __syncthreads();
  if (!live) return;  // no barrier follows
  for (int i = T.tid; i < NX + NU; i += T.nt) {
    double acc = 0.0;
    if (i < NX) {
      for (int t = 0; t <= N; ++t) acc += TX[t * NX + i];
      a.g_wx[b * NX + i] = -acc;
    } else {
      const int k = i - NX;
      for (int t = 0; t < N; ++t) acc += TU[t * NU + k];
      a.g_wu[b * NU + k] = -acc;
    }
  }
  if (a.adj_x) {
    for (int k = T.tid; k < NBs * NX; k += T.nt) a.adj_x[(size_t)b * NBs * NX + k] = k < NB * NX ? XT[k] : 0.0;
  }
  if (a.adj_u) {
    for (int k = T.tid; k < NBs * NU; k += T.nt) a.adj_u[(size_t)b * NBs * NU + k] = k < N * NU ? UT[k] : 0.0;
  }
  if (T.tid == 0) a.pcg_iters[b] = npcg;
}

constexpr size_t kMaxLds = 160 * 1024;
constexpr int kMaxTeams = 8;  // instances per workgroup at most

struct Launch {
  int K, threads;
  size_t lds;
  int ipw = 1;        // instances per workgroup (teams)
  int team_threads = 0;
  size_t team_lds = 0;
};

// K blocks per thread group -> at most 1024 / 512 / 256 threads, so the
// compiler may give each thread 128 / 256 / 512 registers.
inline int max_threads(int K) { return K == 1 ? 1024 : (K == 2 ? 512 : 256); }

// Largest number of time blocks per lane group a model is dispatched with.
// The attitude model stops at two: its four-block backward
// (fused_backward_kernel<Spacecraft7Model, 8, 4, 256>, 256 VGPRs + AGPRs,
// SGPRs spilled to VGPR lanes) once returned run-to-run varying dl/dw_x
// when its adjoint outputs were stored to global memory inside the block
// loop, and the cause was not found in its code or its assembly; the
// variant is not built rather than kept on a workaround.
template <class Model>
constexpr int max_blocks_per_group() {
  return std::is_same<Model, Spacecraft7Model>::value ? 2 : 4;
}

template <class Model, int G>
Launch plan(int N, int P, int ls = 0, int forced_K = 0) {
  constexpr int NX = Model::NX, NU = Model::NU;
  const int NB = N + 1;
  Launch l{0, 0, LdsPlan<NX, NU>::bytes(N, P, ls)};
  // one block per thread group first (the most waves per instance; with the
  // fused DPP-FMA chains it measured at least as fast as two for every
  // model), then two, then four
  const int cand[3] = {1, 2, 4};
  for (int c = 0; c < 3; ++c) {
    const int K = cand[c];
    if (K > max_blocks_per_group<Model>()) continue;
    const int want = forced_K > 0 ? forced_K : g_forced_K;
    if (want > 0 && K != want) continue;
    const int groups = (NB + K - 1) / K;
    const int threads = ((groups * G + 63) / 64) * 64;
    if (threads <= max_threads(K)) {
      l.K = K;
      l.threads = threads;
      return l;
    }
  }
  return l;  // K = 0: does not fit
}

// Several instances per workgroup (ipw > 1): teams of one block per lane
// group, each of tt = (N + 1) G threads rounded up to a DPP row of 16, with
// its own LDS region (the one-instance plan's bytes rounded to 16) and one
// int flag per team after the regions.  Only for 4- and 8-lane groups,
// whose short-horizon instances leave most of a wavefront's lanes idle when
// alone; K = 0 when the request does not fit one workgroup.
template <class Model, int G>
Launch plan_teams(int N, int P, int ls, int forced_K, int ipw) {
  Launch l{0, 0, 0};
  const int want = forced_K > 0 ? forced_K : g_forced_K;
  if (G > 8 || ipw < 2 || ipw > kMaxTeams || (want > 0 && want != 1)) return l;
  const int tt = ((N + 1) * G + 15) / 16 * 16;
  const size_t team_lds = (LdsPlan<Model::NX, Model::NU>::bytes(N, P, ls) + 15) / 16 * 16;
  if (ipw * tt > max_threads(1)) return l;
  l.K = 1;
  l.threads = ipw * tt;
  l.lds = (size_t)ipw * team_lds + kMaxTeams * sizeof(int);
  l.ipw = ipw;
  l.team_threads = tt;
  l.team_lds = team_lds;
  return l;
}

// The whole solve as one launch of B * maxit workgroups (see pq_take), the
// schedule's counter in status[0].  DMPC_ERR_UNSUPPORTED when the
// per-iteration launches are to be used instead (the kernel's static LDS
// does not fit next to the instance's, or the schedule would not fit an int).
template <class Model, int G, int K, int MAXT, bool LS>
int launch_fwd_persistent(const SolveArgs& a, const Launch& l, hipStream_t st) {
  const void* kern = reinterpret_cast<const void*>(&qp_iteration_kernel<Model, G, K, MAXT, LS, false, true>);
  constexpr size_t kStatic = 256;  // the kernel's own LDS (the task slot), rounded up
  if (l.lds + kStatic > kMaxLds || (long long)a.B * a.maxit >= (1LL << 30)) return DMPC_ERR_UNSUPPORTED;
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kMaxLds - kStatic)) != hipSuccess)
      return DMPC_ERR_HIP;
    attr = true;
  }
  const unsigned groups = (unsigned)(a.B * a.maxit);
  SolveArgs ai = a;
  ai.order = nullptr;
  ai.ipw = 1;
  ai.team_threads = l.team_threads;
  ai.team_lds = (int)l.team_lds;
  hipLaunchKernelGGL(sqp_reset_kernel, dim3((a.B + 255) / 256), dim3(256), 0, st, a.B, a.num_iter, a.status);
  prof_begin(PROF_QP_ITERATION, st);
  hipLaunchKernelGGL((qp_iteration_kernel<Model, G, K, MAXT, LS, false, true>), dim3(groups), dim3(l.threads), l.lds,
                     st, ai);
  prof_end(PROF_QP_ITERATION, st);
  hipLaunchKernelGGL(sqp_finish_kernel, dim3(a.B), dim3(256), 0, st, a, Model::NX, Model::NU);
  return hipGetLastError() == hipSuccess ? DMPC_OK : DMPC_ERR_HIP;
}

template <class Model, int G, int K, int MAXT, bool LS, bool TEAMS = false>
int launch_fwd_k(const SolveArgs& a, const Launch& l, hipStream_t st) {
  if constexpr (!TEAMS) {
    // batches of one horizon with more instances than CUs: one launch per
    // solve.  With fewer, every instance runs at once in each per-iteration
    // launch, so there is no launch tail to recover, and the separate
    // linearization kernel spreads that work over the whole device
    // (measured: the 64-instance rollout and the 32-instance imitation step
    // 3-6 % slower in one launch)
    if (!a.horizons && a.schedule == 0 && a.B > device_cus()) {
      const int rc = launch_fwd_persistent<Model, G, K, MAXT, LS>(a, l, st);
      if (rc != DMPC_ERR_UNSUPPORTED) return rc;
    }
  }
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&qp_iteration_kernel<Model, G, K, MAXT, LS, TEAMS>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds) != hipSuccess)
      return DMPC_ERR_HIP;
    attr = true;
  }
  const unsigned groups = (unsigned)((a.B + l.ipw - 1) / l.ipw);
  hipLaunchKernelGGL(sqp_reset_kernel, dim3((a.B + 255) / 256), dim3(256), 0, st, a.B, a.num_iter, nullptr);
  const long long nlin = linearize_threads<Model>(a.B, a.N);
  for (int it = 0; it < a.maxit; ++it) {
    prof_begin(PROF_LINEARIZE, st);
    hipLaunchKernelGGL(linearize_kernel<Model>, dim3((unsigned)((nlin + 255) / 256)), dim3(256), 0, st, a);
    prof_end(PROF_LINEARIZE, st);
    // the solution's status array is free until sqp_finish_kernel writes it
    SolveArgs ai = a;
    ai.order = launch_order(a, it, a.status, st) ? a.status : nullptr;
    ai.ipw = l.ipw;
    ai.team_threads = l.team_threads;
    ai.team_lds = (int)l.team_lds;
    prof_begin(PROF_QP_ITERATION, st);
    hipLaunchKernelGGL((qp_iteration_kernel<Model, G, K, MAXT, LS, TEAMS>), dim3(groups), dim3(l.threads), l.lds,
                       st, ai);
    prof_end(PROF_QP_ITERATION, st);
  }
  hipLaunchKernelGGL(sqp_finish_kernel, dim3(a.B), dim3(256), 0, st, a, Model::NX, Model::NU);
  return hipGetLastError() == hipSuccess ? DMPC_OK : DMPC_ERR_HIP;
}

template <class Model, int G, int K, int MAXT, bool TEAMS = false>
int launch_bwd_k(const BackwardArgs& a, const Launch& l, hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&fused_backward_kernel<Model, G, K, MAXT, TEAMS>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds) != hipSuccess)
      return DMPC_ERR_HIP;
    attr = true;
  }
  BackwardArgs ai = a;
  ai.ipw = l.ipw;
  ai.team_threads = l.team_threads;
  ai.team_lds = (int)l.team_lds;
  const unsigned groups = (unsigned)((a.B + l.ipw - 1) / l.ipw);
  prof_begin(PROF_BACKWARD, st);
  hipLaunchKernelGGL((fused_backward_kernel<Model, G, K, MAXT, TEAMS>), dim3(groups), dim3(l.threads), l.lds, st,
                     ai);
  prof_end(PROF_BACKWARD, st);
  return hipGetLastError() == hipSuccess ? DMPC_OK : DMPC_ERR_HIP;
}

template <class Model, int G>
int fwd_model(const SolveArgs& a, hipStream_t st) {
  const int lsd = linesearch_doubles(a.ls_eval, a.n_alphas, a.N);
  if (a.ipw > 1) {
    if constexpr (G <= 8) {
      const Launch l = plan_teams<Model, G>(a.N, a.P, lsd, a.forced_K, a.ipw);
      if (l.K != 1 || l.lds > kMaxLds) return DMPC_ERR_UNSUPPORTED;
      return a.ls_eval ? launch_fwd_k<Model, G, 1, 1024, true, true>(a, l, st)
                       : launch_fwd_k<Model, G, 1, 1024, false, true>(a, l, st);
    }
    return DMPC_ERR_UNSUPPORTED;
  }
  Launch l = plan<Model, G>(a.N, a.P, lsd, a.forced_K);
  if (l.lds > kMaxLds) return DMPC_ERR_UNSUPPORTED;
  switch (l.K) {
    case 1:
      return a.ls_eval ? launch_fwd_k<Model, G, 1, 1024, true>(a, l, st) : launch_fwd_k<Model, G, 1, 1024, false>(a, l, st);
    case 2:
      return a.ls_eval ? launch_fwd_k<Model, G, 2, 512, true>(a, l, st) : launch_fwd_k<Model, G, 2, 512, false>(a, l, st);
    case 4:
      if constexpr (max_blocks_per_group<Model>() >= 4)
        return a.ls_eval ? launch_fwd_k<Model, G, 4, 256, true>(a, l, st) : launch_fwd_k<Model, G, 4, 256, false>(a, l, st);
      return DMPC_ERR_UNSUPPORTED;
    default: return DMPC_ERR_UNSUPPORTED;
  }
}

template <class Model, int G>
int bwd_model(const BackwardArgs& a, hipStream_t st) {
  if (a.ipw > 1) {
    if constexpr (G <= 8) {
      const Launch l = plan_teams<Model, G>(a.N, 0, 0, a.forced_K, a.ipw);
      if (l.K != 1 || l.lds > kMaxLds) return DMPC_ERR_UNSUPPORTED;
      return launch_bwd_k<Model, G, 1, 1024, true>(a, l, st);
    }
    return DMPC_ERR_UNSUPPORTED;
  }
  Launch l = plan<Model, G>(a.N, 0, 0, a.forced_K);
  if (l.lds > kMaxLds) return DMPC_ERR_UNSUPPORTED;
  switch (l.K) {
    case 1: return launch_bwd_k<Model, G, 1, 1024>(a, l, st);
    case 2: return launch_bwd_k<Model, G, 2, 512>(a, l, st);
    case 4:
      if constexpr (max_blocks_per_group<Model>() >= 4) return launch_bwd_k<Model, G, 4, 256>(a, l, st);
      return DMPC_ERR_UNSUPPORTED;
    default: return DMPC_ERR_UNSUPPORTED;
  }
}

template <class Model, int G>
bool fits(int N, int lsd) {
  Launch l = plan<Model, G>(N, Model::NPARAM, lsd, 0);
  return l.K > 0 && l.lds <= kMaxLds;
}

bool register_fits(int model, int nx, int nu, int N, int lsd) {
  if (N < 1) return false;
  switch (model) {
    case MODEL_LINEAR:
      if (nx == 8 && nu == 4) return fits<LinearModel<8, 4>, 8>(N, lsd);
      return nx == 4 && nu == 2 && fits<LinearModel<4, 2>, 4>(N, lsd);
    case MODEL_QUADROTOR: return nx == 13 && nu == 4 && fits<QuadrotorModel, 16>(N, lsd);
    case MODEL_SPACECRAFT3: return nx == 3 && nu == 3 && fits<Spacecraft3Model, 4>(N, lsd);
    case MODEL_SPACECRAFT7: return nx == 7 && nu == 3 && fits<Spacecraft7Model, 8>(N, lsd);
    case MODEL_CARTPOLE: return nx == 4 && nu == 1 && fits<CartpoleModel, 4>(N, lsd);
#ifdef DMPC_CUSTOM_MODEL_FILE
    case MODEL_CUSTOM:
      if constexpr (kCustomG <= 16)
        return nx == CustomModel::NX && nu == CustomModel::NU && fits<CustomModel, kCustomG>(N, lsd);
      return false;
#endif
    default: return false;
  }
}

// register-resident kernels when the instance fits on chip (or when forced),
// else the streamed ones
bool use_stream(int model, int nx, int nu, int N, int lsd, int variant) {
  const int v = variant > 0 ? variant : g_fused_variant;
  if (v == 1) return false;
  if (v >= 2) return stream_supported(model, nx, nu);
  return !register_fits(model, nx, nu, N, lsd) && stream_supported(model, nx, nu);
}

}  // namespace

int g_fused_variant = 0;

// The fused kernels give each time block one lane group of G >= nx lanes,
// lane i owning state row i; control rows are recovered by lanes i < nu of
// the block's group and staged in the per-block vector buffers, so they
// need nu <= nx.  Models with more controls than states run in the plain mode.
bool fused_supported(int model, int nx, int nu, int N) {
  if (N < 1 || nu > nx) return false;
  return register_fits(model, nx, nu, N, 0) || stream_supported(model, nx, nu);
}

size_t fused_work_stride(int model, int nx, int nu, int N, int ls_eval, int n_alphas, int variant) {
  const int lsd = linesearch_doubles(ls_eval, n_alphas, N);
  if (N < 1 || !use_stream(model, nx, nu, N, lsd, variant)) return 0;
  return stream_work_stride(nx, nu, N, ls_eval, n_alphas);
}

// launches the streamed variant shares with the register-resident one
const int32_t* stream_launch_order(const SolveArgs& a, int it, hipStream_t st) {
  // the solution's status array is free until sqp_finish_kernel writes it
  return launch_order(a, it, a.status, st) ? a.status : nullptr;
}

// Longest-first order of an adjoint launch with ragged horizons, into
// order[i * stride]; false (identity order) for a batch with one horizon.
bool stream_backward_order(const BackwardArgs& a, int32_t* order, size_t stride, hipStream_t st) {
  if (!a.horizons) return false;
  return enqueue_order(HorizonOrderKeys{a.horizons, a.N}, a.B, order, stride, st);
}

int stream_launch_reset(const SolveArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(sqp_reset_kernel, dim3((a.B + 255) / 256), dim3(256), 0, st, a.B, a.num_iter, nullptr);
  return hipGetLastError() == hipSuccess ? DMPC_OK : DMPC_ERR_HIP;
}

template <class Model>
int launch_linearize(const SolveArgs& a, hipStream_t st) {
  const long long nlin = linearize_threads<Model>(a.B, a.N);
  prof_begin(PROF_LINEARIZE, st);
  hipLaunchKernelGGL(linearize_kernel<Model>, dim3((unsigned)((nlin + 255) / 256)), dim3(256), 0, st, a);
  prof_end(PROF_LINEARIZE, st);
  return hipGetLastError() == hipSuccess ? DMPC_OK : DMPC_ERR_HIP;
}

int stream_launch_linearize(int model, const SolveArgs& a, hipStream_t st) {
  switch (model) {
    case MODEL_LINEAR:
      if (a.nx == 8) return launch_linearize<LinearModel<8, 4>>(a, st);
      return launch_linearize<LinearModel<4, 2>>(a, st);
    case MODEL_QUADROTOR: return launch_linearize<QuadrotorModel>(a, st);
    case MODEL_SPACECRAFT3: return launch_linearize<Spacecraft3Model>(a, st);
    case MODEL_SPACECRAFT7: return launch_linearize<Spacecraft7Model>(a, st);
    case MODEL_CARTPOLE: return launch_linearize<CartpoleModel>(a, st);
    case MODEL_CHAIN: return launch_linearize<ChainModel>(a, st);
    DMPC_CUSTOM_CASE(return launch_linearize<CustomModel>(a, st))
    default: return DMPC_ERR_UNSUPPORTED;
  }
}

int stream_launch_finish(int model, const SolveArgs& a, hipStream_t st) {
  int nx, nu;
  switch (model) {
    case MODEL_LINEAR: nx = a.nx; nu = nx == 8 ? 4 : 2; break;
    case MODEL_QUADROTOR: nx = QuadrotorModel::NX; nu = QuadrotorModel::NU; break;
    case MODEL_SPACECRAFT3: nx = Spacecraft3Model::NX; nu = Spacecraft3Model::NU; break;
    case MODEL_SPACECRAFT7: nx = Spacecraft7Model::NX; nu = Spacecraft7Model::NU; break;
    case MODEL_CARTPOLE: nx = CartpoleModel::NX; nu = CartpoleModel::NU; break;
    case MODEL_CHAIN: nx = ChainModel::NX; nu = ChainModel::NU; break;
    DMPC_CUSTOM_CASE(nx = CustomModel::NX; nu = CustomModel::NU; break)
    default: return DMPC_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(sqp_finish_kernel, dim3(a.B), dim3(256), 0, st, a, nx, nu);
  return hipGetLastError() == hipSuccess ? DMPC_OK : DMPC_ERR_HIP;
}

int fused_forward(int model, int nx, int nu, const SolveArgs& a, hipStream_t st) {
  if (nu > nx) return DMPC_ERR_UNSUPPORTED;  // see fused_supported
  const int lsd = linesearch_doubles(a.ls_eval, a.n_alphas, a.N);
  if (use_stream(model, nx, nu, a.N, lsd, a.variant)) {
    if (a.ipw > 1) return DMPC_ERR_UNSUPPORTED;  // teams are a register-resident launch shape
    return stream_forward(model, nx, nu, a, st);
  }
  if (!register_fits(model, nx, nu, a.N, lsd)) return DMPC_ERR_UNSUPPORTED;
  switch (model) {
    case MODEL_LINEAR:
      if (nx == 8) return fwd_model<LinearModel<8, 4>, 8>(a, st);
      return fwd_model<LinearModel<4, 2>, 4>(a, st);
    case MODEL_QUADROTOR: return fwd_model<QuadrotorModel, 16>(a, st);
    case MODEL_SPACECRAFT3: return fwd_model<Spacecraft3Model, 4>(a, st);
    case MODEL_SPACECRAFT7: return fwd_model<Spacecraft7Model, 8>(a, st);
    case MODEL_CARTPOLE: return fwd_model<CartpoleModel, 4>(a, st);
#ifdef DMPC_CUSTOM_MODEL_FILE
    case MODEL_CUSTOM:
      if constexpr (kCustomG <= 16) return fwd_model<CustomModel, kCustomG>(a, st);
      return DMPC_ERR_UNSUPPORTED;
#endif
    default: return DMPC_ERR_UNSUPPORTED;
  }
}

int fused_backward(int model, int nx, int nu, const BackwardArgs& a, hipStream_t st) {
  if (nu > nx) return DMPC_ERR_UNSUPPORTED;  // see fused_supported
  if (use_stream(model, nx, nu, a.N, 0, a.variant)) {
    if (a.ipw > 1) return DMPC_ERR_UNSUPPORTED;
    return stream_backward(model, nx, nu, a, st);
  }
  if (!register_fits(model, nx, nu, a.N, 0)) return DMPC_ERR_UNSUPPORTED;
  switch (model) {
    case MODEL_LINEAR:
      if (nx == 8) return bwd_model<LinearModel<8, 4>, 8>(a, st);
      return bwd_model<LinearModel<4, 2>, 4>(a, st);
    case MODEL_QUADROTOR: return bwd_model<QuadrotorModel, 16>(a, st);
    case MODEL_SPACECRAFT3: return bwd_model<Spacecraft3Model, 4>(a, st);
    case MODEL_SPACECRAFT7: return bwd_model<Spacecraft7Model, 8>(a, st);
    case MODEL_CARTPOLE: return bwd_model<CartpoleModel, 4>(a, st);
#ifdef DMPC_CUSTOM_MODEL_FILE
    case MODEL_CUSTOM:
      if constexpr (kCustomG <= 16) return bwd_model<CustomModel, kCustomG>(a, st);
      return DMPC_ERR_UNSUPPORTED;
#endif
    default: return DMPC_ERR_UNSUPPORTED;
  }
}

}  // namespace dmpc
