// Internal declarations shared by the fused kernels, the per-launch kernels
// and the C ABI.  Not part of the public interface (include/dmpc.h is).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dmpc.h"

namespace dmpc {

constexpr double kPsdMinEig = 1e-9;  // diffmpc/solvers/sqp.py:191-196
constexpr int kMaxAlphas = 8;

// Everything one batched SQP solve needs, flattened for kernel arguments.
struct SolveArgs {
  int B, N, P, scheme, penalize;
  int nx, nu;             // model sizes (the linear model has two builds)
  int forced_K, variant;  // per-call tuning (0 = library default)
  float dt;       // step of the float32 linearization
  double dt64;    // step of the float64 line-search evaluations
  int maxit;
  double tol;
  int pcg_max;
  double pcg_eps;
  int linesearch;  // record alphas in the history
  int ls_eval;     // evaluate the merit line search (off for the alphas = [1] no-op)
  double ls_eta;
  int n_alphas;
  double alphas[kMaxAlphas];
  // problem data (device)
  const double *x0, *xf, *xref, *uref, *wx, *wu, *wf;
  const float* dyn;
  const int32_t* horizons;  // (B,) per-instance horizon, or null (all N)
  // solution (device, in: guess and guess histories, out: solution)
  double *x, *u, *lam;
  int32_t *status, *num_iter;
  double *conv_err, *conv_errs, *ls_alphas;
  int32_t* pcg_iters;
  // last linearization point, kept for the backward
  float* AB;      // (B, N, nx, nx+nu): [A_t | B_t]
  double* fd;     // (B, N, nx): f_d(x_t, u_t) in float64
  double* qinv;   // (B, N+1, nx) diagonal of Qinv
  double* rinv;   // (B, nu) diagonal of Rinv (the same for every t)
  // workspace of the streamed variant: work_stride bytes per instance
  char* work;
  size_t work_stride;
  // launch order of the per-instance workgroups of one QP launch (workgroup
  // i solves instance order[i]), or null for the identity
  const int32_t* order;
  // streamed variant, time-invariant model: one copy of the blocks (see
  // shared_block in dmpc_stream.hip)
  int shared_blocks;
  // instances per workgroup of the register-resident kernels (0 or 1: one);
  // with more, each instance is a team of team_threads threads with
  // team_lds bytes of LDS (set at launch)
  int ipw, team_threads, team_lds;
  // forward schedule of the register-resident kernels (0 automatic, 1 one
  // launch per SQP iteration)
  int schedule;
};

struct BackwardArgs {
  int B, N, penalize, pcg_max;
  int forced_K, variant;  // per-call tuning (0 = library default)
  double pcg_eps;
  const double *x0, *xf, *xref, *uref, *wx, *wu, *wf;
  const int32_t* horizons;             // (B,) or null
  const double *x, *u;                 // solution
  const float* AB;                     // saved linearization
  const double *qinv, *rinv;
  const double *dl_dx, *dl_du;         // cotangents (B,N+1,nx), (B,N+1,nu)
  double* lam_bwd;                     // in: adjoint warm start, out: lam~
  double *g_wx, *g_wu, *g_wf, *g_x0, *g_xref, *g_uref, *g_xf;
  double *adj_x, *adj_u;               // optional: x~ (B,N+1,nx), u~ (B,N+1,nu)
  int32_t* pcg_iters;                  // (B,)
  char* work;                          // streamed variant workspace
  size_t work_stride;
  // launch order: workgroup i solves instance order[i * order_stride], or
  // null for the identity
  const int32_t* order;
  size_t order_stride;
  int shared_blocks;                   // as in SolveArgs
  int ipw, team_threads, team_lds;     // as in SolveArgs
};

// SQP continue rule of sqp.py:460-466 for instance b:
// (num_iter < maxit && err > tol) || num_iter < 1.
__device__ __forceinline__ bool sqp_active(const SolveArgs& a, int b) {
  const int it = a.num_iter[b];
  return (it < a.maxit && a.conv_err[b] > a.tol) || it < 1;
}

// Instance b's horizon, clamped to [1, N] so a bad entry cannot index out of
// the padded arrays.
__device__ __forceinline__ int horizon_of(const int32_t* h, int N, int b) {
  if (!h) return N;
  const int v = h[b];
  return v < 1 ? 1 : (v > N ? N : v);
}

// fused path (dmpc_fused.hip)
int fused_forward(int model, int nx, int nu, const SolveArgs& a, hipStream_t s);
int fused_backward(int model, int nx, int nu, const BackwardArgs& a, hipStream_t s);
bool fused_supported(int model, int nx, int nu, int N);

// per-launch path (dmpc_plain.hip)
int plain_forward(int model, int nx, int nu, const SolveArgs& a, hipStream_t s);
int plain_backward(int model, int nx, int nu, const BackwardArgs& a, hipStream_t s);

// tuning knobs: thread groups handle K time blocks each (0 = automatic);
// fused variant (0 automatic, 1 register-resident, 2 streamed)
extern int g_forced_K;
extern int g_fused_variant;

// bytes of workspace per instance the fused forward / backward need (0 when
// the register-resident variant runs)
size_t fused_work_stride(int model, int nx, int nu, int N, int ls_eval, int n_alphas, int variant);

// optional HIP-event timing of individual kernel launches (dmpc_profile_*)
enum ProfKernel { PROF_QP_ITERATION = 0, PROF_BACKWARD = 1, PROF_LINEARIZE = 2, PROF_DYN_GRADS = 3, PROF_NKERNELS = 4 };
void prof_begin(int kernel, hipStream_t s);
void prof_end(int kernel, hipStream_t s);

}  // namespace dmpc
