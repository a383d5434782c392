// Gradient of the loss with respect to the dynamics parameters.
//
// The reference differentiates through any `weights` leaf, dynamics
// parameters included (diffmpc/solvers/sqp.py:605-660: jacfwd over the
// weights of the cost gradient, of the constraints and of the
// multiplier-weighted constraint Jacobian; contraction at :714-737).  The
// cost does not depend on the dynamics parameters and only the dynamics rows
// g_{t+1} = f_d(x_t, u_t; p) - x_{t+1} of the constraints do
// (optimal_control_problem.py:104-151), so for parameter j
//
//   dL/dp_j = - sum_{t < N} [ lam*_{t+1} . d/dp_j (J_t z~_t) + lam~_{t+1} . d f_d / dp_j ]
//
// with J_t = d f_d / d(x_t, u_t) at the solution (x*, u*), lam* its
// multipliers, and (z~, lam~) = (x~, u~, lam~) the adjoint KKT solution of the
// backward.  One hyper-dual evaluation of f_d per (instance, step, parameter)
// gives both terms: the e1 tangent carries z~_t, the e2 tangent parameter j;
// the e2 part is d f_d / dp_j and the e1 e2 part is d/dp_j (J_t z~_t).
// Evaluated in float64.
//
// Work mapping: one workgroup per (instance, parameter); its threads take the
// time steps and reduce in a fixed order (deterministic).
#include "dmpc_device.h"
#include "dmpc_internal.h"
#include "dmpc_models.h"

namespace dmpc {
namespace {

struct DynGradArgs {
  int B, N, P, scheme;
  double dt;
  const float* dyn;          // (B, P)
  const int32_t* horizons;   // (B,) or null
  const double *x, *u, *lam;  // solution (B, N+1, nx|nu|nx)
  const double *xt, *ut, *lamt;  // adjoint solution (B, N+1, nx|nu|nx)
  double* g;                 // (B, P): columns j < NPARAM
};

template <class Model>
__global__ void __launch_bounds__(256) dyn_param_grad_kernel(DynGradArgs a) {
  constexpr int NX = Model::NX, NU = Model::NU;
  __shared__ double red[16];
  const int b = blockIdx.x, j = blockIdx.y;
  const int NBs = a.N + 1;
  const int N = horizon_of(a.horizons, a.N, b);
  const ParamTangent prm{a.dyn + (size_t)b * a.P, j};
  double acc = 0.0;
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    const size_t kx = ((size_t)b * NBs + t) * NX, ku = ((size_t)b * NBs + t) * NU;
    HD x[NX], u[NU], out[NX];
#pragma unroll
    for (int i = 0; i < NX; ++i) x[i] = HD(a.x[kx + i], a.xt[kx + i], 0.0, 0.0);
#pragma unroll
    for (int k = 0; k < NU; ++k) u[k] = HD(a.u[ku + k], a.ut[ku + k], 0.0, 0.0);
    step<Model, HD>(x, u, prm, HD(a.dt), a.scheme, out);
    const size_t kl = kx + NX;  // knot t + 1
#pragma unroll
    for (int i = 0; i < NX; ++i) acc += a.lam[kl + i] * out[i].ab + a.lamt[kl + i] * out[i].b;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) a.g[(size_t)b * a.P + j] = -s;
}

template <class Model>
int launch(const DynGradArgs& a, hipStream_t st) {
  int threads = ((a.N + 63) / 64) * 64;
  threads = threads > 256 ? 256 : threads;
  prof_begin(PROF_DYN_GRADS, st);
  hipLaunchKernelGGL(dyn_param_grad_kernel<Model>, dim3(a.B, Model::NPARAM), dim3(threads), 0, st, a);
  prof_end(PROF_DYN_GRADS, st);
  return hipGetLastError() == hipSuccess ? DMPC_OK : DMPC_ERR_HIP;
}

}  // namespace

int dyn_param_grads_any(const dmpc_problem* p, const dmpc_problem_data* d, const double* x, const double* u,
                        const double* lam, const double* xt, const double* ut, const double* lamt, double* g,
                        hipStream_t st) {
  DynGradArgs a;
  a.B = p->batch;
  a.N = p->horizon;
  a.P = p->n_params;
  a.scheme = p->scheme;
  a.dt = p->dt;
  a.dyn = d->dyn_params;
  a.horizons = d->horizons;
  a.x = x;
  a.u = u;
  a.lam = lam;
  a.xt = xt;
  a.ut = ut;
  a.lamt = lamt;
  a.g = g;
  switch (p->model) {
    case MODEL_LINEAR:
      if (p->nx == 4 && p->nu == 2) return launch<LinearModel<4, 2>>(a, st);
      if (p->nx == 8 && p->nu == 4) return launch<LinearModel<8, 4>>(a, st);
      return DMPC_ERR_UNSUPPORTED;
    case MODEL_QUADROTOR: return launch<QuadrotorModel>(a, st);
    case MODEL_SPACECRAFT3: return launch<Spacecraft3Model>(a, st);
    case MODEL_SPACECRAFT7: return launch<Spacecraft7Model>(a, st);
    case MODEL_CARTPOLE: return launch<CartpoleModel>(a, st);
    case MODEL_CHAIN: return launch<ChainModel>(a, st);
    DMPC_CUSTOM_CASE(return launch<CustomModel>(a, st))
    default: return DMPC_ERR_UNSUPPORTED;
  }
}

}  // namespace dmpc
