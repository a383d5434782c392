// Open-loop simulation of a batch of control sequences.
//
// The imitation-learning benchmark builds each solve's initial guess by
// rolling the demonstrated controls through the model
// (benchmarking/imitation-learning/utils.py:20-31: x_{t+1} = x_t + dt f(x_t, u_t),
// an explicit Euler step) and the closed-loop drivers advance their plant
// with the problem's own integrator (diffmpc/dynamics/integrators.py:20-71).
// Both are this kernel with the problem's scheme: x_0 = x0,
// x_{t+1} = f_d(x_t, u_t) for t < N_b, float64 states at the model's float32
// parameters, knots past an instance's horizon set to zero.
//
// Work mapping: one thread per instance (the recursion is sequential in t);
// the state stays in registers and each step writes one nx-vector.
#include "dmpc_device.h"
#include "dmpc_internal.h"
#include "dmpc_models.h"

namespace dmpc {
namespace {

struct SimArgs {
  int B, N, P, scheme;
  double dt;
  const float* dyn;         // (B, P)
  const int32_t* horizons;  // (B,) or null
  const double* x0;         // (B, nx)
  const double* u;          // (B, N+1, nu)
  double* x;                // (B, N+1, nx)
};

template <class Model>
__global__ void __launch_bounds__(64) simulate_kernel(SimArgs a) {
  constexpr int NX = Model::NX, NU = Model::NU;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.B) return;
  const int NBs = a.N + 1;
  const int N = horizon_of(a.horizons, a.N, b);
  const float* p = a.dyn + (size_t)b * a.P;
  double x[NX], u[NU], nxt[NX];
#pragma unroll
  for (int i = 0; i < NX; ++i) x[i] = a.x0[(size_t)b * NX + i];
  double* xo = a.x + (size_t)b * NBs * NX;
  const double* ui = a.u + (size_t)b * NBs * NU;
#pragma unroll
  for (int i = 0; i < NX; ++i) xo[i] = x[i];
  for (int t = 0; t < N; ++t) {
#pragma unroll
    for (int k = 0; k < NU; ++k) u[k] = ui[(size_t)t * NU + k];
    step<Model, double>(x, u, p, a.dt, a.scheme, nxt);
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      x[i] = nxt[i];
      xo[(size_t)(t + 1) * NX + i] = x[i];
    }
  }
  for (int t = N + 1; t < NBs; ++t)
#pragma unroll
    for (int i = 0; i < NX; ++i) xo[(size_t)t * NX + i] = 0.0;
}

template <class Model>
int launch(const SimArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(simulate_kernel<Model>, dim3((a.B + 63) / 64), dim3(64), 0, st, a);
  return hipGetLastError() == hipSuccess ? DMPC_OK : DMPC_ERR_HIP;
}

}  // namespace

int simulate_any(const dmpc_problem* p, const dmpc_problem_data* d, const double* controls, double* states,
                 hipStream_t st) {
  SimArgs a;
  a.B = p->batch;
  a.N = p->horizon;
  a.P = p->n_params;
  a.scheme = p->scheme;
  a.dt = p->dt;
  a.dyn = d->dyn_params;
  a.horizons = d->horizons;
  a.x0 = d->x0;
  a.u = controls;
  a.x = states;
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
