// C ABI of libdmpc.so (declared in include/dmpc.h): argument checks and
// dispatch to the fused or per-launch implementation.
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "dmpc_internal.h"
#include "dmpc_models.h"

namespace dmpc {
int pcg_layer_assemble(int B, int N, int nx, int nu, const double* As_next, const double* As, const double* Bs,
                       const double* Qinv, const double* Rinv, double* S, double* Phiinv, hipStream_t st);
int pcg_layer_gamma(int B, int N, int nx, int nu, const double* As_next, const double* As, const double* Bs,
                    const double* Qinv, const double* Rinv, const double* Cs, const double* qs, const double* rs,
                    double* gamma, hipStream_t st);
int pcg_layer_solve(int B, int N, int nx, const double* S, const double* Phiinv, const double* gamma, double* lam,
                    int max_iter, double eps, int32_t* iters, double* eta, hipStream_t st);
int pcg_layer_recover(int B, int N, int nx, int nu, const double* lam, const double* As_next, const double* As,
                      const double* Bs, const double* Qinv, const double* qs, const double* Rinv, const double* rs,
                      double* x, double* u, hipStream_t st);
int linearize_any(const dmpc_problem* p, const dmpc_problem_data* d, const double* x, const double* u, double* fd,
                  double* A, double* Bm, double* C, hipStream_t st);
int dyn_param_grads_any(const dmpc_problem* p, const dmpc_problem_data* d, const double* x, const double* u,
                        const double* lam, const double* xt, const double* ut, const double* lamt, double* g,
                        hipStream_t st);
int simulate_any(const dmpc_problem* p, const dmpc_problem_data* d, const double* controls, double* states,
                 hipStream_t st);
}  // namespace dmpc

using namespace dmpc;

// ---------------------------------------------------------------------------
// Launch timing with HIP events on the launch stream (off by default).
namespace dmpc {
namespace {
struct ProfState {
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs[PROF_NKERNELS];
  hipEvent_t open[PROF_NKERNELS] = {};
};
ProfState g_prof;
hipEvent_t take_event() {
  if (!g_prof.pool.empty()) {
    hipEvent_t e = g_prof.pool.back();
    g_prof.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace
void prof_begin(int k, hipStream_t s) {
  if (!g_prof.on) return;
  hipEvent_t e = take_event();
  if (e && hipEventRecord(e, s) == hipSuccess) g_prof.open[k] = e;
}
void prof_end(int k, hipStream_t s) {
  if (!g_prof.on || !g_prof.open[k]) return;
  hipEvent_t e = take_event();
  if (e && hipEventRecord(e, s) == hipSuccess) g_prof.pairs[k].push_back({g_prof.open[k], e});
  g_prof.open[k] = nullptr;
}
}  // namespace dmpc

namespace {

// Sizes of a model; the linear model is built for (nx, nu) = (4, 2) and
// (8, 4), so for it the requested sizes pick the instantiation.
int model_sizes(int model, int* nx, int* nu, int* np, int want_nx = 0, int want_nu = 0) {
  switch (model) {
    case DMPC_MODEL_LINEAR:
      if (want_nx == 8 && want_nu == 4) {
        *nx = 8; *nu = 4;
      } else {
        *nx = 4; *nu = 2;
      }
      *np = *nx * *nx + *nx * *nu + *nx;
      return 0;
    case DMPC_MODEL_QUADROTOR: *nx = 13; *nu = 4; *np = 10; return 0;
    case DMPC_MODEL_SPACECRAFT3: *nx = 3; *nu = 3; *np = 3; return 0;
    case DMPC_MODEL_SPACECRAFT7: *nx = 7; *nu = 3; *np = 3; return 0;
    case DMPC_MODEL_CARTPOLE: *nx = 4; *nu = 1; *np = 4; return 0;
    case DMPC_MODEL_CHAIN: *nx = 48; *nu = 12; *np = 24; return 0;
#ifdef DMPC_CUSTOM_MODEL_FILE
    case DMPC_MODEL_CUSTOM: *nx = CustomModel::NX; *nu = CustomModel::NU; *np = CustomModel::NPARAM; return 0;
#endif
    default: return -1;
  }
}

int check_problem(const dmpc_problem* p) {
  if (!p) return DMPC_ERR_ARG;
  int nx, nu, np;
  if (model_sizes(p->model, &nx, &nu, &np, p->nx, p->nu)) return DMPC_ERR_UNSUPPORTED;
  if (p->nx != nx || p->nu != nu) return DMPC_ERR_UNSUPPORTED;
  if (p->n_params < np) return DMPC_ERR_ARG;
  if (p->horizon < 1 || p->batch < 1) return DMPC_ERR_ARG;
  if (p->scheme < 0 || p->scheme > 2) return DMPC_ERR_ARG;
  if (!(p->dt > 0.0)) return DMPC_ERR_ARG;
  return DMPC_OK;
}

bool tuning_ok(const dmpc_solver_params* s) {
  const int k = s->blocks_per_group, v = s->fused_variant, ipw = s->instances_per_workgroup;
  return (k == 0 || k == 1 || k == 2 || k == 4) && v >= 0 && v <= 3 && ipw >= 0 && ipw <= 8 &&
         (s->forward_schedule == 0 || s->forward_schedule == 1);
}

bool linesearch_is_noop(const dmpc_solver_params* s) {
  // with a single candidate alpha = 1 the merit test cannot change the step
  // (sqp.py:389-403 always returns alpha = 1)
  return s->n_alphas == 1 && s->linesearch_alphas[0] == 1.0;
}

}  // namespace

extern "C" {

const char* dmpc_version(void) { return "dmpc 0.1 gfx950"; }

int dmpc_custom_model_info(int32_t* nx, int32_t* nu, int32_t* n_params) {
  if (!nx || !nu || !n_params) return DMPC_ERR_ARG;
#ifdef DMPC_CUSTOM_MODEL_FILE
  *nx = CustomModel::NX;
  *nu = CustomModel::NU;
  *n_params = CustomModel::NPARAM;
  return DMPC_OK;
#else
  *nx = *nu = *n_params = 0;
  return DMPC_ERR_UNSUPPORTED;
#endif
}

int dmpc_supported(int32_t mode, int32_t model, int32_t nx, int32_t nu, int32_t horizon) {
  int mx, mu, np;
  if (model_sizes(model, &mx, &mu, &np, nx, nu) || mx != nx || mu != nu || horizon < 1) return 0;
  if (mode == DMPC_MODE_FUSED) return fused_supported(model, nx, nu, horizon) ? 1 : 0;
  if (mode == DMPC_MODE_PLAIN) return 1;
  return 0;
}

int dmpc_profile_enable(int32_t on) {
  g_prof.on = on != 0;
  return DMPC_OK;
}

int dmpc_profile_read(int32_t kernel, double* total_ms, int32_t* launches) {
  if (kernel < 0 || kernel >= PROF_NKERNELS || !total_ms || !launches) return DMPC_ERR_ARG;
  double tot = 0.0;
  int n = 0;
  for (auto& pr : g_prof.pairs[kernel]) {
    if (hipEventSynchronize(pr.second) != hipSuccess) return DMPC_ERR_HIP;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, pr.first, pr.second) != hipSuccess) return DMPC_ERR_HIP;
    tot += ms;
    ++n;
    g_prof.pool.push_back(pr.first);
    g_prof.pool.push_back(pr.second);
  }
  g_prof.pairs[kernel].clear();
  *total_ms = tot;
  *launches = n;
  return DMPC_OK;
}

int dmpc_set_blocks_per_group(int32_t k) {
  if (k != 0 && k != 1 && k != 2 && k != 4) return DMPC_ERR_ARG;
  g_forced_K = k;
  return DMPC_OK;
}

int dmpc_set_fused_variant(int32_t v) {
  if (v < 0 || v > 3) return DMPC_ERR_ARG;
  g_fused_variant = v;
  return DMPC_OK;
}

int dmpc_workspace_bytes(int32_t mode, const dmpc_problem* prob, const dmpc_solver_params* params, uint64_t* bytes) {
  int rc = check_problem(prob);
  if (rc) return rc;
  if (!params || !bytes) return DMPC_ERR_ARG;
  *bytes = 0;
  if (mode == DMPC_MODE_PLAIN) return DMPC_OK;
  if (mode != DMPC_MODE_FUSED) return DMPC_ERR_ARG;
  const int ls_eval = params->linesearch && !linesearch_is_noop(params);
  *bytes = (uint64_t)fused_work_stride(prob->model, prob->nx, prob->nu, prob->horizon, ls_eval, params->n_alphas,
                                       params->fused_variant) *
           (uint64_t)prob->batch;
  return DMPC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Library-owned scratch of dmpc_sqp_solve_noalloc: the streamed fused
// variant's workspace (preconditioner blocks, PCG vectors, line-search
// scratch), one copy per device, held until the process ends.  Host code
// never passes `scratch` itself (its host-side name is not a device
// address): the device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_ptr.  The kernels write every
// word of the workspace they read (vectors are cleared, blocks assembled)
// before reading it, so nothing needs clearing between calls.
static __device__ __align__(256) unsigned char scratch[DMPC_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call per device, with that device current).
int scratch_on_current_device(char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return DMPC_ERR_HIP;
  if (dev < 0 || dev >= kMaxDevices) return DMPC_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return DMPC_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<char*>(g_scratch_ptr[dev]);
  return DMPC_OK;
}

// The body of dmpc_sqp_solve and dmpc_sqp_solve_noalloc: with `noalloc` the
// streamed variant's workspace is the device's `scratch` (saved->work is not
// read) and only the fused mode is accepted (the plain mode allocates).
int sqp_solve_impl(const dmpc_problem* prob, const dmpc_solver_params* params, const dmpc_problem_data* data,
                   dmpc_solution* sol, dmpc_saved* saved, int32_t mode, void* stream, bool noalloc) {
  int rc = check_problem(prob);
  if (rc) return rc;
  if (!params || !data || !sol || !saved) return DMPC_ERR_ARG;
  if (!data->x0 || !data->x_final || !data->x_ref || !data->u_ref || !data->w_x || !data->w_u || !data->w_f ||
      !data->dyn_params)
    return DMPC_ERR_ARG;
  if (!sol->states || !sol->controls || !sol->dual || !sol->status || !sol->num_iter || !sol->convergence_error ||
      !sol->convergence_errors || !sol->linesearch_alphas || !sol->num_pcg_iters)
    return DMPC_ERR_ARG;
  if (!saved->AB || !saved->fd || !saved->qinv || !saved->rinv) return DMPC_ERR_ARG;
  if (params->num_scp_iteration_max < 1 || params->pcg_max_iter < 1) return DMPC_ERR_ARG;
  if (params->n_alphas < 0 || params->n_alphas > kMaxAlphas) return DMPC_ERR_ARG;
  if (params->linesearch && (params->n_alphas < 1 || params->n_alphas > 8)) return DMPC_ERR_ARG;
  if (!tuning_ok(params)) return DMPC_ERR_ARG;
  SolveArgs a;
  std::memset(&a, 0, sizeof(a));
  a.B = prob->batch;
  a.N = prob->horizon;
  a.P = prob->n_params;
  a.scheme = prob->scheme;
  a.penalize = prob->penalize_control_reference;
  a.nx = prob->nx;
  a.nu = prob->nu;
  a.forced_K = params->blocks_per_group;
  a.variant = params->fused_variant;
  a.ipw = params->instances_per_workgroup;
  a.schedule = params->forward_schedule;
  a.dt = (float)prob->dt;
  a.dt64 = prob->dt;
  a.maxit = params->num_scp_iteration_max;
  a.tol = params->tol_convergence;
  a.pcg_max = params->pcg_max_iter;
  a.pcg_eps = params->pcg_tol_epsilon;
  a.linesearch = params->linesearch;
  a.ls_eval = params->linesearch && !linesearch_is_noop(params);
  a.ls_eta = params->linesearch_eta;
  a.n_alphas = params->n_alphas;
  for (int k = 0; k < params->n_alphas; ++k) a.alphas[k] = params->linesearch_alphas[k];
  a.x0 = data->x0;
  a.xf = data->x_final;
  a.xref = data->x_ref;
  a.uref = data->u_ref;
  a.wx = data->w_x;
  a.wu = data->w_u;
  a.wf = data->w_f;
  a.dyn = data->dyn_params;
  a.horizons = data->horizons;
  a.x = sol->states;
  a.u = sol->controls;
  a.lam = sol->dual;
  a.status = sol->status;
  a.num_iter = sol->num_iter;
  a.conv_err = sol->convergence_error;
  a.conv_errs = sol->convergence_errors;
  a.ls_alphas = sol->linesearch_alphas;
  a.pcg_iters = sol->num_pcg_iters;
  a.AB = saved->AB;
  a.fd = saved->fd;
  a.qinv = saved->qinv;
  a.rinv = saved->rinv;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (mode == DMPC_MODE_FUSED) {
    a.work_stride = fused_work_stride(prob->model, prob->nx, prob->nu, prob->horizon, a.ls_eval, a.n_alphas, a.variant);
    if (a.work_stride) {
      const uint64_t need = (uint64_t)a.work_stride * (uint64_t)a.B;
      if (noalloc) {
        if (need > (uint64_t)DMPC_NOALLOC_SCRATCH_BYTES) return DMPC_ERR_TOO_LARGE;
        rc = scratch_on_current_device(&a.work);
        if (rc) return rc;
      } else {
        if (!saved->work || saved->work_bytes < need) return DMPC_ERR_ARG;
        a.work = static_cast<char*>(saved->work);
      }
    }
    return fused_forward(prob->model, prob->nx, prob->nu, a, st);
  }
  if (noalloc) return DMPC_ERR_UNSUPPORTED;  // the per-launch mode allocates its scratch
  if (mode == DMPC_MODE_PLAIN) {
    if (data->horizons) return DMPC_ERR_UNSUPPORTED;  // ragged batches: fused mode only
    return plain_forward(prob->model, prob->nx, prob->nu, a, st);
  }
  return DMPC_ERR_ARG;
}
}  // namespace

extern "C" {

int dmpc_sqp_solve(const dmpc_problem* prob, const dmpc_solver_params* params, const dmpc_problem_data* data,
                   dmpc_solution* sol, dmpc_saved* saved, int32_t mode, void* stream) {
  return sqp_solve_impl(prob, params, data, sol, saved, mode, stream, false);
}

int dmpc_sqp_solve_noalloc(const dmpc_problem* prob, const dmpc_solver_params* params, const dmpc_problem_data* data,
                           dmpc_solution* sol, dmpc_saved* saved, void* stream) {
  return sqp_solve_impl(prob, params, data, sol, saved, DMPC_MODE_FUSED, stream, true);
}

int dmpc_sqp_backward(const dmpc_problem* prob, const dmpc_solver_params* params, const dmpc_problem_data* data,
                      const double* states, const double* controls, const dmpc_saved* saved, const double* dl_dstates,
                      const double* dl_dcontrols, dmpc_grads* g, int32_t mode, void* stream) {
  int rc = check_problem(prob);
  if (rc) return rc;
  if (!params || !data || !states || !controls || !saved || !dl_dstates || !dl_dcontrols || !g) return DMPC_ERR_ARG;
  if (!saved->AB || !saved->qinv || !saved->rinv) return DMPC_ERR_ARG;
  if (!tuning_ok(params)) return DMPC_ERR_ARG;
  if (!g->w_x || !g->w_u || !g->w_f || !g->x0 || !g->x_ref || !g->u_ref || !g->x_final || !g->dual_backward ||
      !g->pcg_iters)
    return DMPC_ERR_ARG;
  if (params->pcg_max_iter < 1) return DMPC_ERR_ARG;
  BackwardArgs a;
  std::memset(&a, 0, sizeof(a));
  a.B = prob->batch;
  a.N = prob->horizon;
  a.penalize = prob->penalize_control_reference;
  a.pcg_max = params->pcg_max_iter;
  a.pcg_eps = params->pcg_tol_epsilon;
  a.forced_K = params->blocks_per_group;
  a.variant = params->fused_variant;
  a.ipw = params->instances_per_workgroup;
  a.x0 = data->x0;
  a.xf = data->x_final;
  a.xref = data->x_ref;
  a.uref = data->u_ref;
  a.wx = data->w_x;
  a.wu = data->w_u;
  a.wf = data->w_f;
  a.horizons = data->horizons;
  a.x = states;
  a.u = controls;
  a.AB = saved->AB;
  a.qinv = saved->qinv;
  a.rinv = saved->rinv;
  a.dl_dx = dl_dstates;
  a.dl_du = dl_dcontrols;
  a.lam_bwd = g->dual_backward;
  a.g_wx = g->w_x;
  a.g_wu = g->w_u;
  a.g_wf = g->w_f;
  a.g_x0 = g->x0;
  a.g_xref = g->x_ref;
  a.g_uref = g->u_ref;
  a.g_xf = g->x_final;
  a.adj_x = g->adj_states;
  a.adj_u = g->adj_controls;
  a.pcg_iters = g->pcg_iters;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (mode == DMPC_MODE_FUSED) {
    a.work_stride = fused_work_stride(prob->model, prob->nx, prob->nu, prob->horizon, 0, 0, a.variant);
    if (a.work_stride) {
      if (!saved->work || saved->work_bytes < (uint64_t)a.work_stride * (uint64_t)a.B) return DMPC_ERR_ARG;
      a.work = static_cast<char*>(saved->work);
    }
    return fused_backward(prob->model, prob->nx, prob->nu, a, st);
  }
  if (mode == DMPC_MODE_PLAIN) {
    if (data->horizons) return DMPC_ERR_UNSUPPORTED;
    return plain_backward(prob->model, prob->nx, prob->nu, a, st);
  }
  return DMPC_ERR_ARG;
}

int dmpc_dynamics_param_grads(const dmpc_problem* prob, const dmpc_problem_data* data, const double* states,
                              const double* controls, const double* dual, const double* adj_states,
                              const double* adj_controls, const double* adj_dual, double* grads, void* stream) {
  int rc = check_problem(prob);
  if (rc) return rc;
  if (!data || !data->dyn_params || !states || !controls || !dual || !adj_states || !adj_controls || !adj_dual ||
      !grads)
    return DMPC_ERR_ARG;
  return dyn_param_grads_any(prob, data, states, controls, dual, adj_states, adj_controls, adj_dual, grads,
                             reinterpret_cast<hipStream_t>(stream));
}

int dmpc_simulate(const dmpc_problem* prob, const dmpc_problem_data* data, const double* controls, double* states,
                  void* stream) {
  int rc = check_problem(prob);
  if (rc) return rc;
  if (!data || !data->x0 || !data->dyn_params || !controls || !states) return DMPC_ERR_ARG;
  return simulate_any(prob, data, controls, states, reinterpret_cast<hipStream_t>(stream));
}

int dmpc_linearize(const dmpc_problem* prob, const dmpc_problem_data* data, const double* states,
                   const double* controls, double* fd, double* A, double* Bm, double* C, void* stream) {
  int rc = check_problem(prob);
  if (rc) return rc;
  if (!data || !data->x0 || !data->dyn_params || !states || !controls || !fd || !A || !Bm || !C) return DMPC_ERR_ARG;
  return linearize_any(prob, data, states, controls, fd, A, Bm, C, reinterpret_cast<hipStream_t>(stream));
}

int dmpc_pcg_assemble(int32_t B, int32_t N, int32_t nx, int32_t nu, const double* As_next, const double* As,
                      const double* Bs, const double* Qinv, const double* Rinv, double* S, double* Phiinv,
                      void* stream) {
  if (B < 1 || N < 1 || nx < 1 || nu < 1 || nx > 64) return DMPC_ERR_ARG;
  if (!As_next || !As || !Bs || !Qinv || !Rinv || !S || !Phiinv) return DMPC_ERR_ARG;
  return pcg_layer_assemble(B, N, nx, nu, As_next, As, Bs, Qinv, Rinv, S, Phiinv,
                            reinterpret_cast<hipStream_t>(stream));
}

int dmpc_pcg_gamma(int32_t B, int32_t N, int32_t nx, int32_t nu, const double* As_next, const double* As,
                   const double* Bs, const double* Qinv, const double* Rinv, const double* Cs, const double* qs,
                   const double* rs, double* gamma, void* stream) {
  if (B < 1 || N < 1 || nx < 1 || nu < 1) return DMPC_ERR_ARG;
  if (!As_next || !As || !Bs || !Qinv || !Rinv || !Cs || !qs || !rs || !gamma) return DMPC_ERR_ARG;
  return pcg_layer_gamma(B, N, nx, nu, As_next, As, Bs, Qinv, Rinv, Cs, qs, rs, gamma,
                         reinterpret_cast<hipStream_t>(stream));
}

int dmpc_pcg_solve(int32_t B, int32_t N, int32_t nx, const double* S, const double* Phiinv, const double* gamma,
                   double* lam, int32_t max_iter, double tol_epsilon, int32_t* iters, double* eta, void* stream) {
  if (B < 1 || N < 1 || nx < 1 || nx > 64 || max_iter < 1) return DMPC_ERR_ARG;
  if (!S || !Phiinv || !gamma || !lam || !iters || !eta) return DMPC_ERR_ARG;
  return pcg_layer_solve(B, N, nx, S, Phiinv, gamma, lam, max_iter, tol_epsilon, iters, eta,
                         reinterpret_cast<hipStream_t>(stream));
}

int dmpc_pcg_recover(int32_t B, int32_t N, int32_t nx, int32_t nu, const double* lam, const double* As_next,
                     const double* As, const double* Bs, const double* Qinv, const double* qs, const double* Rinv,
                     const double* rs, double* states, double* controls, void* stream) {
  if (B < 1 || N < 1 || nx < 1 || nu < 1 || nx > 64) return DMPC_ERR_ARG;
  if (!lam || !As_next || !As || !Bs || !Qinv || !qs || !Rinv || !rs || !states || !controls) return DMPC_ERR_ARG;
  return pcg_layer_recover(B, N, nx, nu, lam, As_next, As, Bs, Qinv, qs, Rinv, rs, states, controls,
                           reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
