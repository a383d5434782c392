/* Host-side checks of the C ABI under AddressSanitizer (no GPU needed).
 *
 * Every entry point of include/dmpc.h is called with arguments its checks
 * must refuse (null pointers, sizes out of range, unknown options, a
 * workspace that is too small, a problem too large for the allocation-free
 * entry point), and the host-only queries (dmpc_supported,
 * dmpc_workspace_bytes, dmpc_version) are called over the built models and
 * a range of horizons.  None of these calls reaches a kernel launch or a
 * device pointer, so the program runs on a machine without a GPU; built
 * against libdmpc_asan.so (host code instrumented) it reports any host
 * memory error in the argument checks and the launch planning.
 * Exit status 0 = every call returned the expected code. */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "dmpc.h"

static int failures = 0;

#define EXPECT(call, want)                                                        \
  do {                                                                            \
    int rc_ = (call);                                                             \
    if (rc_ != (want)) {                                                          \
      fprintf(stderr, "%s:%d: %s returned %d, expected %d\n", __FILE__, __LINE__, \
              #call, rc_, (want));                                                \
      ++failures;                                                                 \
    }                                                                             \
  } while (0)

static dmpc_problem quad(int N, int B) {
  dmpc_problem p;
  memset(&p, 0, sizeof(p));
  p.model = DMPC_MODEL_QUADROTOR;
  p.nx = 13;
  p.nu = 4;
  p.horizon = N;
  p.batch = B;
  p.scheme = 2;
  p.n_params = 10;
  p.penalize_control_reference = 1;
  p.dt = 0.02;
  return p;
}

static dmpc_solver_params sqp_params(void) {
  dmpc_solver_params s;
  memset(&s, 0, sizeof(s));
  s.num_scp_iteration_max = 5;
  s.tol_convergence = 0.0;
  s.linesearch = 0;
  s.linesearch_eta = 0.4;
  s.n_alphas = 1;
  s.linesearch_alphas[0] = 1.0;
  s.pcg_max_iter = 100;
  s.pcg_tol_epsilon = 1e-6;
  s.warm_start_backward = 1;
  return s;
}

int main(void) {
  /* host pointers standing in for device buffers: no call below uses them */
  static double dbuf[64];
  static int32_t ibuf[64];
  static float fbuf[64];
  double* d = dbuf;
  int32_t* i = ibuf;
  dmpc_problem p = quad(50, 4);
  dmpc_solver_params s = sqp_params();
  dmpc_problem_data data = {d, d, d, d, d, d, d, fbuf, NULL};
  dmpc_solution sol = {d, d, d, i, i, d, d, d, i};
  dmpc_saved saved = {fbuf, d, d, d, NULL, 0};
  dmpc_grads g;
  memset(&g, 0, sizeof(g));

  if (strncmp(dmpc_version(), "dmpc", 4) != 0) ++failures;

  /* dmpc_sqp_solve: null arguments and every refused field */
  EXPECT(dmpc_sqp_solve(NULL, &s, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_sqp_solve(&p, NULL, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_sqp_solve(&p, &s, NULL, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_sqp_solve(&p, &s, &data, NULL, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_sqp_solve(&p, &s, &data, &sol, NULL, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
  {
    dmpc_problem q = p;
    q.horizon = 0;
    EXPECT(dmpc_sqp_solve(&q, &s, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    q = p;
    q.batch = 0;
    EXPECT(dmpc_sqp_solve(&q, &s, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    q = p;
    q.nx = 12;
    EXPECT(dmpc_sqp_solve(&q, &s, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_UNSUPPORTED);
    q = p;
    q.n_params = 9;
    EXPECT(dmpc_sqp_solve(&q, &s, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    q = p;
    q.scheme = 3;
    EXPECT(dmpc_sqp_solve(&q, &s, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    q = p;
    q.dt = 0.0;
    EXPECT(dmpc_sqp_solve(&q, &s, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    q = p;
    q.model = 42;
    EXPECT(dmpc_sqp_solve(&q, &s, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_UNSUPPORTED);
  }
  {
    dmpc_solver_params t = s;
    t.n_alphas = 9;
    EXPECT(dmpc_sqp_solve(&p, &t, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    t = s;
    t.linesearch = 1;
    t.n_alphas = 0;
    EXPECT(dmpc_sqp_solve(&p, &t, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    t = s;
    t.pcg_max_iter = 0;
    EXPECT(dmpc_sqp_solve(&p, &t, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    t = s;
    t.num_scp_iteration_max = 0;
    EXPECT(dmpc_sqp_solve(&p, &t, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    t = s;
    t.blocks_per_group = 3;
    EXPECT(dmpc_sqp_solve(&p, &t, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    t = s;
    t.fused_variant = 5;
    EXPECT(dmpc_sqp_solve(&p, &t, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    t = s;
    t.instances_per_workgroup = 9;
    EXPECT(dmpc_sqp_solve(&p, &t, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    t.instances_per_workgroup = -1;
    EXPECT(dmpc_sqp_solve(&p, &t, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    t = s;
    t.forward_schedule = 2;
    EXPECT(dmpc_sqp_solve(&p, &t, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
  }
  {
    dmpc_problem_data e = data;
    e.w_u = NULL;
    EXPECT(dmpc_sqp_solve(&p, &s, &e, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    dmpc_solution o = sol;
    o.num_pcg_iters = NULL;
    EXPECT(dmpc_sqp_solve(&p, &s, &data, &o, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    dmpc_saved v = saved;
    v.fd = NULL;
    EXPECT(dmpc_sqp_solve(&p, &s, &data, &sol, &v, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
  }
  EXPECT(dmpc_sqp_solve(&p, &s, &data, &sol, &saved, 7, NULL), DMPC_ERR_ARG);
  {
    /* the streamed variant (N = 200) needs a workspace: none, then too small */
    dmpc_problem q = quad(200, 4);
    EXPECT(dmpc_sqp_solve(&q, &s, &data, &sol, &saved, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    dmpc_saved v = saved;
    v.work = dbuf;
    v.work_bytes = sizeof(dbuf);
    EXPECT(dmpc_sqp_solve(&q, &s, &data, &sol, &v, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
    /* the allocation-free entry point refuses a workspace beyond its scratch */
    q.batch = 4096;
    EXPECT(dmpc_sqp_solve_noalloc(&q, &s, &data, &sol, &saved, NULL), DMPC_ERR_TOO_LARGE);
    EXPECT(dmpc_sqp_solve_noalloc(NULL, &s, &data, &sol, &saved, NULL), DMPC_ERR_ARG);
  }
  /* more controls than states: not built in the fused mode */
  EXPECT(dmpc_supported(DMPC_MODE_FUSED, DMPC_MODEL_LINEAR, 2, 3, 10), 0);

  /* dmpc_sqp_backward */
  EXPECT(dmpc_sqp_backward(NULL, &s, &data, d, d, &saved, d, d, &g, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_sqp_backward(&p, &s, &data, d, d, &saved, d, d, &g, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_sqp_backward(&p, &s, &data, NULL, d, &saved, d, d, &g, DMPC_MODE_FUSED, NULL), DMPC_ERR_ARG);

  /* the other entry points */
  EXPECT(dmpc_linearize(NULL, &data, d, d, d, d, d, d, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_simulate(NULL, &data, d, d, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_simulate(&p, &data, NULL, d, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_dynamics_param_grads(NULL, &data, d, d, d, d, d, d, d, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_pcg_assemble(0, 1, 1, 1, d, d, d, d, d, d, d, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_pcg_assemble(1, 1, 65, 1, d, d, d, d, d, d, d, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_pcg_gamma(1, 0, 2, 1, d, d, d, d, d, d, d, d, d, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_pcg_solve(0, 1, 1, d, d, d, d, 1, 0.0, i, d, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_pcg_solve(1, 1, 1, NULL, d, d, d, 1, 0.0, i, d, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_pcg_recover(1, 1, 1, 1, NULL, d, d, d, d, d, d, d, d, d, NULL), DMPC_ERR_ARG);
  EXPECT(dmpc_set_blocks_per_group(3), DMPC_ERR_ARG);
  EXPECT(dmpc_set_blocks_per_group(0), DMPC_OK);
  EXPECT(dmpc_set_fused_variant(4), DMPC_ERR_ARG);
  EXPECT(dmpc_set_fused_variant(0), DMPC_OK);
  {
    double ms = 0.0;
    int32_t n = 0;
    EXPECT(dmpc_profile_read(9, &ms, &n), DMPC_ERR_ARG);
    EXPECT(dmpc_profile_read(0, NULL, &n), DMPC_ERR_ARG);
  }
  {
    int32_t a = 0, b = 0, c = 0;
    EXPECT(dmpc_custom_model_info(NULL, &b, &c), DMPC_ERR_ARG);
    EXPECT(dmpc_custom_model_info(&a, &b, &c), DMPC_ERR_UNSUPPORTED); /* no custom model in this build */
  }

  /* host-side planning over every built model and a range of horizons */
  {
    const int models[][3] = {{DMPC_MODEL_LINEAR, 4, 2},      {DMPC_MODEL_LINEAR, 8, 4},   {DMPC_MODEL_QUADROTOR, 13, 4},
                             {DMPC_MODEL_SPACECRAFT3, 3, 3}, {DMPC_MODEL_SPACECRAFT7, 7, 3},
                             {DMPC_MODEL_CARTPOLE, 4, 1},    {DMPC_MODEL_CHAIN, 48, 12}};
    const int nparams[] = {4 * 4 + 4 * 2 + 4, 8 * 8 + 8 * 4 + 8, 10, 3, 3, 4, 24};
    for (size_t m = 0; m < sizeof(models) / sizeof(models[0]); ++m) {
      for (int N = 1; N <= 300; N += 13) {
        for (int mode = 0; mode <= 1; ++mode) {
          dmpc_problem q;
          memset(&q, 0, sizeof(q));
          q.model = models[m][0];
          q.nx = models[m][1];
          q.nu = models[m][2];
          q.horizon = N;
          q.batch = 3;
          q.scheme = 2;
          q.n_params = nparams[m];
          q.dt = 0.1;
          const int sup = dmpc_supported(mode, q.model, q.nx, q.nu, N);
          if (sup != 0 && sup != 1) ++failures;
          for (int ls = 0; ls <= 1; ++ls) {
            dmpc_solver_params t = sqp_params();
            t.linesearch = ls;
            t.n_alphas = ls ? 5 : 1;
            for (int k = 0; k < t.n_alphas; ++k) t.linesearch_alphas[k] = 1.0 / (k + 1);
            uint64_t bytes = 0;
            EXPECT(dmpc_workspace_bytes(mode, &q, &t, &bytes), DMPC_OK);
          }
        }
      }
    }
    uint64_t bytes = 0;
    EXPECT(dmpc_workspace_bytes(5, &p, &s, &bytes), DMPC_ERR_ARG);
    EXPECT(dmpc_workspace_bytes(0, &p, &s, NULL), DMPC_ERR_ARG);
  }

  if (failures) {
    fprintf(stderr, "%d unexpected return codes\n", failures);
    return 1;
  }
  printf("check_abi: all argument checks returned the expected codes\n");
  return 0;
}
