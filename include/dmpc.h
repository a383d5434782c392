/* dmpc.h -- C ABI of the MI355X batched SQP/PCG solver (libdmpc.so).
 *
 * All data pointers are DEVICE pointers to contiguous, batch-major arrays;
 * every call is asynchronous on the given HIP stream (passed as void*) and
 * returns DMPC_OK or a negative error code.  The fused mode allocates
 * nothing: every buffer is the caller's.  The per-launch (plain) mode keeps a
 * device scratch per device that grows on demand and lives until the process
 * ends.  N is the reference's `horizon`: trajectories have N+1 knots and the
 * last control row is zero.
 *
 * Reference interfaces replaced (ToyotaResearchInstitute/diffmpc):
 *   dmpc_sqp_solve      SQPSolver.solve / custom_vjp primal   diffmpc/solvers/sqp.py:572-603 (+ _solve :436-560)
 *   dmpc_sqp_backward   custom_vjp backward                    diffmpc/solvers/sqp.py:668-798
 *   dmpc_linearize      get_dynamics_linearized_matrices_...   diffmpc/problems/optimal_control_problem.py:325-398
 *   dmpc_pcg_assemble   PCGOptimalControl.compute_S_Phiinv     diffmpc/solvers/pcg.py:160-243
 *   dmpc_pcg_gamma      PCGOptimalControl.compute_gamma        diffmpc/solvers/pcg.py:245-288
 *   dmpc_pcg_solve      PCGOptimalControl.solve_KKT_Schur      diffmpc/solvers/pcg.py:350-445
 *   dmpc_pcg_recover    get_states_controls_from_kkt_mult...   diffmpc/solvers/pcg.py:290-348
 *   dmpc_dynamics_param_grads  _compute_mixed_derivatives + dL_dtheta for
 *                       dynamics parameters among the weights  diffmpc/solvers/sqp.py:605-660, 714-737
 *   dmpc_simulate       rollout of demonstrated controls       benchmarking/imitation-learning/utils.py:20-31
 *                       (and predict_next_state per step)      diffmpc/dynamics/integrators.py:20-71
 */
#ifndef DMPC_H_
#define DMPC_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMPC_OK 0
#define DMPC_ERR_ARG (-1)          /* null pointer, bad size or option */
#define DMPC_ERR_UNSUPPORTED (-2)  /* model / size / option not built in this mode */
#define DMPC_ERR_HIP (-3)          /* a HIP call failed */
#define DMPC_ERR_TOO_LARGE (-4)    /* workspace beyond DMPC_NOALLOC_SCRATCH_BYTES (dmpc_sqp_solve_noalloc) */

/* Size of the library-owned device scratch of dmpc_sqp_solve_noalloc, per device. */
#define DMPC_NOALLOC_SCRATCH_BYTES (64u << 20)

/* dynamics models; flat per-instance parameter layouts (float32, n_params) */
#define DMPC_MODEL_LINEAR 0      /* A (nx*nx, row-major), B (nx*nu), b (nx)   linear_dynamics.py:46-70 */
#define DMPC_MODEL_QUADROTOR 1   /* mass, inertia 3x3 row-major  (13/4)        quadrotor_dynamics.py:84-133 */
#define DMPC_MODEL_SPACECRAFT3 2 /* inertia_vector (3)           (3/3)         spacecraft_dynamics.py:31-66 */
#define DMPC_MODEL_SPACECRAFT7 3 /* inertia diagonal (3)         (7/3)  quaternion + rate, no reference counterpart */
#define DMPC_MODEL_CARTPOLE 4    /* masscart, masspole, length, gravity (4/1)  cartpole_dynamics.py:36-80 */
#define DMPC_MODEL_CHAIN 5       /* spring k[12], damper c[12]   (48/12) planar chain, no reference counterpart */
#define DMPC_MODEL_CUSTOM 6      /* the user-defined model a custom build was compiled with (any nx <= 64);
                                    stands in for Dynamics subclasses with their own state_dot, whose
                                    Jacobians the reference takes by autodiff (ocp.py:360-381) */

#define DMPC_MODE_FUSED 0 /* one workgroup per instance per SQP iteration, no host synchronisation */
#define DMPC_MODE_PLAIN 1 /* one kernel per operation, host-driven loops (the reference's op graph) */

typedef struct dmpc_problem {
  int32_t model;     /* DMPC_MODEL_* */
  int32_t nx, nu;    /* must match the model (any for LINEAR with a built instantiation) */
  int32_t horizon;   /* N >= 1 */
  int32_t batch;     /* B >= 1 */
  int32_t scheme;    /* 0 Euler, 1 midpoint, 2 RK4 (integrators.py:12-17) */
  int32_t n_params;  /* floats per instance in dyn_params */
  int32_t penalize_control_reference;
  double dt;         /* discretization_resolution */
} dmpc_problem;

typedef struct dmpc_solver_params { /* keys of diffmpc/solvers/params/sqp.yaml */
  int32_t num_scp_iteration_max;
  double tol_convergence;
  int32_t linesearch;
  double linesearch_eta;
  int32_t n_alphas; /* <= 8 */
  double linesearch_alphas[8];
  int32_t pcg_max_iter;
  double pcg_tol_epsilon;
  int32_t warm_start_backward;
  /* Fused-mode tuning for this call (no reference counterpart; 0 = the
   * library default, see dmpc_set_blocks_per_group / dmpc_set_fused_variant):
   * time blocks per lane group of the register-resident kernels (1, 2, 4) and
   * the kernel variant (1 register-resident, 2 streamed, 3 streamed with a
   * copy of the blocks per knot even for time-invariant models). */
  int32_t blocks_per_group;
  int32_t fused_variant;
  /* Instances per workgroup of the register-resident kernels, forward and
   * adjoint (0 or 1: one; up to 8): each instance is a team of threads
   * aligned to 16 lanes, for models with nx <= 8 and one block per lane
   * group.  Without the line search the results are bitwise those of one
   * instance per workgroup.  DMPC_ERR_UNSUPPORTED where the teams do not fit
   * one workgroup or the streamed variant runs.  No reference counterpart. */
  int32_t instances_per_workgroup;
  /* Forward schedule of the register-resident kernels: 0 = automatic (a
   * batch of one horizon with more instances than the device has CUs is
   * solved by one launch, in which each instance's next SQP iteration starts
   * as soon as its previous one is done; other batches by one linearization
   * and one QP launch per SQP iteration),
   * 1 = one linearization and one QP launch per SQP iteration always.  The
   * results are the same bits.  No reference counterpart. */
  int32_t forward_schedule;
} dmpc_solver_params;

typedef struct dmpc_problem_data {
  const double* x0;        /* (B, nx)      initial_state */
  const double* x_final;   /* (B, nx)      final_state */
  const double* x_ref;     /* (B, N+1, nx) reference_state_trajectory */
  const double* u_ref;     /* (B, N+1, nu) reference_control_trajectory */
  const double* w_x;       /* (B, nx)      weights_penalization_reference_state_trajectory */
  const double* w_u;       /* (B, nu)      weights_penalization_control_squared */
  const double* w_f;       /* (B, nx)      weights_penalization_final_state */
  const float* dyn_params; /* (B, n_params) */
  /* Optional (NULL = every instance has `horizon`).  Ragged batches:
   * horizons[b] in [1, horizon] is instance b's own horizon N_b; every
   * per-knot array keeps the padded stride horizon + 1 and instance b uses
   * its first N_b + 1 knots (N_b for the saved linearization).  Outputs past
   * N_b are zero.  No counterpart in the reference, which solves one horizon
   * per compiled solve. */
  const int32_t* horizons; /* (B,) */
} dmpc_problem_data;

typedef struct dmpc_solution { /* SQPSolution fields, sqp.py:38-57 */
  double* states;             /* (B, N+1, nx)  in: initial guess, out: solution */
  double* controls;           /* (B, N+1, nu)  in: initial guess, out: solution */
  double* dual;               /* (B, N+1, nx)  in: multiplier guess, out: multipliers */
  int32_t* status;            /* (B,) 0 SUCCESS, -100 ERROR (fused mode: holds the QP
                                 launches' instance order, or the one-launch schedule's
                                 counter, until the solve's last kernel) */
  int32_t* num_iter;          /* (B,) */
  double* convergence_error;  /* (B,) */
  double* convergence_errors; /* (B, maxit) in: carried from the guess, out */
  double* linesearch_alphas;  /* (B, maxit) in/out */
  int32_t* num_pcg_iters;     /* (B, maxit) in/out */
} dmpc_solution;

typedef struct dmpc_saved { /* QP data of the last linearization point (sqp.py:487-492) */
  float* AB;    /* (B, N, nx, nx+nu): [A_t | B_t] */
  double* fd;   /* (B, N, nx): f_d(x_t, u_t), evaluated in float64 */
  double* qinv; /* (B, N+1, nx): diagonal of Qinv_t */
  double* rinv; /* (B, nu): diagonal of Rinv_t (the same for every t) */
  /* Device workspace of dmpc_workspace_bytes() bytes (NULL when that is 0):
   * the streamed fused variant keeps its preconditioner blocks there, and
   * the PCG vectors that do not fit in LDS.  Contents need not survive
   * between calls. */
  void* work;
  uint64_t work_bytes;
} dmpc_saved;

typedef struct dmpc_grads { /* d loss / d input, per instance */
  double* w_x;     /* (B, nx) */
  double* w_u;     /* (B, nu) */
  double* w_f;     /* (B, nx) */
  double* x0;      /* (B, nx) */
  double* x_ref;   /* (B, N+1, nx) */
  double* u_ref;   /* (B, N+1, nu) */
  double* x_final; /* (B, nx) */
  double* dual_backward; /* (B, N+1, nx) in: adjoint warm start, out: lambda~ */
  int32_t* pcg_iters;    /* (B,) */
  /* Optional (NULL = not written): primal part of the adjoint KKT solution,
   * x~ (B, N+1, nx) and u~ (B, N+1, nu) with a zero last row -- the
   * linsys_sol.states / .controls of sqp.py:703-737; the input of
   * dmpc_dynamics_param_grads. */
  double* adj_states;
  double* adj_controls;
} dmpc_grads;

/* Library version and build target ("gfx950"). */
const char* dmpc_version(void);

/* Sizes of the user-defined model (DMPC_MODEL_CUSTOM) of a custom build;
 * DMPC_ERR_UNSUPPORTED in a build without one. */
int dmpc_custom_model_info(int32_t* nx, int32_t* nu, int32_t* n_params);

/* 1 if the mode supports (model, nx, nu, horizon), else 0. */
int dmpc_supported(int32_t mode, int32_t model, int32_t nx, int32_t nu, int32_t horizon);

/* Batched SQP solve (sqp.py:436-560).  Solution arrays carry the initial guess in. */
int dmpc_sqp_solve(const dmpc_problem* prob, const dmpc_solver_params* params, const dmpc_problem_data* data,
                   dmpc_solution* sol, dmpc_saved* saved, int32_t mode, void* stream);

/* dmpc_sqp_solve in the fused mode without any allocation, for callers that
 * cannot allocate device memory (code under HIP graph capture, real-time
 * control loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The caller passes inputs and outputs as for dmpc_sqp_solve
 * (saved->work and saved->work_bytes are not read); the workspace the
 * streamed variant needs (dmpc_workspace_bytes for DMPC_MODE_FUSED) comes
 * from a library-owned device array of DMPC_NOALLOC_SCRATCH_BYTES per
 * device, held until the process ends.  Returns DMPC_ERR_TOO_LARGE, before
 * launching anything, when the problem needs more.  The calling thread's
 * current device must be the stream's.  Every call on a device shares that
 * device's array: calls that overlap in time on one device (other streams,
 * other threads) race and must be serialised by the caller; the array keeps
 * the last call's contents, which no call reads before writing.  Calls no
 * hipMalloc/hipFree (with dmpc_profile_enable off). */
int dmpc_sqp_solve_noalloc(const dmpc_problem* prob, const dmpc_solver_params* params, const dmpc_problem_data* data,
                           dmpc_solution* sol, dmpc_saved* saved, void* stream);

/* Implicit-differentiation backward (sqp.py:668-798): one adjoint PCG on the
 * saved linearization with q = -dl/dx, r = -dl/du[:-1], C = 0, then the
 * closed-form contraction with the cost's mixed derivatives. */
int dmpc_sqp_backward(const dmpc_problem* prob, const dmpc_solver_params* params, const dmpc_problem_data* data,
                      const double* states, const double* controls, const dmpc_saved* saved, const double* dl_dstates,
                      const double* dl_dcontrols, dmpc_grads* grads, int32_t mode, void* stream);

/* d loss / d dynamics parameters, per instance: the contraction of
 * sqp.py:714-737 for parameters that enter only the dynamics constraints
 * g_{t+1} = f_d(x_t, u_t; p) - x_{t+1} (optimal_control_problem.py:104-151):
 *   grads[b][j] = -sum_{t < N_b} ( dual_{t+1} . d/dp_j (J_t [adj_x_t; adj_u_t])
 *                                  + adj_dual_{t+1} . d f_d / dp_j ),
 * J_t = d f_d / d(x_t, u_t) at (states, controls), evaluated in float64
 * hyper-dual arithmetic at the model's float32 parameters.  states, controls,
 * dual: the forward solution; adj_*: the adjoint solution written by
 * dmpc_sqp_backward (adj_dual = grads->dual_backward).  grads: (B, n_params),
 * the flat layout of DMPC_MODEL_*; columns past the model's own parameter
 * count are not written.  data uses dyn_params and horizons only. */
int dmpc_dynamics_param_grads(const dmpc_problem* prob, const dmpc_problem_data* data, const double* states,
                              const double* controls, const double* dual, const double* adj_states,
                              const double* adj_controls, const double* adj_dual, double* grads, void* stream);

/* Open-loop rollout of controls (B, N+1, nu) from data->x0 with the
 * problem's integrator: states (B, N+1, nx), x_0 = x0, x_{t+1} = f_d(x_t, u_t)
 * for t < N_b, float64 at the float32 dyn_params; knots past N_b are zero.
 * data uses x0, dyn_params and horizons only.  The imitation-learning
 * initial guess of utils.py:20-31 is this with scheme 0 (Euler). */
int dmpc_simulate(const dmpc_problem* prob, const dmpc_problem_data* data, const double* controls, double* states,
                  void* stream);

/* f_d, A, B, C at (states, controls): fd (B,N,nx), A (B,N,nx,nx), Bm (B,N,nx,nu), C (B,N+1,nx). */
int dmpc_linearize(const dmpc_problem* prob, const dmpc_problem_data* data, const double* states,
                   const double* controls, double* fd, double* A, double* Bm, double* C, void* stream);

/* PCGOptimalControl layer, dense general blocks, float64, 1 <= nx <= 64.
 * As_next, As (B,N,nx,nx); Bs (B,N,nx,nu); Qinv (B,N+1,nx,nx); Rinv (B,N,nu,nu);
 * S, Phiinv (B,N+1,nx,3nx). */
int dmpc_pcg_assemble(int32_t B, int32_t N, int32_t nx, int32_t nu, const double* As_next, const double* As,
                      const double* Bs, const double* Qinv, const double* Rinv, double* S, double* Phiinv,
                      void* stream);
/* gamma (B,N+1,nx) from Cs (B,N+1,nx), qs (B,N+1,nx), rs (B,N,nu). */
int dmpc_pcg_gamma(int32_t B, int32_t N, int32_t nx, int32_t nu, const double* As_next, const double* As,
                   const double* Bs, const double* Qinv, const double* Rinv, const double* Cs, const double* qs,
                   const double* rs, double* gamma, void* stream);
/* PCG on S lam = gamma; lam (B,N+1,nx) in: guess, out: solution; iters (B,), eta (B,).
 * Synchronises the stream once per PCG iteration (host-driven loop). */
int dmpc_pcg_solve(int32_t B, int32_t N, int32_t nx, const double* S, const double* Phiinv, const double* gamma,
                   double* lam, int32_t max_iter, double tol_epsilon, int32_t* iters, double* eta, void* stream);
/* states (B,N+1,nx), controls (B,N+1,nu) from lam. */
int dmpc_pcg_recover(int32_t B, int32_t N, int32_t nx, int32_t nu, const double* lam, const double* As_next,
                     const double* As, const double* Bs, const double* Qinv, const double* qs, const double* Rinv,
                     const double* rs, double* states, double* controls, void* stream);

/* Bytes of dmpc_saved.work that dmpc_sqp_solve / dmpc_sqp_backward need for
 * this problem and mode (0 when the fused kernels keep everything on chip).
 * No reference counterpart (the reference allocates inside XLA). */
int dmpc_workspace_bytes(int32_t mode, const dmpc_problem* prob, const dmpc_solver_params* params, uint64_t* bytes);

/* Tuning: time blocks per thread group of the fused kernels (0 = automatic). */
int dmpc_set_blocks_per_group(int32_t k);

/* Tuning: fused variant, 0 = automatic (register-resident blocks when they fit
 * on chip, else streamed), 1 = register-resident only, 2 = streamed only,
 * 3 = streamed only, and a copy of the preconditioner blocks per knot even for
 * models whose linearization is the same at every knot (linear and chain
 * models, which otherwise store and read one copy: same results, for
 * comparison). */
int dmpc_set_fused_variant(int32_t v);

/* Instrumentation (no reference counterpart): when enabled, the library
 * records a HIP event pair on the launch stream around every launch of the
 * fused kernels; dmpc_profile_read sums the elapsed times of one kernel
 * (0 = QP iteration, 1 = backward, 2 = linearize, 3 = dynamics-parameter
 * contraction) and clears them. */
int dmpc_profile_enable(int32_t on);
int dmpc_profile_read(int32_t kernel, double* total_ms, int32_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* DMPC_H_ */
