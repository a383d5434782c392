"""FP64 CPU oracle for the diffmpc SQP/PCG hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in batched float64 numpy, the algorithm of the reference
library's SQP solver (diffmpc/solvers/sqp.py), its Schur-complement PCG QP
solver (diffmpc/solvers/pcg.py), the optimal-control problem linearization
(diffmpc/problems/optimal_control_problem.py) and the dynamics models
(diffmpc/dynamics/*.py).  Each function cites the reference lines it follows.

It is the checker, never the product: only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it.  The product path
(``diffmpc_amd``) never imports or calls anything from here.

Independent techniques are used on purpose so that a mistake shared with the GPU
kernels is unlikely: Jacobians come from complex-step differentiation (the GPU
uses forward-mode dual numbers), block inverses from LU (``np.linalg.inv``; the GPU
uses Cholesky), and the whole QP path is pinned against a dense KKT solve.

Parity pinning (what checks each function; see tests/test_oracle_*.py):
  dynamics.*            reference invariants (skew, rotation, q_left), central FD,
                        and the mechanics each model obeys (tests/test_oracle_invariants.py):
                        cart-pole and chain energy balance, attitude kinematics
                        (q.q' = 0, R' = R S(w)) and Euler's equation (energy and
                        angular-momentum balance), integrator closed forms on
                        linear flows; each check is shown to fail on planted mistakes
  qp.compute_S_Phiinv   dense -G H^-1 G^T, symmetry, Phi^-1 == (-S)^-1 when phi=0
  qp.compute_gamma      dense C + G H^-1 q
  qp.solve_kkt_schur    dense KKT solve; warm-start invariant of the reference test
  qp.recover            stationarity H z + q + G^T lam = 0
  sqp.sqp_solve         LTI closed form (one iteration == exact QP), notebook value
  sqp.sqp_backward      central finite differences of sqp_solve (FP64)
  sqp.linesearch        the alpha=[1.0] no-op, a hand-checked decrease rule on
                        a 1-D case, and the quadrotor notebook's recorded loss
                        (5e-4, the run's measured rounding spread: its solves stop unconverged at their
                        iteration cap, so it pins no tighter)
  rollout.*             the notebooks' printed losses and updated weights
  sqp.dynamics_param_gradient  central FD of full solves (LTI, chain) and of the
                        converged QP's sensitivities (nonlinear models)
  imitation.*           explicit Euler steps; central FD of the loss of
                        converged solves
"""
