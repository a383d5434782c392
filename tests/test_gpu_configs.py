"""GPU parity at the benchmark configurations' own sizes and on the cost
model's PSD-floor branch, through the C ABI, against the float64 oracle.

* the ragged mix (T ~ U{10..200}, LTI / attitude / quadrotor, 3 SQP
  iterations) on the exact path the ragged benchmark times: split each
  system at the register-resident horizon limit, autotune each group's
  launch shape, capture forward + backward of every group as one hipGraph;
  every instance is compared with the oracle at its own horizon;
* the 48-state chain at its benchmark horizon N = 256 (about 1500 PCG
  iterations per solve);
* zero cost weights: Q = 2 diag(w) and R = 2 diag(w_u) are projected onto
  eigenvalues >= 1e-9 (diffmpc/solvers/sqp.py:191-196,
  diffmpc/utils/jax_utils.py:75-95), so Qinv = Rinv = 1e9 on those entries.
"""
import sys

import numpy as np
import pytest
import torch

from diffmpc_amd import _lib
from diffmpc_amd import workloads as W
from oracle import sqp as osqp
from tests.helpers import oracle_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _np(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


class _variant:
    """Forces a fused kernel variant (1 register-resident, 2 streamed)."""

    def __init__(self, v):
        self.lib, self.v = _lib.load(), v

    def __enter__(self):
        _lib.check(self.lib.dmpc_set_fused_variant(self.v), "dmpc_set_fused_variant")

    def __exit__(self, *exc):
        self.lib.dmpc_set_fused_variant(0)


MODEL_IDS = {"linear": (_lib.MODEL_LINEAR, 4, 2), "spacecraft7": (_lib.MODEL_SPACECRAFT7, 7, 3),
             "quadrotor": (_lib.MODEL_QUADROTOR, 13, 4)}


def test_ragged_benchmark_path_matches_oracle_per_instance(cuda):
    """48 instances of the ragged mix (seed 13: attitude and quadrotor
    instances on both sides of their register-resident limits, so both
    kernel variants run), through split_by_horizon, autotune.tune and one
    GraphedStep of forward + backward; each instance equals the oracle at its
    own horizon (states 1e-5, gradient w.r.t. its state weights 1e-4) and its
    knots past that horizon are zero."""
    from diffmpc_amd import autotune
    from diffmpc_amd.graph import GraphedStep

    say("[ragged bench path] 48 instances, T ~ U{10..200}")
    wkey = "weights_penalization_reference_state_trajectory"
    groups = []
    for name, (prob, sp, h) in W.ragged(batch=48, seed=13).items():
        limit = autotune.register_horizon_limit(*MODEL_IDS[name])
        for sub, hs in W.split_by_horizon(prob, h, limit):
            solver, params, guess = W.make_solver(name, sub, sp, device=cuda, horizons=hs)
            best, _ = autotune.tune(solver, params, guess, reps=1)
            say(f"  {name}: {len(hs)} instances, N <= {int(hs.max())}, launch shape {best}")
            w = params[wkey].clone().requires_grad_(True)
            groups.append((name, sub, sp, hs, solver, params, guess, w))
    assert {g[0] for g in groups} == set(MODEL_IDS)

    def step():
        outs = []
        for _, _, _, _, solver, params, guess, w in groups:
            guess.dual_backward_guess.zero_()
            sol = solver.solve(guess, params, {wkey: w})
            loss = 0.5 * (sol.states - params["reference_state_trajectory"]).square().sum()
            (g,) = torch.autograd.grad(loss, [w])
            outs.append((sol.states, sol.controls, g))
        return outs

    graphed = GraphedStep(step)
    outs = graphed.replay()
    torch.cuda.synchronize()
    worst_x = worst_g = 0.0
    for (name, sub, sp, hs, *_), (xs, us, gw) in zip(groups, outs):
        xs, us, gw = _np(xs), _np(us), _np(gw)
        for b, n in enumerate(hs):
            one = oracle_problem(name, W.instance_problem(sub, b, n))
            ref = osqp.sqp_solve(one, sp)
            g = osqp.sqp_backward(one, sp, ref, ref["states"] - one["x_ref"], np.zeros_like(ref["controls"]))
            ex = rel(xs[b, : n + 1], ref["states"][0])
            eu = rel(us[b, : n + 1], ref["controls"][0])
            eg = rel(gw[b], g["w_x"][0])
            worst_x, worst_g = max(worst_x, ex, eu), max(worst_g, eg)
            assert ex <= 1e-5 and eu <= 1e-5, (name, b, int(n), ex, eu)
            assert eg <= 1e-4, (name, b, int(n), eg)
            assert np.all(xs[b, n + 1:] == 0.0) and np.all(us[b, n + 1:] == 0.0)
    say(f"  worst rel: trajectories {worst_x:.2e}, gradients {worst_g:.2e}")


def test_chain_at_benchmark_horizon_matches_oracle(cuda):
    """The 48-state chain at N = 256 (the chain benchmark's horizon; one SQP
    iteration, PCG 1e-8, about 1500 PCG iterations), forward and the
    gradient w.r.t. the state weights and x0."""
    say("[chain] N=256 B=2")
    name, prob, sp = W.chain(batch=2, seed=4, horizon=256)
    solver, params, guess = W.make_solver(name, prob, sp, device=cuda)
    wx = params["weights_penalization_reference_state_trajectory"].clone().requires_grad_(True)
    x0 = params["initial_state"].clone().requires_grad_(True)
    sol = solver.solve(guess, params, {"weights_penalization_reference_state_trajectory": wx, "initial_state": x0})
    gwx, gx0 = torch.autograd.grad(0.5 * sol.states.square().sum(), [wx, x0])
    torch.cuda.synchronize()
    op = oracle_problem(name, prob)
    ref = osqp.sqp_solve(op, sp)
    g = osqp.sqp_backward(op, sp, ref, ref["states"], np.zeros_like(ref["controls"]))
    say("  pcg gpu", _np(sol.num_pcg_iters).tolist(), "oracle", ref["num_pcg_iters"].tolist())
    say("  rel states", rel(_np(sol.states), ref["states"]), "w_x", rel(_np(gwx), g["w_x"]),
        "x0", rel(_np(gx0), g["x0"]))
    assert rel(_np(sol.states), ref["states"]) <= 1e-5
    assert rel(_np(sol.controls), ref["controls"]) <= 1e-5
    assert rel(_np(sol.dual), ref["dual"]) <= 1e-5
    assert np.max(np.abs(_np(sol.num_pcg_iters) - ref["num_pcg_iters"])) <= 2
    assert rel(_np(gwx), g["w_x"]) <= 1e-4
    assert rel(_np(gx0), g["x0"]) <= 1e-4


def _zero_weight_lti():
    """Config-0 LTI (B=3) with w_x[0] = 0, w_u[1] = 0 and a terminal weight
    w_x[1] + w_f[1] = 0.  PCG eps 1e-20 (at most 3000 iterations): the
    solve stops at the rounding floor rather than at a tolerance."""
    name, prob, sp = W.lti(batch=3, seed=0)
    prob = dict(prob)
    for k in ("w_x", "w_u", "w_f"):
        prob[k] = prob[k].copy()
    prob["w_x"][:, 0] = 0.0
    prob["w_u"][:, 1] = 0.0
    prob["w_f"][:, 1] = -prob["w_x"][:, 1]
    sp = dict(sp)
    sp["pcg"] = dict(sp["pcg"], tol_epsilon=1e-20, max_iter=3000)
    return name, prob, sp


@pytest.mark.parametrize("variant", ["register", "streamed", "plain"])
def test_psd_floor_branch_matches_oracle(cuda, variant):
    """Zero weights run the eigenvalue floor: the saved diagonal Qinv / Rinv
    are exactly 1 / 1e-9 on the zero-weight entries (and 1 / (2w) elsewhere),
    in every execution path.

    The QP is then ill-conditioned (cond(S) ~ 2e12) and the recovery
    x = -Qinv (...) multiplies the multipliers' error by 1e9 on the
    zero-weight state components: two float64 algorithms on the identical
    QP (this PCG and a dense solve of S) already differ there by 10-70 %,
    so those components (and dl/dw_x, which contracts them) are rounding
    noise in the reference's algorithm as much as here and are not compared.
    A 1e-8 relative perturbation of A (float32 rounding is ~6e-8) moves the
    well-determined quantities -- multipliers, controls, the weighted state
    components, dl/dx0, dl/dw_u, dl/dx_ref -- by <= 2e-5 relative in the
    oracle, so they are held to 1e-3."""
    say(f"[psd floor] LTI zero weights, {variant}")
    name, prob, sp = _zero_weight_lti()
    mode = "plain" if variant == "plain" else "fused"
    solver, params, guess = W.make_solver(name, prob, sp, device=cuda, mode=mode)
    keys = ["initial_state", "weights_penalization_control_squared", "reference_state_trajectory"]
    w = {k: params[k].clone().requires_grad_(True) for k in keys}
    with _variant({"register": 1, "streamed": 2, "plain": 0}[variant]):
        sol = solver.solve(guess, params, w)
        # tracking loss on the weighted state components and the controls
        mask = torch.ones_like(sol.states)
        mask[:, :, 0] = 0.0
        mask[:, -1, 1] = 0.0
        loss = 0.5 * (mask * sol.states).square().sum() + 0.5 * sol.controls.square().sum()
        grads = torch.autograd.grad(loss, [w[k] for k in keys])
    torch.cuda.synchronize()
    qinv, rinv = _np(sol.qp_parameters.qinv), _np(sol.qp_parameters.rinv)
    floor = 1.0 / 1e-9  # 1 / max(2 w, 1e-9) at w = 0, as the device computes it
    assert np.all(qinv[:, :-1, 0] == floor) and np.all(qinv[:, -1, 1] == floor) and np.all(rinv[:, 1] == floor)
    assert np.array_equal(qinv[:, :-1, 1:], 1.0 / (2.0 * prob["w_x"][:, None, 1:]).repeat(prob["horizon"], 1))
    assert np.array_equal(rinv[:, 0], 1.0 / (2.0 * prob["w_u"][:, 0]))
    op = oracle_problem(name, prob)
    ref = osqp.sqp_solve(op, sp)
    m = _np(mask).astype(bool)
    dl_dx = np.where(m, ref["states"], 0.0)
    g = osqp.sqp_backward(op, sp, ref, dl_dx, ref["controls"])
    for b in range(prob["x0"].shape[0]):
        ex = rel(_np(sol.states)[b][m[b]], ref["states"][b][m[b]])
        eu = rel(_np(sol.controls)[b], ref["controls"][b])
        el = rel(_np(sol.dual)[b], ref["dual"][b])
        say(f"  b={b}: weighted states {ex:.1e}, controls {eu:.1e}, dual {el:.1e}")
        assert ex <= 1e-3 and eu <= 1e-3 and el <= 1e-3
        for k, gk, ok in zip(keys, grads, ("x0", "w_u", "x_ref")):
            e = rel(_np(gk)[b], g[ok][b])
            say(f"    d/d{ok}: {e:.1e}")
            assert e <= 1e-3, (k, e)


@pytest.mark.parametrize("variant,hmin,hmax", [("register", 20, 63), ("streamed", 64, 120)])
def test_longest_first_launch_order_keeps_every_result(cuda, variant, hmin, hmax):
    """With more instances than CUs each QP launch hands the instances to its
    workgroups longest-first (estimated from the horizons, then from the
    previous SQP iteration's PCG counts).  Instances are independent, so 600
    ragged quadrotor instances solved as one batch equal, bit for bit, the
    same instances solved in batches of 200 (at most one instance per CU,
    launched in index order): trajectories, multipliers, PCG counts and the
    gradient w.r.t. the state weights."""
    say(f"[launch order] quadrotor {variant}, 600 instances, N ~ U{{{hmin}..{hmax}}}")
    name, prob, sp = W.quadrotor(batch=600, seed=5, horizon=hmax, sqp_iters=3)
    hs = np.random.default_rng(5).integers(hmin, hmax + 1, size=600).astype(np.int32)
    key = "weights_penalization_reference_state_trajectory"

    def run(idx):
        sub = W.slice_problem(prob, idx)
        solver, params, guess = W.make_solver(name, sub, sp, device=cuda, horizons=hs[idx])
        wx = params[key].clone().requires_grad_(True)
        with _variant(1 if variant == "register" else 2):
            sol = solver.solve(guess, params, {key: wx})
            loss = 0.5 * (sol.states - params["reference_state_trajectory"]).square().sum()
            (g,) = torch.autograd.grad(loss, [wx])
        torch.cuda.synchronize()
        return [_np(sol.states), _np(sol.controls), _np(sol.dual), _np(sol.num_pcg_iters), _np(g)]

    whole = run(np.arange(600))
    parts = [run(np.arange(k, k + 200)) for k in (0, 200, 400)]
    for j, a in enumerate(whole):
        b = np.concatenate([p[j] for p in parts])
        assert np.array_equal(a, b), j
    say("  identical")


@pytest.mark.parametrize("name,horizon,batch", [("chain", 64, 8), ("linear", 300, 16)])
def test_time_invariant_shared_blocks_match_per_knot_blocks(cuda, name, horizon, batch):
    """Linear models linearize to the same [A_t | B_t] at every knot, so the
    streamed variant stores and reads one copy of the blocks (and D_0, D_1,
    D_N of the preconditioner) instead of one per knot.  The results are the
    same bits as the streamed variant with a copy per knot (variant 3):
    trajectories, multipliers, PCG counts and the gradient w.r.t. the state
    weights, on a ragged batch."""
    say(f"[shared blocks] {name} N={horizon} B={batch}")
    if name == "chain":
        _, prob, sp = W.chain(batch=batch, seed=4, horizon=horizon)
    else:
        _, prob, sp = W.lti(batch=batch, seed=8, horizon=horizon)
        sp = dict(sp)
        sp["pcg"] = dict(sp["pcg"], max_iter=2000)
    hs = np.random.default_rng(3).integers(horizon // 2, horizon + 1, size=batch).astype(np.int32)
    key = "weights_penalization_reference_state_trajectory"

    def run(v):
        solver, params, guess = W.make_solver(name, prob, sp, device=cuda, horizons=hs)
        wx = params[key].clone().requires_grad_(True)
        with _variant(v):
            sol = solver.solve(guess, params, {key: wx})
            loss = 0.5 * (sol.states - params["reference_state_trajectory"]).square().sum()
            (g,) = torch.autograd.grad(loss, [wx])
        torch.cuda.synchronize()
        return [_np(sol.states), _np(sol.controls), _np(sol.dual), _np(sol.num_pcg_iters), _np(g)]

    shared, per_knot = run(2), run(3)
    for j, (a, b) in enumerate(zip(shared, per_knot)):
        assert np.array_equal(a, b), j
    say("  identical")
