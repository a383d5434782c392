"""Host-side pieces of the imitation-learning pipeline (no GPU): the seeded
initial-state draws of collect_expert_data.py:35-39, the solver settings of
collect_expert_data.py:28-33 and train_imitation_diffmpc.py:47-53, the
.npz demonstrations format and the oracle's cart-pole problem."""
import numpy as np

from diffmpc_amd import imitation as il
from oracle import imitation as oil


def test_initial_states_follow_the_global_generator_draw_order():
    np.random.seed(123)
    want = []
    for _ in range(5):
        want.append([np.random.uniform(low=-0.5, high=0.5), np.random.uniform(low=-0.5, high=0.5),
                     np.random.uniform(low=-3.14, high=3.14), np.random.uniform(low=-1.0, high=1.0)])
    assert np.array_equal(il.sample_initial_states(5, 123), np.array(want))


def test_solver_settings_of_the_two_stages():
    ex = il.expert_solver_params()
    assert (ex["num_scp_iteration_max"], ex["linesearch"], ex["pcg"]["tol_epsilon"]) == (501, True, 1e-24)
    assert ex["tol_convergence"] == 1e-8  # sqp.yaml default, kept
    tr = il.training_solver_params()
    assert (tr["num_scp_iteration_max"], tr["tol_convergence"], tr["linesearch"], tr["pcg"]["tol_epsilon"]) == \
        (5, 1e-1, True, 1e-14)
    assert tr["linesearch_alphas"] == [0.01, 0.1, 0.3, 0.7, 1.0]


def test_demonstrations_round_trip_through_npz(tmp_path):
    rng = np.random.default_rng(0)
    x0, u = rng.normal(size=(3, 4)), rng.normal(size=(3, 101, 1))
    path = str(tmp_path / "demos.npz")
    il.save_demonstrations(path, x0, u)
    x0b, ub = il.load_dataset(path)
    assert np.array_equal(x0, x0b) and np.array_equal(u, ub)


def test_starting_weights_and_problem():
    assert il.INITIAL_WEIGHTS.shape == (20, 4)
    assert np.all((il.INITIAL_WEIGHTS > 0.5) & (il.INITIAL_WEIGHTS < 2.5))
    problem = il.cartpole_problem()
    assert problem.horizon == 100 and int(problem.discretization_scheme) == 0
    prob = oil.cartpole_problem(np.zeros((2, 4)), problem.params)
    assert prob["x_ref"].shape == (2, 101, 4) and prob["u_ref"].shape == (2, 101, 1)
    assert np.allclose(prob["w_x"], [[1.0, 2.0, 1.5, 1.0]] * 2)
    assert np.allclose(prob["dyn"], [[1.0, 0.1, 0.5, 9.81]] * 2, rtol=1e-7)
    assert prob["dt"] == 0.05 and prob["scheme"] == 0
