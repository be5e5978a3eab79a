"""The planner's MLP cost model: the numpy restatement (oracle/mlp.py) is pinned by finite
differences on the CPU; the HIP forward, loss, gradient and training kernels are checked
against it on the GPU."""

import numpy as np
import pytest

from oracle import mlp as ref


def random_problem(rows, seed):
    rng = np.random.default_rng(seed)
    w = rng.normal(scale=0.3, size=ref.N_WEIGHTS)
    X = rng.uniform(0, 1, size=(rows, ref.N_IN))
    T = np.sin(X @ rng.normal(size=ref.N_IN)) + 0.1 * rng.normal(size=rows)
    return w, X, T


def test_oracle_gradient_finite_differences():
    w, X, T = random_problem(40, 1)
    _, g = ref.loss_grad(w, X, T)
    rng = np.random.default_rng(2)
    for i in rng.choice(ref.N_WEIGHTS, size=60, replace=False):
        e = np.zeros_like(w)
        e[i] = 1e-6
        fd = (ref.loss(w + e, X, T) - ref.loss(w - e, X, T)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd))


def test_oracle_layout_matches_header():
    # W1[64][24], b1[64], W2[64][64], b2[64], w3[64], b3 -> 5825 weights (projectq_amd/csrc/mlp.hpp)
    assert ref.N_WEIGHTS == 64 * 24 + 64 + 64 * 64 + 64 + 64 + 1


@pytest.fixture(scope="module")
def gpu_mlp():
    import projectq_amd
    from projectq_amd._binding import MLP

    m = MLP(projectq_amd.library())
    assert m.n_weights == ref.N_WEIGHTS
    return m


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle(gpu_mlp):
    w, X, T = random_problem(700, 3)
    y_gpu = gpu_mlp.predict(w, X)
    y_ref, _ = ref.forward(w.astype(np.float32).astype(np.float64), X.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(y_gpu - y_ref)) <= 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    loss, g = gpu_mlp.loss_grad(w, X, T)
    loss_ref, g_ref = ref.loss_grad(
        w.astype(np.float32).astype(np.float64), X.astype(np.float32).astype(np.float64), T.astype(np.float32)
    )
    assert abs(loss - loss_ref) <= 1e-4 * max(1.0, loss_ref)
    assert np.max(np.abs(g - g_ref)) <= 1e-4 * max(1.0, np.max(np.abs(g_ref)))


@pytest.mark.gpu
def test_gpu_training_fits_and_is_deterministic(gpu_mlp):
    _, X, T = random_problem(2048, 4)
    base = float(np.mean((T - T.mean()) ** 2))
    w1, loss1 = gpu_mlp.train(X, T, epochs=60, batch=128, lr=2e-3, seed=7)
    w2, loss2 = gpu_mlp.train(X, T, epochs=60, batch=128, lr=2e-3, seed=7)
    assert np.array_equal(w1, w2) and loss1 == loss2
    assert loss1 < 0.25 * base
    assert abs(ref.loss(w1, X, T) - loss1) <= 1e-3 * max(1.0, loss1)
