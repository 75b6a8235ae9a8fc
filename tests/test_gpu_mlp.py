"""GPU: the hand-written BF16 MLP kernels against the CPU oracle — forward bit for bit,
gradients within float32 tolerance, and training that fits a cost function."""
import os
import sys

import numpy as np
import pytest

import pgs

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def test_forward_bit_exact(hip_device):
    rng = np.random.default_rng(11)
    p = M.init_params(11)
    p[M.N_IN * M.N_HID:M.N_IN * M.N_HID + M.N_HID] = rng.normal(0, 0.2, M.N_HID)
    X = rng.normal(0, 2, (1000, M.N_IN)).astype(np.float32)
    m = pgs.CostModel(1)
    m.set_params(p)
    got = m.predict(X)
    ref = M.forward(p, X)
    assert np.array_equal(got.view(np.uint32), ref.astype(np.float32).view(np.uint32))


@pytest.mark.parametrize("n", [1, 37, 512])
def test_gradients_match_oracle(hip_device, n):
    rng = np.random.default_rng(n)
    p = M.init_params(n)
    X = rng.normal(0, 1, (n, M.N_IN)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    m = pgs.CostModel(2)
    m.set_params(p)
    got = m.gradients(X, t).astype(np.float64)
    ref = M.gradients(p, X, t)
    scale = np.max(np.abs(ref)) + 1e-12
    assert np.max(np.abs(got - ref)) <= 1e-4 * scale, np.max(np.abs(got - ref)) / scale


def test_training_fits_a_cost_function(hip_device):
    """Targets follow a smooth nonlinear function of the features; Adam on the GPU lowers the
    mean squared error well below the variance of the targets."""
    rng = np.random.default_rng(5)
    X = rng.uniform(0, 1, (20000, M.N_IN)).astype(np.float32)
    t = (3.0 * X[:, 0] + 2.0 * X[:, 1] * X[:, 2] + np.log2(1 + 8 * X[:, 3]) + 0.5).astype(np.float32)
    m = pgs.CostModel(3)
    before = float(np.mean((m.predict(X) - t) ** 2))
    mse = m.train(X, t, epochs=15, batch=256, lr=2e-3, seed=9)
    print("mse before", before, "after", mse, "target variance", float(np.var(t)), flush=True)
    assert mse < 0.05 * float(np.var(t))
    assert mse < before
