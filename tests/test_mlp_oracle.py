"""The CPU oracle of the tile-cost MLP pinned without the GPU: BF16 rounding against known
values, and the backward formulas against central finite differences of the unquantized
float64 network."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_oracle as M  # noqa: E402


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 0.0, 65504.0], np.float32)
    got = M.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1). 1 + 3*2^-8 ties to 1 + 2^-6.
    exp = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 0.0, 65536.0], np.float32)
    assert np.array_equal(got, exp)


def test_parameter_count():
    assert M.N_PARAMS == 24 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 5825


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_match_finite_differences(seed):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed).astype(np.float64)
    p[M.N_IN * M.N_HID:M.N_IN * M.N_HID + M.N_HID] = rng.normal(0, 0.1, M.N_HID)  # nonzero biases
    X = rng.normal(0, 1, (16, M.N_IN))
    t = rng.normal(0, 1, 16)
    g = M.gradients(p, X, t, quantize=False)

    def loss(q):
        y = M.forward(q, X, quantize=False)
        return np.mean((y - t) ** 2)

    idx = rng.choice(M.N_PARAMS, 60, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (loss(p + e) - loss(p - e)) / 2e-6
        assert abs(fd - g[i]) <= 1e-5 + 1e-4 * abs(fd), (i, fd, g[i])


def test_quantized_forward_is_close_to_float64():
    rng = np.random.default_rng(3)
    p = M.init_params(3)
    X = rng.normal(0, 1, (32, M.N_IN)).astype(np.float32)
    yq = M.forward(p, X)
    y64 = M.forward(p, X, quantize=False)
    assert np.max(np.abs(yq - y64)) < 0.05 * (np.max(np.abs(y64)) + 1)
