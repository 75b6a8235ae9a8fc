"""GPU: pgs_k_mlp_gradients_noalloc (fixed device scratch, no allocation) against
pgs_k_mlp_gradients (caller workspace) through the C ABI, bit for bit, on two calls in a row,
the second with fewer samples. Device buffers come from torch; the entry point itself takes
only the caller's inputs and outputs."""
import ctypes
import os
import sys

import numpy as np
import pytest

import pgs

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu

NOALLOC_MAX_SAMPLES = 4096
SAMPLE_FLOATS = 281


class GradArgs(ctypes.Structure):
    _fields_ = [("weights", ctypes.c_void_p), ("features", ctypes.c_void_p), ("targets", ctypes.c_void_p),
                ("scratch", ctypes.c_void_p), ("grads", ctypes.c_void_p), ("loss", ctypes.c_void_p),
                ("count", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


def bf16_bits(p):
    """Round-to-nearest-even BF16 of a float32 vector, as uint16 bits."""
    u = np.asarray(p, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return (u >> 16).astype(np.uint16)


def run_pair(L, torch, weights, n, seed):
    rng = np.random.default_rng(seed)
    X = torch.from_numpy(rng.normal(0, 1, (n, M.N_IN)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.normal(0, 1, n).astype(np.float32)).cuda()
    ws = torch.empty(n * SAMPLE_FLOATS, dtype=torch.float32, device="cuda")
    g_ref = torch.full((M.N_PARAMS,), float("nan"), dtype=torch.float32, device="cuda")
    g_new = torch.full((M.N_PARAMS,), float("nan"), dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    a = GradArgs(weights.data_ptr(), X.data_ptr(), t.data_ptr(), ws.data_ptr(), g_ref.data_ptr(), None, n, 0)
    assert L.pgs_k_mlp_gradients(ctypes.byref(a), stream) == 0
    b = GradArgs(weights.data_ptr(), X.data_ptr(), t.data_ptr(), None, g_new.data_ptr(), None, n, 0)
    assert L.pgs_k_mlp_gradients_noalloc(ctypes.byref(b), stream) == 0
    torch.cuda.synchronize()
    ref, got = g_ref.cpu().numpy(), g_new.cpu().numpy()
    assert np.isfinite(ref).all()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"n={n}: gradients differ"
    return ref


def test_noalloc_gradients_match_workspace_entry(hip_device):
    import torch

    L = pgs.lib()
    L.pgs_k_mlp_gradients.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.pgs_k_mlp_gradients_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    p = M.init_params(17)
    weights = torch.from_numpy(bf16_bits(p).view(np.int16)).cuda()
    print("[case] noalloc gradients n=4096 then n=300", flush=True)
    run_pair(L, torch, weights, NOALLOC_MAX_SAMPLES, 1)
    g_small = run_pair(L, torch, weights, 300, 2)
    # The gradients of the second (smaller) call are those of its own samples: the oracle's
    # float64 gradients agree within float32 tolerance.
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (300, M.N_IN)).astype(np.float32)
    t = rng.normal(0, 1, 300).astype(np.float32)
    ref = M.gradients(p, X, t)
    scale = np.max(np.abs(ref)) + 1e-12
    assert np.max(np.abs(g_small.astype(np.float64) - ref)) <= 1e-4 * scale


def test_noalloc_rejects_oversize(hip_device):
    import torch

    L = pgs.lib()
    L.pgs_k_mlp_gradients_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    dummy = torch.zeros(16, dtype=torch.float32, device="cuda")
    a = GradArgs(dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), None, dummy.data_ptr(), None,
                 NOALLOC_MAX_SAMPLES + 1, 0)
    assert L.pgs_k_mlp_gradients_noalloc(ctypes.byref(a), None) != 0
