"""GPU: the device reciprocals triangle setup uses are bit-identical to the oracle's, over every
32-bit input of rcp_fixed and every mantissa of rcp_float at two exponents (gs/shaders/
math_utils.h:27-101): the device evaluates them (pgs_k_rcp_eval) and sums a hash of
(input, output bits) that the oracle's exhaustive sweep computes the same way."""
import ctypes

import pytest

import pgs
from test_pinning_oracle import sweep

pytestmark = pytest.mark.gpu


def device_hash(mode, first, count):
    import torch

    L = pgs.lib()
    L.pgs_k_rcp_eval.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
    h = torch.zeros(1, dtype=torch.int64, device="cuda")
    rc = L.pgs_k_rcp_eval(mode, first, count, h.data_ptr(), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    return int(h.item()) % 2**64


def test_rcp_fixed_bit_identical_over_all_inputs(hip_device):
    with pgs.Context():  # uploads the seed tables to the device
        got = device_hash(0, 1, (1 << 32) - 1)
    ref, _ = sweep("orc_rcp_fixed_sweep", 1, (1 << 32) - 1, threads=16)
    print(f"[case] rcp_fixed 2^32 - 1 inputs: device {got:#x} oracle {ref:#x}", flush=True)
    assert got == ref


@pytest.mark.parametrize("first", [0x3F800000, 0x41000000, 0x2A000000])
def test_rcp_float_bit_identical_over_mantissas(hip_device, first):
    with pgs.Context():
        got = device_hash(1, first, 1 << 23)
    ref, _ = sweep("orc_rcp_float_sweep", first, 1 << 23)
    assert got == ref, (hex(first), hex(got), hex(ref))
