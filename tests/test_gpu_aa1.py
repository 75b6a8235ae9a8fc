"""GPU parity of AA1 edge anti-aliasing (intersect.h:156-244, ubershader.comp:848-851, 1497-1505):
every plane against the oracle on random streams with AA1 primitives (triangles, strips, fans,
lines, line strips; sprites and points ignore AA1) at 1x and super-sampled, and on the closed-form
set-ups of test_aa1_oracle.py."""
import numpy as np
import pytest

import pgs
import render_streams as rs
import test_aa1_oracle as T
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

LOG2 = {1: 0, 2: 1, 4: 2, 8: 3, 16: 4}


def both(stream, ssaa):
    print(f"[case] AA1 stream {len(stream)} bytes ssaa={ssaa}", flush=True)
    with pgs.Context(super_sampling=ssaa) as ctx:
        ctx.gif_transfer(stream)
        got = ctx.vram_planes()
    o = Oracle(LOG2[ssaa])
    o.gif_transfer(stream)
    o.flush()
    ref = o.planes()
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


@pytest.mark.parametrize("ssaa,seed", [(1, 0), (1, 1), (1, 2), (4, 3), (16, 4), (2, 5)])
def test_random_aa1_streams(hip_device, ssaa, seed):
    both(rs.random_render_stream(100 + seed, n_batches=10, prims_per_batch=30, aa1=True), ssaa)


def test_closed_form_setups_on_gpu(hip_device, monkeypatch):
    """The closed-form AA1 tests, each drawn on the GPU too: VRAM must match before the
    closed-form assertions run on the oracle's result."""
    captured = []

    class Mirrored(Oracle):
        def gif_transfer(self, data, path=3):
            captured.append(np.array(data, np.uint8))
            super().gif_transfer(data, path)

        def vram(self):
            ref = super().vram()
            with pgs.Context() as ctx:
                for c in captured:
                    ctx.gif_transfer(c)
                got = ctx.vram()
            captured.clear()
            assert np.array_equal(got, ref)
            return ref

    monkeypatch.setattr(T, "Oracle", Mirrored)
    for fn in (T.test_triangle_vertical_edge_feather, T.test_partial_coverage_does_not_write_depth,
               T.test_feather_falls_off_with_distance, T.test_aa1_line_columns_add_to_one_pixel):
        print(f"[case] {fn.__name__}", flush=True)
        fn()
