"""A lattice close to the 32-bit population-index limit (9 x places < 2^32):
lptl, FP64, each layout, 70 000 000 x 4 inner nodes (420 M places, 9 x the
direction stride = 3.8e9; 61 GB of lattices).  Influence travels one column
per step, so after 20 steps the first and the last 100 columns must equal a
200 x 4 channel's bit for bit: the inlet side checks the smallest indices, the
outlet side (Zou-He, uy history) the largest."""
import numpy as np
import pytest

import portlbm_amd as pa

pytestmark = pytest.mark.gpu

WIDE = 70_000_000
STEPS = 20


def _doc(width, layout="stream"):
    return pa.settings("lptl", layout, inner_h=width, inner_v=4, scenario="walls", work_group_size=256,
                       time_steps=STEPS, tau=0.6, rho_in=1.05, rho_out=1.0, storage="f64")


def _channel(width):
    """Plain channel phase table (the walls scenario also places a few
    obstacles at width-dependent positions): ghost frame -1, walls on the
    first and last inner rows, fluid between."""
    H = width + 2
    ph = np.full((6, H), -1, dtype=np.int8)
    ph[1, 1:-1] = 1
    ph[4, 1:-1] = 1
    ph[2:4, 1:-1] = 0
    return ph.ravel()


def _lattice(width, layout):
    lat = pa.Lattice(pa.desc_from_settings(_doc(width, layout)))
    lat.set_phase(_channel(width))
    lat.init()
    return lat


@pytest.mark.parametrize("layout", ["stream", "collision", "bundle"])
def test_lattice_near_the_index_limit_matches_a_small_one_at_both_ends(layout):
    desc = pa.desc_from_settings(_doc(WIDE, layout))
    places = (63 + desc.expanded_horizontal_nodes + 63) // 64 * 64 * desc.expanded_vertical_nodes
    assert 9 * places > 0.85 * 2**32  # close to the limit, and accepted
    with _lattice(200, layout) as small:
        small.run(STEPS)
        ref = [a.reshape(4, 200) for a in small.macros()]
    print(f"creating {WIDE} x 4 (9 x places = {9 * places:.3e})", flush=True)
    with _lattice(WIDE, layout) as big:
        big.run(STEPS)
        got = [a.reshape(4, WIDE) for a in big.macros()]
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a[:, :100], b[:, :100])
        np.testing.assert_array_equal(a[:, -100:], b[:, -100:])
    # far from both ends every column of the fluid rows is the same (rest, to rounding)
    mid = got[0][1:3, 1000:2000]
    np.testing.assert_array_equal(mid, np.broadcast_to(mid[:, :1], mid.shape))
    assert np.all(np.abs(mid - 1.0) < 1e-12)
