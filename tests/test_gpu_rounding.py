"""Stochastic rounding of the FP16/BF16 population stores (lbm_desc.rounding,
settings key mi355x.rounding): deterministic, and on a steady flow closer to
FP64 than round-to-nearest, whose identical error every step acts as a
spurious source of mass and momentum."""
import numpy as np
import pytest

import portlbm_amd as pa
from portlbm_amd import workloads

pytestmark = pytest.mark.gpu


def _run(storage, rounding, steps=6000, algorithm="nptl", layout="stream"):
    doc = workloads.porous_channel(algorithm, layout, storage=storage, steps=steps, inner_h=512, inner_v=128,
                                   tau=0.8)
    doc["mi355x"]["rounding"] = rounding
    doc["mi355x"]["obstacleMargin"] = 16
    with pa.Lattice.from_settings(doc) as lat:
        lat.run(steps)
        return lat.macros()


def _u_err(a, ref):
    f = ref[0] > -1
    d = np.hypot(a[1][f] - ref[1][f], a[2][f] - ref[2][f])
    return float(np.linalg.norm(d) / np.linalg.norm(np.hypot(ref[1][f], ref[2][f])))


def test_stochastic_rounding_is_deterministic_and_differs_from_nearest():
    a = _run("f16-shifted", "stochastic", 300)
    b = _run("f16-shifted", "stochastic", 300)
    c = _run("f16-shifted", "nearest", 300)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[1], c[1])


@pytest.mark.parametrize("storage,bound", [("f16-shifted", 1.5e-3), ("bf16-shifted", 1e-2)])
def test_stochastic_rounding_beats_nearest_on_a_steady_flow(storage, bound):
    ref = _run("f64", "nearest")
    sr = _u_err(_run(storage, "stochastic"), ref)
    rn = _u_err(_run(storage, "nearest"), ref)
    print(storage, "u L2 error: stochastic", sr, "nearest", rn)
    assert sr < 0.5 * rn and sr < bound


@pytest.mark.parametrize("algorithm,layout", [("lptl", "bundle"), ("npol", "collision"), ("nsol", "stream")])
def test_stochastic_rounding_in_every_scheme(algorithm, layout):
    ref = _run("f64", "nearest", 800, algorithm, layout)
    sr = _u_err(_run("f16-shifted", "stochastic", 800, algorithm, layout), ref)
    assert sr < 3e-3


def test_unknown_rounding_is_rejected():
    desc = pa.desc_from_settings(workloads.porous_channel(storage="f16-shifted", inner_h=64, inner_v=32))
    desc.rounding = 2
    with pytest.raises(pa.LbmError):
        pa.Lattice(desc)
