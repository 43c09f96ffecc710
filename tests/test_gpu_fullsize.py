"""The measured workloads at their own sizes on the GPU, through the C ABI.

* Porous channel 8192 x 2048 (seeded discs, radii U[4, 16), porosity 0.7,
  tau 0.6, inlet 1.05 / outlet 1.0): every algorithm x layout against the
  oracle after 20 steps -- FP64 against the FP64 instance (1e-12 relative),
  FP32 against the float instance (the reference's USE_FLOAT build).
* The same channel in FP32 and 16-bit storage against the GPU's FP64 after
  500 steps: relative L2 of density and momentum.
* Cylinder channel 8192 x 2048 (r 64 at x 1024), tau 0.52, 0.6 and 0.8,
  inlet 1.02 / outlet 1.0: FP64 against the oracle after 20 steps, and at
  tau 0.52 (where the float 1 / tau differs from 1 / tau rounded from
  double) FP32 against the float instance; 400 steps in every storage at
  tau 0.6 and 0.8, finite populations and bounded L2 / mass difference
  against FP64.
* 512 uneven channels (lbm_batch): persistent and graph launches equal
  per-domain launches bit for bit after 200 steps; sampled domains equal the
  oracle (FP32: the float instance, bit for bit).
"""
import numpy as np
import pytest

import portlbm_amd as pa
from oracle import binding as ob
from portlbm_amd import workloads
from tests.parity import FP32_RTOL, FP64_RTOL, compare

pytestmark = pytest.mark.gpu

ALGS = ["nptl", "lptl", "npol", "nsol"]
LAYOUTS = ["collision", "stream", "bundle"]


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("alg", ALGS)
def test_porous_8192x2048_fp64_matches_oracle(alg, layout):
    doc = workloads.porous_channel(alg, layout, "f64", steps=20)
    r = compare(doc, 20)
    print("porous 8192x2048 f64", alg, layout, r, flush=True)
    assert r["phase_equal"] and r["codes_equal"]
    assert r["pop_err"] <= FP64_RTOL and r["field_err"] <= FP64_RTOL


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("alg", ALGS)
def test_porous_8192x2048_fp32_matches_float_oracle(alg, layout):
    doc = workloads.porous_channel(alg, layout, "f32", steps=20)
    r = compare(doc, 20)
    print("porous 8192x2048 f32", alg, layout, r, flush=True)
    assert r["phase_equal"] and r["codes_equal"]
    assert r["pop_err"] <= FP32_RTOL and r["field_err"] <= FP32_RTOL


def _fields(doc, steps):
    """Fields after `steps`; `finite` from the FP64 sums of the fluid
    populations themselves (the reported fields clamp NaN to 0)."""
    with pa.Lattice.from_settings(doc) as lat:
        lat.run(steps)
        rho, jx, jy, _ = lat.macros()
        red = lat.reduce()
    finite = bool(np.isfinite([red["mass"], red["jx"], red["jy"]]).all())
    return rho, jx, jy, red, finite


def _l2(a, b, mask):
    return float(np.linalg.norm(a[mask] - b[mask]) / np.linalg.norm(b[mask]))


def _against_fp64(make, steps, bounds, label):
    """Each storage's fields after `steps` against FP64 on the GPU: relative
    L2 of density and momentum over fluid nodes and the difference of the FP64
    mass sums; all storages are measured before any bound is checked."""
    rho0, jx0, jy0, red0, ok0 = _fields(make("f64"), steps)
    assert ok0
    fluid = rho0 > -1
    j0 = np.hypot(jx0, jy0)
    out = {}
    for storage in bounds:
        rho, jx, jy, red, ok = _fields(make(storage), steps)
        e_rho = _l2(rho, rho0, fluid)
        e_u = float(np.linalg.norm(np.hypot(jx - jx0, jy - jy0)[fluid]) / np.linalg.norm(j0[fluid]))
        dm = abs(red["mass"] - red0["mass"]) / red0["mass"]
        out[storage] = (ok, e_rho, e_u, dm)
        print(f"{label} {storage} vs f64 after {steps}: rho L2 {e_rho:.3e} u L2 {e_u:.3e} mass {dm:.3e}",
              flush=True)
    for storage, (rho_bound, u_bound) in bounds.items():
        ok, e_rho, e_u, dm = out[storage]
        assert ok and e_rho <= rho_bound and e_u <= u_bound and dm <= rho_bound, (storage, out[storage])


# (rho L2, u L2) bounds, about 3x the values measured on an MI355X after
# 500 steps: FP32 6.8e-7 / 7.7e-5, FP16 5.5e-7 / 3.2e-4, BF16 4.6e-6 / 2.6e-3
POROUS_BOUNDS = {"f32": (2e-6, 2.5e-4), "f16-shifted": (2e-6, 1e-3), "bf16-shifted": (1.5e-5, 8e-3)}


def test_porous_8192x2048_reduced_storage_against_fp64():
    _against_fp64(lambda st: workloads.porous_channel("nptl", "stream", st, steps=500), 500, POROUS_BOUNDS,
                  "porous 8192x2048")


@pytest.mark.parametrize("tau", [0.52, 0.6, 0.8])
def test_cylinder_8192x2048_fp64_matches_oracle(tau):
    r = compare(workloads.cylinder_channel(tau, "f64", steps=20), 20)
    print("cylinder 8192x2048 f64 tau", tau, r, flush=True)
    assert r["ok"], r


def test_cylinder_8192x2048_fp32_tau_052_matches_float_oracle():
    r = compare(workloads.cylinder_channel(0.52, "f32", steps=20), 20)
    print("cylinder 8192x2048 f32 tau 0.52", r, flush=True)
    assert r["phase_equal"] and r["codes_equal"]
    assert r["pop_err"] <= FP32_RTOL and r["field_err"] <= FP32_RTOL


# measured after 400 steps (tau 0.6 / 0.8): FP32 3.4e-7 / 1.7e-4 and 2.3e-7 /
# 1.4e-4, FP16 8.6e-7 / 1.2e-3 and 1.2e-6 / 1.2e-3, BF16 1.2e-5 / 1.0e-2 and
# 1.9e-5 / 1.6e-2; bounds about 3x the larger of each
CYL_BOUNDS = {"f32": (1e-6, 5e-4), "f16-shifted": (4e-6, 4e-3), "bf16-shifted": (6e-5, 5e-2)}


@pytest.mark.parametrize("tau", [0.6, 0.8])
def test_cylinder_8192x2048_storages_against_fp64(tau):
    _against_fp64(lambda st: workloads.cylinder_channel(tau, st, steps=400), 400, CYL_BOUNDS,
                  f"cylinder 8192x2048 tau {tau}")


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_batch_512_channels_modes_and_oracle(storage):
    steps = 200
    docs = workloads.batch_channels(512, storage=storage, steps=steps)
    results = {}
    for mode in ("per-domain", "persistent", "graph"):
        with pa.Batch(docs, storage=storage) as b:
            b.run(steps, mode)
            results[mode] = [b.macros(i) for i in range(len(docs))]
        print(f"batch 512 {storage} {mode} done", flush=True)
    for mode in ("persistent", "graph"):
        for i in range(len(docs)):
            for a, r in zip(results[mode][i], results["per-domain"][i]):
                assert np.array_equal(a, r), (mode, i)
    rtol = FP32_RTOL if storage == "f32" else FP64_RTOL
    for i in (0, 1, 255, 511):
        s = ob.Simulation(docs[i], real=storage)
        s.run(steps)
        for a, r in zip(results["graph"][i], s.results()):
            den = np.maximum(np.abs(r), 1e-300)
            assert float(np.max(np.abs(a - r) / den)) <= rtol, (storage, i)
        s.close()
