"""Shared comparison helpers of the parity tests (GPU library vs CPU oracle)."""
from __future__ import annotations

import numpy as np

import portlbm_amd as pa
from oracle import binding as ob

# Plain FP64 path: populations, density and velocity within 1e-12 relative.
FP64_RTOL = 1e-12
# FP32 storage against the oracle's float instance (the reference's USE_FLOAT
# build): the same operations in the same types and order.
FP32_RTOL = 1e-6


def oracle_real(doc: dict) -> str:
    """The oracle instance a settings document is checked against."""
    storage = doc.get("mi355x", {}).get("storage", "f64")
    if storage not in ("f64", "f32"):
        raise ValueError(f"no oracle instance for storage {storage}")
    return storage


def fluid_slots(phase: np.ndarray, layout: str, N: int) -> np.ndarray:
    """Flat population indices of the nine slots of every fluid node."""
    nodes = np.flatnonzero(phase == 0).astype(np.int64)
    d = np.arange(9, dtype=np.int64)[None, :]
    n = nodes[:, None]
    if layout == "collision":
        idx = 9 * n + d
    elif layout == "stream":
        idx = N * d + n
    else:
        idx = 3 * (d // 3) * N + (d % 3) + 3 * n
    return idx.ravel()


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / max(|b|, tiny), elementwise; NaNs must coincide, equal
    values (infinities included) count as no error."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    if not np.array_equal(nan_a, nan_b):
        return float("inf")
    m = ~nan_a & (a != b)
    if not m.any():
        return 0.0
    den = np.maximum(np.abs(b[m]), 1e-300)
    return float(np.max(np.abs(a[m] - b[m]) / den))


def run_pair(doc: dict, steps: int):
    """Run the same settings on the GPU library and on the oracle."""
    lat = pa.Lattice.from_settings(doc)
    sim = ob.Simulation(doc, real=oracle_real(doc))
    gpu_phase = lat.phase()
    if steps:
        lat.run(steps, doc["algorithmic"]["frameUpdateInterval"])
        sim.run(steps)
    return lat, sim, gpu_phase


def compare(doc: dict, steps: int, rtol: float | None = None) -> dict:
    if rtol is None:
        rtol = FP64_RTOL if oracle_real(doc) == "f64" else FP32_RTOL
    lat, sim, gpu_phase = run_pair(doc, steps)
    try:
        ophase = sim.phase()
        layout = doc["algorithmic"]["dataLayout"]
        out = {"phase_equal": bool(np.array_equal(gpu_phase, ophase)),
               "codes_equal": bool(np.array_equal(lat.node_codes(), sim.node_codes()))}
        idx = fluid_slots(ophase, layout, sim.N)
        out["pop_err"] = rel_err(lat.populations()[idx], sim.populations()[idx])
        g, o = lat.macros(), sim.results()
        out["field_err"] = max(rel_err(a, b) for a, b in zip(g, o))
        out["fluid_nodes"] = int((ophase == 0).sum())
        out["ok"] = (out["phase_equal"] and out["codes_equal"] and out["pop_err"] <= rtol
                     and out["field_err"] <= rtol)
        return out
    finally:
        lat.close()
        sim.close()
