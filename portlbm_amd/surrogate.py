"""Learned warm start for the steady solve (include/lbm_surrogate.h).

A convolutional encoder-decoder, trained on the GPU by this package's own
HIP kernels on steady states that the batched solver produces, maps an
obstacle mask, the boundary densities and tau to a steady (density,
momentum) field.  The field only warm-starts the lattice from its
equilibrium (lbm_batch_set_equilibrium); the plain solve decides
convergence, and a domain whose residual fails to fall is reset to the
at-rest start (lbm_batch_reset_domain).  The reference solver always starts
from rest (include/lbm/core/domain_initialization.hpp:311-403).

Encoding on an inner grid of ny x nx nodes (the network's H x W):
  inputs   0: fluid mask (1 fluid, 0 solid)
           1: x ramp, 2 (x + 0.5) / nx - 1
           2: (tau - 0.725) / 0.175
           3: (drho - 0.03) / 0.02, drho = rho_in - rho_out
  targets  0: (rho - rho_out) / drho * rho_scale
           1, 2: j (tau - 1/2) / drho * j_scale
The momentum normalisation is Stokes scaling: at low Reynolds number the
steady momentum is proportional to the pressure drop over the viscosity,
nu = (tau - 1/2) / 3, so the network mostly learns the geometry's share.
j_scale (one number per model) brings the normalised momentum to unit RMS
over the training set; rho_scale > 1 weights the density in the loss.

    python -m portlbm_amd.surrogate train --out model.npz --report train.json
    python -m portlbm_amd.surrogate evaluate --model model.npz --report eval.json
    python -m portlbm_amd.surrogate select --model model.npz --out model2.npz --report sel.json
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import Batch, _check_hip, _ptr, build_phase, hip_library, settings

INPUTS, OUTPUTS = 4, 3
TAU_MID, TAU_HALF = 0.725, 0.175
DRHO_MID, DRHO_HALF = 0.03, 0.02
DEFAULT_MODEL = Path(__file__).resolve().parent / "data" / "surrogate_porous_256.npz"

_P, _U32, _SZ = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_size_t
_sig_done = False


def _lib() -> ctypes.CDLL:
    global _sig_done
    lib = hip_library()
    if not _sig_done:
        sig = {
            "lbm_surrogate_param_count": ([], _SZ),
            "lbm_surrogate_create": ([_U32, _U32, _U32, ctypes.POINTER(_P)], ctypes.c_int),
            "lbm_surrogate_destroy": ([_P], None),
            "lbm_surrogate_init_params": ([_P, ctypes.c_uint64], ctypes.c_int),
            "lbm_surrogate_set_params": ([_P, _P, _SZ], ctypes.c_int),
            "lbm_surrogate_get_params": ([_P, _P, _SZ], ctypes.c_int),
            "lbm_surrogate_predict": ([_P, _P, _U32, _P], ctypes.c_int),
            "lbm_surrogate_predict_bf16": ([_P, _P, _U32, _P], ctypes.c_int),
            "lbm_surrogate_loss_grad": ([_P, _P, _P, _U32, ctypes.POINTER(ctypes.c_double), _P, _SZ], ctypes.c_int),
            "lbm_surrogate_train_step": ([_P, _P, _P, _U32, ctypes.c_float, ctypes.POINTER(ctypes.c_double)],
                                         ctypes.c_int),
            "lbm_surrogate_dataset_reserve": ([_P, _SZ], ctypes.c_int),
            "lbm_surrogate_dataset_add": ([_P, _P, _P, _U32], ctypes.c_int),
            "lbm_surrogate_dataset_size": ([_P], _SZ),
            "lbm_surrogate_train": ([_P, _U32, _U32, _U32, ctypes.c_float, ctypes.c_float, ctypes.c_uint64, _P],
                                    ctypes.c_int),
            "lbm_surrogate_last_ms": ([_P, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _sig_done = True
    return lib


def param_count() -> int:
    return int(_lib().lbm_surrogate_param_count())


def _f32(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class Surrogate:
    """Network, optimiser state and device-resident training set on one GPU."""

    def __init__(self, height: int, width: int, max_batch: int = 16):
        self.lib = _lib()
        self.height, self.width, self.max_batch = height, width, max_batch
        self.n = param_count()
        self._s = ctypes.c_void_p()
        _check_hip(self.lib.lbm_surrogate_create(height, width, max_batch, ctypes.byref(self._s)),
                   "lbm_surrogate_create")

    def close(self) -> None:
        if self._s:
            self.lib.lbm_surrogate_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_params(self, seed: int) -> None:
        _check_hip(self.lib.lbm_surrogate_init_params(self._s, seed), "lbm_surrogate_init_params")

    def set_params(self, params: np.ndarray) -> None:
        p = _f32(params).ravel()
        _check_hip(self.lib.lbm_surrogate_set_params(self._s, _ptr(p), p.size), "lbm_surrogate_set_params")

    def params(self) -> np.ndarray:
        p = np.empty(self.n, dtype=np.float32)
        _check_hip(self.lib.lbm_surrogate_get_params(self._s, _ptr(p), p.size), "lbm_surrogate_get_params")
        return p

    def _check_shape(self, a: np.ndarray, channels: int) -> None:
        if a.ndim != 4 or a.shape[1:] != (channels, self.height, self.width):
            raise ValueError(f"expected (B, {channels}, {self.height}, {self.width}), got {a.shape}")

    def predict(self, inputs: np.ndarray, bf16: bool = False) -> np.ndarray:
        """Forward pass; bf16: convolutions on the matrix cores in BF16 (FP32 accumulation)."""
        x = _f32(inputs)
        self._check_shape(x, INPUTS)
        out = np.empty((x.shape[0], OUTPUTS, self.height, self.width), dtype=np.float32)
        fn = self.lib.lbm_surrogate_predict_bf16 if bf16 else self.lib.lbm_surrogate_predict
        _check_hip(fn(self._s, _ptr(x), x.shape[0], _ptr(out)), "lbm_surrogate_predict")
        return out

    def loss_grad(self, inputs: np.ndarray, targets: np.ndarray) -> tuple[float, np.ndarray]:
        x, t = _f32(inputs), _f32(targets)
        self._check_shape(x, INPUTS)
        self._check_shape(t, OUTPUTS)
        loss = ctypes.c_double()
        g = np.empty(self.n, dtype=np.float32)
        _check_hip(self.lib.lbm_surrogate_loss_grad(self._s, _ptr(x), _ptr(t), x.shape[0], ctypes.byref(loss),
                                                    _ptr(g), g.size), "lbm_surrogate_loss_grad")
        return loss.value, g

    def train_step(self, inputs: np.ndarray, targets: np.ndarray, lr: float) -> float:
        x, t = _f32(inputs), _f32(targets)
        self._check_shape(x, INPUTS)
        self._check_shape(t, OUTPUTS)
        loss = ctypes.c_double()
        _check_hip(self.lib.lbm_surrogate_train_step(self._s, _ptr(x), _ptr(t), x.shape[0], lr, ctypes.byref(loss)),
                   "lbm_surrogate_train_step")
        return loss.value

    def dataset_reserve(self, capacity: int) -> None:
        _check_hip(self.lib.lbm_surrogate_dataset_reserve(self._s, capacity), "lbm_surrogate_dataset_reserve")

    def dataset_add(self, inputs: np.ndarray, targets: np.ndarray) -> None:
        x, t = _f32(inputs), _f32(targets)
        self._check_shape(x, INPUTS)
        self._check_shape(t, OUTPUTS)
        _check_hip(self.lib.lbm_surrogate_dataset_add(self._s, _ptr(x), _ptr(t), x.shape[0]),
                   "lbm_surrogate_dataset_add")

    def dataset_size(self) -> int:
        return int(self.lib.lbm_surrogate_dataset_size(self._s))

    def train(self, steps: int, schedule_steps: int, batch: int, lr_max: float, lr_min: float,
              seed: int = 1) -> np.ndarray:
        hist = np.empty(steps, dtype=np.float32)
        _check_hip(self.lib.lbm_surrogate_train(self._s, steps, schedule_steps, batch, lr_max, lr_min, seed,
                                                _ptr(hist)), "lbm_surrogate_train")
        return hist

    def last_ms(self) -> float:
        ms = ctypes.c_double()
        _check_hip(self.lib.lbm_surrogate_last_ms(self._s, ctypes.byref(ms)), "lbm_surrogate_last_ms")
        return ms.value


# ---------------------------------------------------------------------------
# encoding
# ---------------------------------------------------------------------------
def _physical(doc: dict) -> tuple[float, float, float]:
    phy = doc["physical"]
    return phy["relaxationTime"], phy["inletDensity"], phy["outletDensity"]


def fluid_mask(doc: dict) -> np.ndarray:
    """Inner fluid mask (ny x nx, bool) of a settings document's unexpanded lattice."""
    lin = json.loads(json.dumps(doc))
    lin["algorithmic"]["algorithm"] = "lptl"
    H, V = doc["domain"]["horizontalNodes"] + 2, doc["domain"]["verticalNodes"] + 2
    return build_phase(lin).reshape(V, H)[1:-1, 1:-1] == 0


def features(mask: np.ndarray, tau: float, rho_in: float, rho_out: float) -> np.ndarray:
    ny, nx = mask.shape
    x = np.empty((INPUTS, ny, nx), dtype=np.float32)
    x[0] = mask
    x[1] = ((np.arange(nx) + 0.5) / nx * 2.0 - 1.0)[None, :]
    x[2] = (tau - TAU_MID) / TAU_HALF
    x[3] = ((rho_in - rho_out) - DRHO_MID) / DRHO_HALF
    return x


def encode_targets(rho: np.ndarray, jx: np.ndarray, jy: np.ndarray, mask: np.ndarray, tau: float, rho_in: float,
                   rho_out: float, j_scale: float, rho_scale: float = 1.0) -> np.ndarray:
    ny, nx = mask.shape
    drho = rho_in - rho_out
    t = np.zeros((OUTPUTS, ny, nx), dtype=np.float32)
    s = (tau - 0.5) / drho * j_scale
    t[0] = np.where(mask, (rho.reshape(ny, nx) - rho_out) / drho * rho_scale, 0.0)
    t[1] = np.where(mask, jx.reshape(ny, nx) * s, 0.0)
    t[2] = np.where(mask, jy.reshape(ny, nx) * s, 0.0)
    return t


def trapped(mask: np.ndarray, axis: int) -> np.ndarray:
    """Fluid nodes whose every population moving along `axis` is bounced back
    (the three upstream nodes on both sides solid).  There the momentum along
    `axis` only changes sign each step and never decays, and a start from
    rest leaves it zero.  Beyond the inlet and outlet columns nothing is
    solid, beyond the top and bottom rows everything is."""
    solid = np.pad(~mask, ((1, 1), (0, 0)), constant_values=True)
    solid = np.pad(solid, ((0, 0), (1, 1)), constant_values=False)
    ny, nx = mask.shape
    side = lambda dy, dx: solid[1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]  # noqa: E731
    t = np.ones_like(mask)
    for k in (-1, 0, 1):
        t &= (side(k, -1) & side(k, 1)) if axis == 1 else (side(-1, k) & side(1, k))
    return t & mask


def _masked_gaussian(a: np.ndarray, mask: np.ndarray, sigma: float) -> np.ndarray:
    from scipy import ndimage
    w = ndimage.gaussian_filter(mask.astype(np.float64), sigma)
    v = ndimage.gaussian_filter(np.where(mask, a, 0.0), sigma)
    return np.where(mask, v / np.maximum(w, 1e-12), a)


def darcy_profile(rho: np.ndarray, jx: np.ndarray, mask: np.ndarray, rho_in: float, rho_out: float,
                  width: float) -> np.ndarray:
    """The predicted density with its column means replaced by a profile that
    carries one flux through every column.  At the steady state each column
    passes the same mass flux Q.  With the prediction's column resistance
    r(x) = g(x) / q(x) -- g the fall of its column-mean density per column
    (central differences, the boundary densities standing beside the first
    and last columns), q its column flux, both Gaussian-smoothed over `width`
    columns -- the column means become
    rho_bar(x) = rho_in - Q (r(0) + ... + r(x-1) + r(x) / 2),
    with Q such that the sum over all columns spans rho_in - rho_out: the
    boundary densities act half a column outside the first and last columns
    (on 128 validation channels this placement gave 1.41x against 1.39x for
    a whole column).  Deviations from the column means are kept; columns
    without fluid take interpolated means and fluxes."""
    from scipy import ndimage
    ny, nx = mask.shape
    count = mask.sum(axis=0)
    if np.count_nonzero(count) < 2:
        return rho
    mean = np.where(mask, rho, 0.0).sum(axis=0) / np.maximum(count, 1)
    q = np.where(mask, jx, 0.0).sum(axis=0)
    cols, has = np.arange(nx), count > 0
    mean = np.interp(cols, cols[has], mean[has])
    q = np.interp(cols, cols[has], q[has])
    ext = np.concatenate([[rho_in], mean, [rho_out]])
    fall = -(ext[2:] - ext[:-2]) / 2.0
    tiny = 1e-12 * max(abs(rho_in - rho_out), 1e-12)
    r = (np.maximum(ndimage.gaussian_filter1d(fall, width), tiny)
         / np.maximum(ndimage.gaussian_filter1d(q, width), tiny))
    cum = np.concatenate([[0.0], np.cumsum(r)])
    profile = rho_in - (rho_in - rho_out) * (cum[:-1] + cum[1:]) / (2.0 * cum[-1])
    return np.where(mask, rho - mean[None, :] + profile[None, :], rho)


def decode(out: np.ndarray, mask: np.ndarray, tau: float, rho_in: float, rho_out: float, j_scale: float,
           rho_scale: float = 1.0, j_blend: float = 1.0, rho_sigma: float = 0.0,
           rho_profile: float = 0.0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(density, x momentum, y momentum) on the inner grid, row-major, for a
    warm start.  Solid nodes get the outlet density at rest (the warm start
    does not write them).  j_blend scales the predicted momentum (0: density
    only); rho_profile > 0 rebuilds the column-mean density from one flux
    through every column (darcy_profile, smoothing width rho_profile);
    rho_sigma > 0 then smooths the density with a masked Gaussian of that
    width; the trapped momentum of one-node gaps is zeroed."""
    drho = rho_in - rho_out
    s = drho / ((tau - 0.5) * j_scale) * j_blend
    o = out.astype(np.float64)
    rho = rho_out + drho * o[0] / rho_scale
    if rho_profile > 0:
        q = drho / ((tau - 0.5) * j_scale)
        rho = darcy_profile(rho, o[1] * q, mask, rho_in, rho_out, rho_profile)
    if rho_sigma > 0:
        rho = _masked_gaussian(rho, mask, rho_sigma)
    rho = np.where(mask, rho, rho_out)
    jx = np.where(mask, o[1] * s, 0.0)
    jy = np.where(mask, o[2] * s, 0.0)
    jx[trapped(mask, 1)] = 0.0
    jy[trapped(mask, 0)] = 0.0
    return rho.ravel(), jx.ravel(), jy.ravel()


# ---------------------------------------------------------------------------
# model files (numpy .npz, no pickled objects)
# ---------------------------------------------------------------------------
class Model:
    """Trained parameters, the target encoding (j_scale, rho_scale) and the
    warm-start options chosen on validation domains (j_blend, rho_sigma,
    rho_profile)."""

    def __init__(self, params: np.ndarray, height: int, width: int, j_scale: float, rho_scale: float = 1.0,
                 j_blend: float = 1.0, rho_sigma: float = 0.0, meta: dict | None = None, bf16: bool = True,
                 rho_profile: float = 0.0):
        self.bf16 = bf16  # warm starts infer with BF16 matrix-core convolutions
        self.params = _f32(params)
        self.height, self.width = int(height), int(width)
        self.j_scale, self.rho_scale = float(j_scale), float(rho_scale)
        self.j_blend, self.rho_sigma = float(j_blend), float(rho_sigma)
        self.rho_profile = float(rho_profile)
        self.meta = meta or {}

    def save(self, path: str | Path) -> None:
        np.savez(path, params=self.params, shape=np.array([self.height, self.width], dtype=np.int64),
                 encoding=np.array([self.j_scale, self.rho_scale, self.j_blend, self.rho_sigma, self.rho_profile]),
                 meta=np.array(json.dumps(self.meta)))

    @classmethod
    def load(cls, path: str | Path = DEFAULT_MODEL) -> "Model":
        with np.load(path, allow_pickle=False) as z:
            h, w = (int(v) for v in z["shape"])
            enc = [float(v) for v in z["encoding"]]
            js, rs, jb, sg_ = enc[:4]
            rp = enc[4] if len(enc) > 4 else 0.0  # files written before the option: none
            return cls(z["params"], h, w, js, rs, jb, sg_, json.loads(str(z["meta"])), rho_profile=rp)

    def network(self, max_batch: int = 16) -> Surrogate:
        net = Surrogate(self.height, self.width, max_batch)
        net.set_params(self.params)
        return net

    def fields(self, out: np.ndarray, doc: dict, mask: np.ndarray, j_blend: float | None = None,
               rho_sigma: float | None = None, rho_profile: float | None = None):
        return decode(out, mask, *_physical(doc), self.j_scale, self.rho_scale,
                      self.j_blend if j_blend is None else j_blend,
                      self.rho_sigma if rho_sigma is None else rho_sigma,
                      self.rho_profile if rho_profile is None else rho_profile)


# ---------------------------------------------------------------------------
# steady states from the batched solver
# ---------------------------------------------------------------------------
def porous_docs(count: int, seed: int, size: int = 256, storage: str = "f64") -> list[dict]:
    """Porous channels of size x size inner nodes: seeded circles (radii
    U[2, 8), obstacle margin 8), porosity U[0.6, 0.9], tau U[0.55, 0.9],
    inlet density U[1.01, 1.05], outlet 1.0."""
    rng = np.random.default_rng(seed)
    docs = []
    for _ in range(count):
        docs.append(settings("lptl", "stream", inner_h=size, inner_v=size, scenario="porous-circles",
                             work_group_size=256, time_steps=1, tau=float(rng.uniform(0.55, 0.9)),
                             rho_in=float(rng.uniform(1.01, 1.05)), rho_out=1.0, storage=storage,
                             seed=int(rng.integers(1, 2**62)), porosity=float(rng.uniform(0.6, 0.9)),
                             radiusMin=2.0, radiusMax=8.0, obstacleMargin=8, maxAttempts=20000))
    return docs


def solve_steady(docs: list[dict], steps: int, storage: str = "f32", final_check: int = 100):
    """Runs every domain `steps` steps in one batch; returns the macros of
    each domain and its residual over the last `final_check` steps."""
    with Batch(docs, storage=storage) as b:
        b.run(steps - final_check, "graph")
        b.residuals()
        b.run(final_check, "graph")
        r = b.residuals()
        macros = [b.macros(i) for i in range(len(docs))]
    return macros, r


def generate(net: Surrogate, docs: list[dict], steps: int, chunk: int, j_scale: float | None,
             rho_scale: float = 1.0, max_residual: float = 1e-3, log=print) -> dict:
    """Adds the steady states of `docs` (FP32 batched solves of `steps`
    steps) to the network's device dataset.  Domains that diverge or whose
    residual stays above `max_residual` are dropped.  j_scale None: taken from
    the first chunk (unit RMS of the normalised momentum)."""
    kept = dropped = 0
    t0 = time.time()
    for c0 in range(0, len(docs), chunk):
        part = docs[c0:c0 + chunk]
        macros, r = solve_steady(part, steps)
        xs, raw = [], []
        for doc, (rho, jx, jy, _), ri in zip(part, macros, r):
            if not np.isfinite(ri) or ri > max_residual or not np.all(np.isfinite(rho)):
                dropped += 1
                continue
            tau, rin, rout = _physical(doc)
            mask = fluid_mask(doc)
            xs.append(features(mask, tau, rin, rout))
            raw.append((rho, jx, jy, mask, tau, rin, rout))
        if j_scale is None and raw:
            acc = [np.concatenate([(jx * (tau - 0.5) / (rin - rout))[mask.ravel()],
                                   (jy * (tau - 0.5) / (rin - rout))[mask.ravel()]])
                   for rho, jx, jy, mask, tau, rin, rout in raw]
            j_scale = float(1.0 / np.sqrt(np.mean(np.concatenate(acc) ** 2)))
        ts = [encode_targets(*rw, j_scale, rho_scale) for rw in raw]
        if xs:
            net.dataset_add(np.stack(xs), np.stack(ts))
        kept += len(xs)
        log(f"dataset: {c0 + len(part)}/{len(docs)} solved, kept {kept}, dropped {dropped}, "
            f"{time.time() - t0:.1f} s")
    return {"kept": kept, "dropped": dropped, "j_scale": j_scale, "rho_scale": rho_scale,
            "seconds": time.time() - t0}


# ---------------------------------------------------------------------------
# warm start and steps to steady state
# ---------------------------------------------------------------------------
def predict(model: Model, docs: list[dict], net: Surrogate | None = None):
    """Network outputs for `docs`, their fluid masks and the inference time (ms)."""
    own = net is None
    net = net or model.network()
    try:
        masks = [fluid_mask(d) for d in docs]
        out = net.predict(np.stack([features(m, *_physical(d)) for m, d in zip(masks, docs)]), bf16=model.bf16)
        return out, masks, net.last_ms()
    finally:
        if own:
            net.close()


def warm_start(batch: Batch, docs: list[dict], model: Model, net: Surrogate | None = None,
               prediction=None, j_blend: float | None = None, rho_sigma: float | None = None,
               rho_profile: float | None = None) -> float:
    """Sets every domain of `batch` to the equilibrium of the predicted field;
    returns the inference time (ms, HIP events)."""
    out, masks, ms = prediction if prediction is not None else predict(model, docs, net)
    for i, (d, m) in enumerate(zip(docs, masks)):
        batch.set_equilibrium(i, *model.fields(out[i], d, m, j_blend, rho_sigma, rho_profile))
    return ms


def steps_to_residual(batch: Batch, tol: float = 1e-6, check: int = 100, max_steps: int = 60000,
                      fallback_after: int | None = None) -> dict:
    """Steps until ||j^{t+check} - j^t|| / ||j^{t+check}|| < tol per domain
    (-1 if not reached within max_steps).  With fallback_after = n, a domain
    whose residual turns NaN, or after n checks has not fallen below half its
    first value, is reset to rest and solved plainly; its count includes the
    steps spent before the reset."""
    n = len(batch.shapes)
    done = np.full(n, -1, dtype=np.int64)
    first = np.full(n, np.nan)
    since = np.zeros(n, dtype=np.int64)   # checks since the (re)start
    fell_back = np.zeros(n, dtype=bool)
    batch.residuals()
    steps = 0
    while steps < max_steps and np.any(done < 0):
        batch.run(check, "graph")
        steps += check
        r = batch.residuals()
        since += 1
        for d in np.nonzero(done < 0)[0]:
            if fallback_after is not None and not fell_back[d]:
                if since[d] == 1:
                    first[d] = r[d]
                bad = not np.isfinite(r[d]) or (since[d] >= fallback_after and not r[d] < 0.5 * first[d])
                if bad:
                    batch.reset_domain(int(d))
                    fell_back[d] = True
                    since[d] = 0
                    continue
            # after a reset the snapshot still holds the pre-reset field: skip that first check
            if np.isfinite(r[d]) and r[d] < tol and since[d] >= (2 if fell_back[d] else 1):
                done[d] = steps
    return {"steps": done, "fell_back": fell_back}


def _summary(rest: np.ndarray, warm: np.ndarray) -> dict:
    ok = (rest > 0) & (warm > 0)
    ratio = rest[ok] / warm[ok]
    return {
        "domains": int(rest.size), "both_converged": int(ok.sum()), "rest_unconverged": int(np.sum(rest < 0)),
        "warm_unconverged": int(np.sum(warm < 0)),
        "rest_median": float(np.median(rest[rest > 0])) if np.any(rest > 0) else None,
        "warm_median": float(np.median(warm[warm > 0])) if np.any(warm > 0) else None,
        "rest_mean": float(np.mean(rest[ok])) if ok.any() else None,
        "warm_mean": float(np.mean(warm[ok])) if ok.any() else None,
        "ratio_median": float(np.median(ratio)) if ok.any() else None,
        "ratio_of_means": float(np.mean(rest[ok]) / np.mean(warm[ok])) if ok.any() else None,
    }


def rest_steps(docs: list[dict], tol: float = 1e-6, max_steps: int = 60000) -> np.ndarray:
    with Batch(docs, storage="f64") as b:
        return steps_to_residual(b, tol, max_steps=max_steps)["steps"]


def warm_steps(docs: list[dict], model: Model, prediction=None, tol: float = 1e-6, max_steps: int = 60000,
               fallback_after: int | None = 20, **options) -> tuple[np.ndarray, dict]:
    with Batch(docs, storage="f64") as b:
        ms = warm_start(b, docs, model, prediction=prediction, **options)
        r = steps_to_residual(b, tol, max_steps=max_steps, fallback_after=fallback_after)
    return r["steps"], {"fallbacks": int(r["fell_back"].sum()), "inference_ms": ms}


def compare(docs: list[dict], model: Model, tol: float = 1e-6, max_steps: int = 60000, fallback_after: int = 20,
            log=print) -> dict:
    """Steps to `tol` from rest and from the model's warm start, same domains, FP64."""
    t0 = time.time()
    rest = rest_steps(docs, tol, max_steps)
    log(f"rest start: {time.time() - t0:.1f} s")
    t0 = time.time()
    warm, info = warm_steps(docs, model, None, tol, max_steps, fallback_after)
    log(f"warm start: {time.time() - t0:.1f} s (inference {info['inference_ms']:.1f} ms)")
    return {**_summary(rest, warm), **info, "rest_steps": rest.tolist(), "warm_steps": warm.tolist()}


WARM_OPTIONS = [(jb, rs, rp) for jb in (0.0, 0.5, 1.0) for rs in (0.0, 1.0, 2.0) for rp in (0.0, 8.0)]


def select_options(docs: list[dict], model: Model, max_steps: int = 40000, log=print,
                   options=WARM_OPTIONS) -> dict:
    """Warm-start options (j_blend, rho_sigma, rho_profile) with the highest
    mean-step gain over rest on validation domains; sets them on `model`."""
    rest = rest_steps(docs, max_steps=max_steps)
    pred = predict(model, docs)
    table = []
    for jb, rs, rp in options:
        warm, info = warm_steps(docs, model, pred, max_steps=max_steps, j_blend=jb, rho_sigma=rs, rho_profile=rp)
        row = {"j_blend": jb, "rho_sigma": rs, "rho_profile": rp, **_summary(rest, warm),
               "fallbacks": info["fallbacks"]}
        table.append(row)
        log(f"validation: j_blend {jb}, rho_sigma {rs}, rho_profile {rp}: ratio of means "
            f"{row['ratio_of_means']}, ratio median {row['ratio_median']}, warm unconverged {row['warm_unconverged']}")
    best = max(table, key=lambda r: (r["ratio_of_means"] or 0.0) - 0.01 * r["warm_unconverged"])
    model.j_blend, model.rho_sigma, model.rho_profile = best["j_blend"], best["rho_sigma"], best["rho_profile"]
    return {"table": table,
            "chosen": {"j_blend": model.j_blend, "rho_sigma": model.rho_sigma, "rho_profile": model.rho_profile}}


def _train_main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m portlbm_amd.surrogate train")
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--data-seed", type=int, default=20261202)
    ap.add_argument("--solve-steps", type=int, default=10000)
    ap.add_argument("--chunk", type=int, default=512)
    ap.add_argument("--rho-scale", type=float, default=4.0)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-min", type=float, default=2e-5)
    ap.add_argument("--init-seed", type=int, default=7)
    ap.add_argument("--validation", type=int, default=128)
    ap.add_argument("--validation-seed", type=int, default=5151)
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", required=True)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    size = 256
    report: dict = {"args": vars(a)}
    net = Surrogate(size, size, a.batch)
    net.init_params(a.init_seed)
    net.dataset_reserve(a.samples)
    docs = porous_docs(a.samples, a.data_seed, size, storage="f32")
    report["dataset"] = generate(net, docs, a.solve_steps, a.chunk, None, a.rho_scale, log=log)
    n = net.dataset_size()
    t0 = time.time()
    hist = []
    seg = 500
    for s0 in range(0, a.train_steps, seg):
        h = net.train(min(seg, a.train_steps - s0), a.train_steps, a.batch, a.lr, a.lr_min, seed=a.init_seed)
        hist.append(h)
        log(f"train: step {s0 + len(h)}/{a.train_steps}, loss {float(np.mean(h[-100:])):.5f}, "
            f"{time.time() - t0:.1f} s")
    hist = np.concatenate(hist)
    report["train"] = {"samples": n, "steps": a.train_steps, "batch": a.batch, "seconds": time.time() - t0,
                       "ms_per_step": (time.time() - t0) * 1e3 / a.train_steps,
                       "epochs": a.train_steps * a.batch / max(n, 1),
                       "loss_first100": float(np.mean(hist[:100])), "loss_last100": float(np.mean(hist[-100:])),
                       "loss_curve": hist[::max(1, len(hist) // 200)].tolist()}
    model = Model(net.params(), size, size, report["dataset"]["j_scale"], a.rho_scale,
                  meta={"samples": n, "train_steps": a.train_steps, "data_seed": a.data_seed,
                        "solve_steps": a.solve_steps, "validation_seed": a.validation_seed})
    net.close()
    if a.validation > 0:
        report["validation"] = select_options(porous_docs(a.validation, a.validation_seed, size), model, log=log)
    model.save(a.out)
    Path(a.report).write_text(json.dumps(report, indent=1))
    return 0


def _evaluate_main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m portlbm_amd.surrogate evaluate")
    ap.add_argument("--model", default=str(DEFAULT_MODEL))
    ap.add_argument("--domains", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=99991)
    ap.add_argument("--max-steps", type=int, default=60000)
    ap.add_argument("--report", required=True)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    model = Model.load(a.model)
    docs = porous_docs(a.domains, a.seed, model.width)
    res = compare(docs, model, max_steps=a.max_steps, log=log)
    res["options"] = {"j_blend": model.j_blend, "rho_sigma": model.rho_sigma, "rho_profile": model.rho_profile}
    log(f"eval: rest median {res['rest_median']}, warm median {res['warm_median']}, ratio median "
        f"{res['ratio_median']}, ratio of means {res['ratio_of_means']}, fallbacks {res['fallbacks']}")
    Path(a.report).write_text(json.dumps({"args": vars(a), "eval": res}, indent=1))
    return 0


def _select_main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m portlbm_amd.surrogate select")
    ap.add_argument("--model", default=str(DEFAULT_MODEL))
    ap.add_argument("--validation", type=int, default=128)
    ap.add_argument("--validation-seed", type=int, default=5151)
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", required=True)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    model = Model.load(a.model)
    res = select_options(porous_docs(a.validation, a.validation_seed, model.width), model, log=log)
    model.save(a.out)
    Path(a.report).write_text(json.dumps({"args": vars(a), "validation": res}, indent=1))
    return 0


if __name__ == "__main__":
    commands = {"train": _train_main, "evaluate": _evaluate_main, "select": _select_main}
    if len(sys.argv) > 1 and sys.argv[1] in commands:
        sys.exit(commands[sys.argv[1]](sys.argv[2:]))
    print(__doc__)
    sys.exit(2)
