"""ctypes bindings of the MI355X D2Q9 BGK library.

Two native libraries, both built in place by ``portlbm_amd.build``:

* ``lib/libportlbm_hip.so`` -- the kernel-level C ABI (include/lbm_hip.h):
  :class:`Lattice` drives one lattice on the GPU.
* ``lib/libportlbm.so`` -- the handler-level C ABI (include/portlbm_c.h):
  :class:`Handler` mirrors ``lbm::create_handler`` / ``AlgorithmHandler``;
  :func:`domain_extents`, :func:`build_phase`, :func:`validate` and
  :func:`roundtrip` run the set-up logic without a GPU.

Nothing here falls back to another implementation: a missing library raises
:class:`LibraryNotBuilt`.
"""
from __future__ import annotations

import ctypes
import json
from pathlib import Path
from typing import Any

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_DIR = _PKG / "lib"

ALGORITHMS = {"nptl": 0, "lptl": 1, "npol": 2, "nsol": 3}
LAYOUTS = {"collision": 0, "stream": 1, "bundle": 2}
STORAGES = {"f64": 0, "f32": 1, "f16-shifted": 2, "bf16-shifted": 3}
ROUNDINGS = {"nearest": 0, "stochastic": 1}
TILE_KERNELS = {"auto": 0, "sweep": 1, "block": 2, "staged": 3}
STORAGE_BYTES = {"f64": 8, "f32": 4, "f16-shifted": 2, "bf16-shifted": 2}

CODE_BOUNCE_MASK = 0x01FF
CODE_FLUID = 0x0200
CODE_OUTLET = 0x0400
CODE_UY_SOURCE = 0x0800
CODE_SOLID = 0x1000


class LibraryNotBuilt(RuntimeError):
    pass


class LbmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class LbmDesc(ctypes.Structure):
    _fields_ = [
        ("algorithm", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("nan_protection", ctypes.c_int32),
        ("horizontal_nodes", ctypes.c_uint32),
        ("vertical_nodes", ctypes.c_uint32),
        ("expanded_horizontal_nodes", ctypes.c_uint32),
        ("expanded_vertical_nodes", ctypes.c_uint32),
        ("subdomain_horizontal_nodes", ctypes.c_uint32),
        ("subdomain_vertical_nodes", ctypes.c_uint32),
        ("subdomain_count_horizontal", ctypes.c_uint32),
        ("subdomain_count_vertical", ctypes.c_uint32),
        ("work_group_size", ctypes.c_uint32),
        ("rounding", ctypes.c_uint32),
        ("tile_kernel", ctypes.c_uint32),
        ("relaxation_time", ctypes.c_double),
        ("inlet_density", ctypes.c_double),
        ("inlet_velocity_x", ctypes.c_double),
        ("inlet_velocity_y", ctypes.c_double),
        ("outlet_density", ctypes.c_double),
        ("outlet_velocity_x", ctypes.c_double),
        ("outlet_velocity_y", ctypes.c_double),
    ]


_hip: ctypes.CDLL | None = None
_host: ctypes.CDLL | None = None

_P = ctypes.c_void_p
# lbm_pass_callback of include/lbm_hip.h
PASS_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64)
PASSES = ("stream", "outlet", "macroscopic", "collide")
_U32 = ctypes.c_uint32
_SZ = ctypes.c_size_t
_DP = ctypes.POINTER(ctypes.c_double)


def _load(name: str) -> ctypes.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise LibraryNotBuilt(f"{path} is missing: run `python -m portlbm_amd.build` first")
    return ctypes.CDLL(str(path), mode=ctypes.RTLD_GLOBAL)


def hip_library() -> ctypes.CDLL:
    """The kernel-level library (include/lbm_hip.h)."""
    global _hip
    if _hip is None:
        lib = _load("libportlbm_hip.so")
        sig = {
            "lbm_last_error": ([], ctypes.c_char_p),
            "lbm_device_max_work_group_size": ([ctypes.POINTER(_U32)], ctypes.c_int),
            "lbm_create": ([ctypes.POINTER(LbmDesc), ctypes.POINTER(_P)], ctypes.c_int),
            "lbm_destroy": ([_P], None),
            "lbm_set_phase": ([_P, _P, _SZ], ctypes.c_int),
            "lbm_get_phase": ([_P, _P, _SZ], ctypes.c_int),
            "lbm_init_populations": ([_P], ctypes.c_int),
            "lbm_get_node_codes": ([_P, _P, _SZ], ctypes.c_int),
            "lbm_run": ([_P, _U32, _U32], ctypes.c_int),
            "lbm_run_split": ([_P, _U32, PASS_CALLBACK, _P], ctypes.c_int),
            "lbm_get_pass_populations": ([_P, _P, _SZ], ctypes.c_int),
            "lbm_get_iteration": ([_P, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
            "lbm_last_run_ms": ([_P, _DP], ctypes.c_int),
            "lbm_get_macros": ([_P, _P, _P, _P, _P, _SZ], ctypes.c_int),
            "lbm_run_frames": ([_P, _U32, _U32], ctypes.c_int),
            "lbm_get_frame": ([_P, _P, _P, _P, _P, _SZ, _P], ctypes.c_int),
            "lbm_get_populations": ([_P, _P, _SZ], ctypes.c_int),
            "lbm_set_populations": ([_P, _P, _SZ], ctypes.c_int),
            "lbm_set_equilibrium_field": ([_P, _P, _P, _P, _SZ], ctypes.c_int),
            "lbm_reduce": ([_P, _P], ctypes.c_int),
            "lbm_residual": ([_P, _DP], ctypes.c_int),
            "lbm_sync": ([_P], ctypes.c_int),
            "lbm_state_size": ([_P, ctypes.POINTER(_SZ)], ctypes.c_int),
            "lbm_get_state": ([_P, _P, _SZ], ctypes.c_int),
            "lbm_set_state": ([_P, _P, _SZ], ctypes.c_int),
            "lbm_measure_copy_bandwidth": ([_SZ, ctypes.c_int, _DP], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _hip = lib
    return _hip


def host_library() -> ctypes.CDLL:
    """The handler-level library (include/portlbm_c.h)."""
    global _host
    if _host is None:
        hip_library()
        lib = _load("libportlbm.so")
        H = _P
        sig = {
            "portlbm_last_error": ([], ctypes.c_char_p),
            "portlbm_create_from_file": ([ctypes.c_char_p, ctypes.POINTER(H)], ctypes.c_int),
            "portlbm_create_from_json": ([ctypes.c_char_p, ctypes.POINTER(H)], ctypes.c_int),
            "portlbm_destroy": ([H], None),
            "portlbm_initialize": ([H], ctypes.c_int),
            "portlbm_start": ([H], ctypes.c_int),
            "portlbm_pause": ([H], ctypes.c_int),
            "portlbm_block_until_finished": ([H], ctypes.c_int),
            "portlbm_is_finished": ([H, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
            "portlbm_get_field": ([H, ctypes.c_int, _P, _SZ], ctypes.c_int),
            "portlbm_get_extents": ([H, _P], ctypes.c_int),
            "portlbm_get_iteration": ([H, ctypes.POINTER(_U32)], ctypes.c_int),
            "portlbm_get_progress": ([H, _DP], ctypes.c_int),
            "portlbm_get_last_frametime": ([H, _DP], ctypes.c_int),
            "portlbm_get_phase": ([H, _P, _SZ], ctypes.c_int),
            "portlbm_get_populations": ([H, _P, _SZ], ctypes.c_int),
            "portlbm_validate_json": ([ctypes.c_char_p], ctypes.c_int),
            "portlbm_domain_extents": ([ctypes.c_char_p, _P], ctypes.c_int),
            "portlbm_build_phase": ([ctypes.c_char_p, _P, _SZ], ctypes.c_int),
            "portlbm_roundtrip_json": ([ctypes.c_char_p, _P, _SZ, ctypes.POINTER(_SZ)], ctypes.c_int),
            "portlbm_properties_string": ([ctypes.c_char_p, _P, _SZ, ctypes.POINTER(_SZ)], ctypes.c_int),
            "portlbm_save_snapshot": ([_P, ctypes.c_char_p], ctypes.c_int),
            "portlbm_load_snapshot": ([_P, ctypes.c_char_p], ctypes.c_int),
            "portlbm_export_csv": ([_P, ctypes.c_char_p], ctypes.c_int),
            "portlbm_debug_history": ([_P, ctypes.c_int, _P, _SZ, ctypes.POINTER(_SZ)], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _host = lib
    return _host


def _check_hip(rc: int, what: str) -> None:
    if rc != 0:
        raise LbmError(rc, f"{what}: {hip_library().lbm_last_error().decode()}")


def _check_host(rc: int, what: str) -> None:
    if rc != 0:
        raise LbmError(rc, f"{what}: {host_library().portlbm_last_error().decode()}")


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------------------
# settings documents (schema of the reference's settings.json)
# ---------------------------------------------------------------------------
def settings(algorithm: str = "nptl", data_layout: str = "stream", *, inner_h: int = 256, inner_v: int = 64,
             scenario: str = "Hagen-Poiseuille", work_group_size: int = 256, time_steps: int = 1000,
             frame_update_interval: int | None = None, debug_mode: bool = False, tau: float = 0.6,
             rho_in: float = 1.2, rho_out: float = 1.0, u_in: tuple[float, float] = (0.0, 0.0),
             u_out: tuple[float, float] = (0.0, 0.0), **extensions: Any) -> dict:
    """A settings document; keyword extensions go to the "mi355x" section
    (storage, rounding, tileKernel, nanProtection, seed, porosity, radiusMin, radiusMax,
    obstacleMargin, cylinderX, cylinderRadius)."""
    doc = {
        "algorithmic": {
            "algorithm": algorithm,
            "dataLayout": data_layout,
            "debugMode": debug_mode,
            "frameUpdateInterval": frame_update_interval if frame_update_interval is not None else time_steps + 1,
            "timeSteps": time_steps,
            "workGroupSize": work_group_size,
        },
        "domain": {"horizontalNodes": inner_h, "verticalNodes": inner_v, "scenario": scenario},
        "physical": {
            "inletDensity": rho_in,
            "inletVelocity": {"x": u_in[0], "y": u_in[1]},
            "outletDensity": rho_out,
            "outletVelocity": {"x": u_out[0], "y": u_out[1]},
            "relaxationTime": tau,
        },
    }
    if extensions:
        doc["mi355x"] = dict(extensions)
    return doc


def _text(doc: dict | str) -> bytes:
    return (doc if isinstance(doc, str) else json.dumps(doc)).encode()


def validate(doc: dict | str) -> None:
    _check_host(host_library().portlbm_validate_json(_text(doc)), "validate")


def domain_extents(doc: dict | str) -> dict:
    """H, V (with ghosts), Hx, Vx, sub_h, sub_v, cnt_h, cnt_v, N of the settings."""
    out = (ctypes.c_uint32 * 9)()
    _check_host(host_library().portlbm_domain_extents(_text(doc), out), "domain_extents")
    keys = ("H", "V", "Hx", "Vx", "sub_h", "sub_v", "cnt_h", "cnt_v", "N")
    return dict(zip(keys, list(out)))


def build_phase(doc: dict | str) -> np.ndarray:
    """The node-type table the handler hands to the GPU (expanded lattice)."""
    n = domain_extents(doc)["N"]
    out = np.empty(n, dtype=np.int8)
    _check_host(host_library().portlbm_build_phase(_text(doc), _ptr(out), n), "build_phase")
    return out


def _string_call(fn, doc) -> str:
    length = ctypes.c_size_t(0)
    _check_host(fn(_text(doc), None, 0, ctypes.byref(length)), fn.__name__)
    buf = ctypes.create_string_buffer(length.value + 1)
    _check_host(fn(_text(doc), ctypes.cast(buf, ctypes.c_void_p), length.value + 1, ctypes.byref(length)),
                fn.__name__)
    return buf.value.decode()


def roundtrip(doc: dict | str) -> dict:
    """settings -> Properties -> settings."""
    return json.loads(_string_call(host_library().portlbm_roundtrip_json, doc))


def properties_string(doc: dict | str) -> str:
    return _string_call(host_library().portlbm_properties_string, doc)


def desc_from_settings(doc: dict) -> LbmDesc:
    """The C-ABI lattice description of a settings document."""
    e = domain_extents(doc)
    alg, phy = doc["algorithmic"], doc["physical"]
    ext = doc.get("mi355x", {})
    d = LbmDesc()
    d.algorithm = ALGORITHMS[alg["algorithm"]]
    d.layout = LAYOUTS[alg["dataLayout"]]
    d.storage = STORAGES[ext.get("storage", "f64")]
    d.nan_protection = 1 if ext.get("nanProtection", True) else 0
    d.rounding = ROUNDINGS[ext.get("rounding", "nearest")]
    d.tile_kernel = TILE_KERNELS[ext.get("tileKernel", "auto")]
    d.horizontal_nodes, d.vertical_nodes = e["H"], e["V"]
    d.expanded_horizontal_nodes, d.expanded_vertical_nodes = e["Hx"], e["Vx"]
    d.subdomain_horizontal_nodes, d.subdomain_vertical_nodes = e["sub_h"], e["sub_v"]
    d.subdomain_count_horizontal, d.subdomain_count_vertical = e["cnt_h"], e["cnt_v"]
    d.work_group_size = alg["workGroupSize"]
    d.relaxation_time = phy["relaxationTime"]
    d.inlet_density = phy["inletDensity"]
    d.inlet_velocity_x, d.inlet_velocity_y = phy["inletVelocity"]["x"], phy["inletVelocity"]["y"]
    d.outlet_density = phy["outletDensity"]
    d.outlet_velocity_x, d.outlet_velocity_y = phy["outletVelocity"]["x"], phy["outletVelocity"]["y"]
    return d


def device_max_work_group_size() -> int:
    out = ctypes.c_uint32(0)
    _check_hip(hip_library().lbm_device_max_work_group_size(ctypes.byref(out)), "device query")
    return out.value


def copy_bandwidth_gbps(nbytes: int = 1 << 30, repeats: int = 5) -> float:
    out = ctypes.c_double(0)
    _check_hip(hip_library().lbm_measure_copy_bandwidth(nbytes, repeats, ctypes.byref(out)), "copy bandwidth")
    return out.value


# ---------------------------------------------------------------------------
# kernel-level lattice
# ---------------------------------------------------------------------------
class Lattice:
    """One lattice on the GPU through include/lbm_hip.h."""

    def __init__(self, desc: LbmDesc, phase: np.ndarray | None = None):
        self.lib = hip_library()
        self.desc = desc
        self._ctx = ctypes.c_void_p()
        _check_hip(self.lib.lbm_create(ctypes.byref(desc), ctypes.byref(self._ctx)), "lbm_create")
        self.H, self.V = desc.horizontal_nodes, desc.vertical_nodes
        self.N = desc.expanded_horizontal_nodes * desc.expanded_vertical_nodes
        self.D = (self.H - 2) * (self.V - 2)
        if phase is not None:
            self.set_phase(phase)
            self.init()

    @classmethod
    def from_settings(cls, doc: dict) -> "Lattice":
        return cls(desc_from_settings(doc), build_phase(doc))

    def close(self) -> None:
        if self._ctx:
            self.lib.lbm_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_phase(self, phase: np.ndarray) -> None:
        phase = np.ascontiguousarray(phase, dtype=np.int8)
        _check_hip(self.lib.lbm_set_phase(self._ctx, _ptr(phase), phase.size), "lbm_set_phase")

    def phase(self) -> np.ndarray:
        out = np.empty(self.N, dtype=np.int8)
        _check_hip(self.lib.lbm_get_phase(self._ctx, _ptr(out), self.N), "lbm_get_phase")
        return out

    def init(self) -> None:
        _check_hip(self.lib.lbm_init_populations(self._ctx), "lbm_init_populations")

    def node_codes(self) -> np.ndarray:
        out = np.empty(self.N, dtype=np.uint16)
        _check_hip(self.lib.lbm_get_node_codes(self._ctx, _ptr(out), self.N), "lbm_get_node_codes")
        return out

    def run(self, steps: int, frame_interval: int = 0) -> float:
        """Advance; returns the device milliseconds of the call."""
        _check_hip(self.lib.lbm_run(self._ctx, steps, frame_interval), "lbm_run")
        ms = ctypes.c_double(0)
        _check_hip(self.lib.lbm_last_run_ms(self._ctx, ctypes.byref(ms)), "lbm_last_run_ms")
        return ms.value

    def iteration(self) -> int:
        it = ctypes.c_uint64(0)
        _check_hip(self.lib.lbm_get_iteration(self._ctx, ctypes.byref(it)), "lbm_get_iteration")
        return it.value

    def macros(self) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        out = [np.empty(self.D) for _ in range(4)]
        _check_hip(self.lib.lbm_get_macros(self._ctx, *map(_ptr, out), self.D), "lbm_get_macros")
        return tuple(out)  # type: ignore[return-value]

    def run_frames(self, steps: int, frame_interval: int) -> float:
        """lbm_run with the frame copies made on a second stream while the
        steps go on; returns the device milliseconds of the call."""
        _check_hip(self.lib.lbm_run_frames(self._ctx, steps, frame_interval), "lbm_run_frames")
        ms = ctypes.c_double(0)
        _check_hip(self.lib.lbm_last_run_ms(self._ctx, ctypes.byref(ms)), "lbm_last_run_ms")
        return ms.value

    def frame(self) -> tuple[int, tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]]:
        """(iteration, (rho, jx, jy, |j|)) of the latest frame copied by run_frames."""
        out = [np.empty(self.D) for _ in range(4)]
        it = ctypes.c_uint64(0)
        _check_hip(self.lib.lbm_get_frame(self._ctx, *map(_ptr, out), self.D, ctypes.byref(it)), "lbm_get_frame")
        return it.value, tuple(out)  # type: ignore[return-value]

    def populations(self) -> np.ndarray:
        out = np.empty(9 * self.N)
        _check_hip(self.lib.lbm_get_populations(self._ctx, _ptr(out), out.size), "lbm_get_populations")
        return out

    def set_populations(self, values: np.ndarray) -> None:
        values = np.ascontiguousarray(values, dtype=np.float64)
        _check_hip(self.lib.lbm_set_populations(self._ctx, _ptr(values), values.size), "lbm_set_populations")

    def state(self) -> np.ndarray:
        """The whole lattice state (lbm_get_state) as bytes, for set_state."""
        n = _SZ(0)
        _check_hip(self.lib.lbm_state_size(self._ctx, ctypes.byref(n)), "lbm_state_size")
        out = np.empty(n.value, dtype=np.uint8)
        _check_hip(self.lib.lbm_get_state(self._ctx, _ptr(out), out.size), "lbm_get_state")
        return out

    def set_state(self, state: np.ndarray) -> None:
        state = np.ascontiguousarray(state, dtype=np.uint8)
        _check_hip(self.lib.lbm_set_state(self._ctx, _ptr(state), state.size), "lbm_set_state")

    def run_split(self, steps: int, callback=None) -> float:
        """Split (debug) steps; callback(pass, iteration) after every pass."""
        cb = PASS_CALLBACK((lambda _u, p, it: callback(p, it)) if callback else (lambda _u, p, it: None))
        _check_hip(self.lib.lbm_run_split(self._ctx, steps, cb, None), "lbm_run_split")
        ms = ctypes.c_double(0)
        _check_hip(self.lib.lbm_last_run_ms(self._ctx, ctypes.byref(ms)), "lbm_last_run_ms")
        return ms.value

    def pass_populations(self) -> np.ndarray:
        out = np.empty(9 * self.N)
        _check_hip(self.lib.lbm_get_pass_populations(self._ctx, _ptr(out), out.size), "lbm_get_pass_populations")
        return out.reshape(9, self.N)

    def set_equilibrium_field(self, rho: np.ndarray, jx: np.ndarray, jy: np.ndarray) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (rho, jx, jy)]
        _check_hip(self.lib.lbm_set_equilibrium_field(self._ctx, *map(_ptr, arrs), self.D),
                   "lbm_set_equilibrium_field")

    def reduce(self) -> dict:
        out = np.zeros(6)
        _check_hip(self.lib.lbm_reduce(self._ctx, _ptr(out)), "lbm_reduce")
        return {"mass": out[0], "jx": out[1], "jy": out[2], "fluid": int(out[3]), "fx": out[4], "fy": out[5]}

    def residual(self) -> float:
        r = ctypes.c_double(0)
        _check_hip(self.lib.lbm_residual(self._ctx, ctypes.byref(r)), "lbm_residual")
        return r.value

    def sync(self) -> None:
        _check_hip(self.lib.lbm_sync(self._ctx), "lbm_sync")


# ---------------------------------------------------------------------------
# handler level
# ---------------------------------------------------------------------------
class Handler:
    """lbm::create_handler(...) through include/portlbm_c.h."""

    FIELDS = {"density": 0, "x_velocity": 1, "y_velocity": 2, "absolute_velocity": 3}

    def __init__(self, doc: dict | str | None = None, *, settings_path: str | Path | None = None):
        self.lib = host_library()
        self._h = ctypes.c_void_p()
        if settings_path is not None:
            rc = self.lib.portlbm_create_from_file(str(settings_path).encode(), ctypes.byref(self._h))
        else:
            rc = self.lib.portlbm_create_from_json(_text(doc), ctypes.byref(self._h))
        _check_host(rc, "create_handler")
        self.extents = self._extents()

    def _extents(self) -> dict:
        out = (ctypes.c_uint32 * 9)()
        _check_host(self.lib.portlbm_get_extents(self._h, out), "get_extents")
        keys = ("H", "V", "Hx", "Vx", "sub_h", "sub_v", "cnt_h", "cnt_v", "N")
        return dict(zip(keys, list(out)))

    def close(self) -> None:
        if self._h:
            self.lib.portlbm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self) -> None:
        _check_host(self.lib.portlbm_initialize(self._h), "initialize")

    def start(self) -> None:
        _check_host(self.lib.portlbm_start(self._h), "start")

    def pause(self) -> None:
        _check_host(self.lib.portlbm_pause(self._h), "pause")

    def block_until_finished(self) -> None:
        _check_host(self.lib.portlbm_block_until_finished(self._h), "block_until_finished")

    def is_finished(self) -> bool:
        f = ctypes.c_int(0)
        _check_host(self.lib.portlbm_is_finished(self._h, ctypes.byref(f)), "is_finished")
        return bool(f.value)

    def field(self, name: str) -> np.ndarray:
        n = (self.extents["H"] - 2) * (self.extents["V"] - 2)
        out = np.empty(n)
        _check_host(self.lib.portlbm_get_field(self._h, self.FIELDS[name], _ptr(out), n), "get_field")
        return out

    def get_densities(self) -> np.ndarray:
        return self.field("density")

    def get_x_velocities(self) -> np.ndarray:
        return self.field("x_velocity")

    def get_y_velocities(self) -> np.ndarray:
        return self.field("y_velocity")

    def get_absolute_velocities(self) -> np.ndarray:
        return self.field("absolute_velocity")

    def get_horizontal_nodes(self) -> int:
        return self.extents["H"]

    def get_vertical_nodes(self) -> int:
        return self.extents["V"]

    def get_current_iteration(self) -> int:
        it = ctypes.c_uint32(0)
        _check_host(self.lib.portlbm_get_iteration(self._h, ctypes.byref(it)), "get_current_iteration")
        return it.value

    def get_progress(self) -> float:
        v = ctypes.c_double(0)
        _check_host(self.lib.portlbm_get_progress(self._h, ctypes.byref(v)), "get_progress")
        return v.value

    def get_last_frametime(self) -> float:
        v = ctypes.c_double(0)
        _check_host(self.lib.portlbm_get_last_frametime(self._h, ctypes.byref(v)), "get_last_frametime")
        return v.value

    def phase(self) -> np.ndarray:
        out = np.empty(self.extents["N"], dtype=np.int8)
        _check_host(self.lib.portlbm_get_phase(self._h, _ptr(out), out.size), "get_phase")
        return out

    def save_snapshot(self, path: str | Path) -> None:
        _check_host(self.lib.portlbm_save_snapshot(self._h, str(path).encode()), "portlbm_save_snapshot")

    def load_snapshot(self, path: str | Path) -> None:
        _check_host(self.lib.portlbm_load_snapshot(self._h, str(path).encode()), "portlbm_load_snapshot")

    def debug_history(self, name: str) -> np.ndarray:
        """Debug mode: the field of every finished iteration, shape (steps, (H-2)(V-2))."""
        n = _SZ(0)
        _check_host(self.lib.portlbm_debug_history(self._h, self.FIELDS[name], None, 0, ctypes.byref(n)),
                    "debug_history")
        out = np.empty(n.value)
        _check_host(self.lib.portlbm_debug_history(self._h, self.FIELDS[name], _ptr(out), out.size,
                                                   ctypes.byref(n)), "debug_history")
        return out

    def export_csv(self, path: str | Path) -> None:
        _check_host(self.lib.portlbm_export_csv(self._h, str(path).encode()), "portlbm_export_csv")

    def populations(self) -> np.ndarray:
        out = np.empty(9 * self.extents["N"])
        _check_host(self.lib.portlbm_get_populations(self._h, _ptr(out), out.size), "get_populations")
        return out


# ---------------------------------------------------------------------------
# batched domains
# ---------------------------------------------------------------------------
class BatchDomain(ctypes.Structure):
    _fields_ = [
        ("horizontal_nodes", ctypes.c_uint32),
        ("vertical_nodes", ctypes.c_uint32),
        ("relaxation_time", ctypes.c_double),
        ("inlet_density", ctypes.c_double),
        ("inlet_velocity_x", ctypes.c_double),
        ("inlet_velocity_y", ctypes.c_double),
        ("outlet_density", ctypes.c_double),
        ("outlet_velocity_x", ctypes.c_double),
        ("outlet_velocity_y", ctypes.c_double),
    ]


BATCH_MODES = {"per-domain": 0, "persistent": 1, "graph": 2}
_batch_sig_done = False


def _batch_sigs() -> ctypes.CDLL:
    global _batch_sig_done
    lib = hip_library()
    if not _batch_sig_done:
        sig = {
            "lbm_batch_create": ([ctypes.POINTER(BatchDomain), _U32, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.POINTER(_P)], ctypes.c_int),
            "lbm_batch_destroy": ([_P], None),
            "lbm_batch_set_phase": ([_P, _U32, _P, _SZ], ctypes.c_int),
            "lbm_batch_init": ([_P], ctypes.c_int),
            "lbm_batch_run": ([_P, _U32, ctypes.c_int32], ctypes.c_int),
            "lbm_batch_last_run_ms": ([_P, _DP], ctypes.c_int),
            "lbm_batch_get_macros": ([_P, _U32, _P, _P, _P, _P, _SZ], ctypes.c_int),
            "lbm_batch_fluid_nodes": ([_P, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
            "lbm_batch_set_tiles": ([_P, _P], ctypes.c_int),
            "lbm_batch_get_tiles": ([_P, _P], ctypes.c_int),
            "lbm_batch_autotune": ([_P, _U32, _DP], ctypes.c_int),
            "lbm_batch_residuals": ([_P, _P, _U32], ctypes.c_int),
            "lbm_batch_residuals_noalloc": ([_P, _P, _U32], ctypes.c_int),
            "lbm_batch_reset_domain": ([_P, _U32], ctypes.c_int),
            "lbm_batch_set_equilibrium": ([_P, _U32, _P, _P, _P, _SZ], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _batch_sig_done = True
    return lib


class Batch:
    """Many independent channels advanced together (include/lbm_hip.h,
    lbm_batch_*).  `docs` are settings documents; each domain uses the
    unexpanded lattice of its document."""

    def __init__(self, docs: list[dict], storage: str = "f32", nan_protection: bool = True):
        self.lib = _batch_sigs()
        n = len(docs)
        arr = (BatchDomain * n)()
        self.shapes = []
        for i, doc in enumerate(docs):
            dom, phy = doc["domain"], doc["physical"]
            H, V = dom["horizontalNodes"] + 2, dom["verticalNodes"] + 2
            arr[i] = BatchDomain(H, V, phy["relaxationTime"], phy["inletDensity"], phy["inletVelocity"]["x"],
                                 phy["inletVelocity"]["y"], phy["outletDensity"], phy["outletVelocity"]["x"],
                                 phy["outletVelocity"]["y"])
            self.shapes.append((H, V))
        self._b = ctypes.c_void_p()
        _check_hip(self.lib.lbm_batch_create(arr, n, STORAGES[storage], 1 if nan_protection else 0,
                                             ctypes.byref(self._b)), "lbm_batch_create")
        for i, doc in enumerate(docs):
            lin = json.loads(json.dumps(doc))
            lin["algorithmic"]["algorithm"] = "lptl"  # the unexpanded lattice
            phase = build_phase(lin)
            _check_hip(self.lib.lbm_batch_set_phase(self._b, i, _ptr(phase), phase.size), "lbm_batch_set_phase")
        _check_hip(self.lib.lbm_batch_init(self._b), "lbm_batch_init")

    def close(self) -> None:
        if self._b:
            self.lib.lbm_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self) -> None:
        _check_hip(self.lib.lbm_batch_init(self._b), "lbm_batch_init")

    def run(self, steps: int, mode: str = "graph") -> float:
        _check_hip(self.lib.lbm_batch_run(self._b, steps, BATCH_MODES[mode]), "lbm_batch_run")
        ms = ctypes.c_double(0)
        _check_hip(self.lib.lbm_batch_last_run_ms(self._b, ctypes.byref(ms)), "lbm_batch_last_run_ms")
        return ms.value

    def macros(self, domain: int):
        H, V = self.shapes[domain]
        D = (H - 2) * (V - 2)
        out = [np.empty(D) for _ in range(4)]
        _check_hip(self.lib.lbm_batch_get_macros(self._b, domain, *map(_ptr, out), D), "lbm_batch_get_macros")
        return tuple(out)

    def fluid_nodes(self) -> int:
        v = ctypes.c_uint64(0)
        _check_hip(self.lib.lbm_batch_fluid_nodes(self._b, ctypes.byref(v)), "lbm_batch_fluid_nodes")
        return v.value

    def tiles(self) -> list[int]:
        out = (ctypes.c_uint32 * 4)()
        _check_hip(self.lib.lbm_batch_get_tiles(self._b, out), "lbm_batch_get_tiles")
        return list(out)

    def set_tiles(self, widths: list[int]) -> None:
        arr = (ctypes.c_uint32 * 4)(*widths)
        _check_hip(self.lib.lbm_batch_set_tiles(self._b, arr), "lbm_batch_set_tiles")

    def residuals(self) -> np.ndarray:
        out = np.empty(len(self.shapes))
        _check_hip(self.lib.lbm_batch_residuals(self._b, _ptr(out), out.size), "lbm_batch_residuals")
        return out

    def residuals_noalloc(self, out: np.ndarray | None = None) -> np.ndarray:
        """residuals() through lbm_batch_residuals_noalloc (the library
        allocates nothing; `out` may be a caller-owned float64 array)."""
        if out is None:
            out = np.empty(len(self.shapes))
        _check_hip(self.lib.lbm_batch_residuals_noalloc(self._b, _ptr(out), out.size), "lbm_batch_residuals_noalloc")
        return out

    def reset_domain(self, domain: int) -> None:
        _check_hip(self.lib.lbm_batch_reset_domain(self._b, domain), "lbm_batch_reset_domain")

    def set_equilibrium(self, domain: int, rho: np.ndarray, jx: np.ndarray, jy: np.ndarray) -> None:
        H, V = self.shapes[domain]
        D = (H - 2) * (V - 2)
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (rho, jx, jy)]
        _check_hip(self.lib.lbm_batch_set_equilibrium(self._b, domain, *map(_ptr, arrs), D),
                   "lbm_batch_set_equilibrium")

    def autotune(self, steps_per_trial: int = 20) -> float:
        best = ctypes.c_double(0)
        _check_hip(self.lib.lbm_batch_autotune(self._b, steps_per_trial, ctypes.byref(best)), "lbm_batch_autotune")
        return best.value
