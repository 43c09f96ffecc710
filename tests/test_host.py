"""Host-side library checks (no GPU): the C++ surface, the settings file, the
scenario tables and the C ABIs' symbol tables."""
import ctypes
import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import portlbm_amd as pa
from oracle import binding as ob

ROOT = Path(__file__).resolve().parents[1]


def test_cpp_unit_tests_pass():
    """tests/cpp/unit_tests.cpp re-expresses the reference's Catch2 suites
    test_access.cpp, test_domain.cpp and test_properties.cpp."""
    exe = ROOT / "tests" / "cpp" / "portlbm_unit_tests"
    assert exe.exists(), "build first: python -m portlbm_amd.build"
    out = subprocess.run([str(exe), str(ROOT / "settings" / "settings.json")], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " 0 failures" in out.stdout


def _declared_functions(header: Path) -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(lbm_[a-z_0-9]+|portlbm_[a-z_0-9]+)\s*\(", text, flags=re.M)


@pytest.mark.parametrize("header,lib", [("lbm_hip.h", "libportlbm_hip.so"), ("lbm_surrogate.h", "libportlbm_hip.so"),
                                        ("portlbm_c.h", "libportlbm.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = _declared_functions(ROOT / "include" / header)
    assert len(names) >= 10
    so = ctypes.CDLL(str(pa.LIB_DIR / lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


@pytest.mark.parametrize("change,message", [
    ({"horizontal_nodes": 3}, "must be >= 4"),
    ({"relaxation_time": 0.0}, "relaxation_time must be > 0"),
    ({"work_group_size": 2048}, r"work_group_size must be in \[1, 1024\]"),
    ({"storage": 7}, "unknown storage type"),
    ({"rounding": 2}, "unknown rounding mode"),
    ({"tile_kernel": 4}, "unknown tile kernel"),
    ({"expanded_horizontal_nodes": 10}, "expanded extents smaller than the lattice"),
    # 9 x 40000 x 20000 populations do not fit 32-bit indices
    ({"horizontal_nodes": 40000, "vertical_nodes": 20000, "expanded_horizontal_nodes": 40000,
      "expanded_vertical_nodes": 20000}, "too large for 32-bit population indices"),
])
def test_lbm_create_rejects_bad_descriptors_before_any_device_call(change, message):
    """lbm_create validates the descriptor first (no HIP call), so the checks
    run here without a GPU; the error surfaces as LbmError with the message."""
    desc = pa.desc_from_settings(pa.settings("nptl", "stream", inner_h=64, inner_v=32))
    for key, value in change.items():
        setattr(desc, key, value)
    with pytest.raises(pa.LbmError, match=message) as err:
        pa.Lattice(desc)
    assert err.value.code == 1  # LBM_ERR_INVALID_ARGUMENT


def test_bundled_settings_file():
    doc = json.loads((ROOT / "settings" / "settings.json").read_text())
    pa.validate(doc)
    e = pa.domain_extents(doc)
    assert (e["H"], e["V"]) == (514, 130)  # 512 x 128 inner + ghosts
    assert doc["algorithmic"]["algorithm"] == "lptl" and doc["algorithmic"]["timeSteps"] == 1000000


def test_roundtrip_keeps_the_reference_schema():
    doc = pa.settings("npol", "bundle", inner_h=100, inner_v=40, time_steps=77, frame_update_interval=7, tau=0.75)
    back = pa.roundtrip(doc)
    assert "mi355x" not in back  # default extensions are not written
    assert back["algorithmic"] == {"workGroupSize": 256, "debugMode": False, "timeSteps": 77, "dataLayout": "bundle",
                                   "algorithm": "npol", "frameUpdateInterval": 7}
    # node counts are written as stored, ghosts included (file_interaction.cpp:23-24)
    assert back["domain"]["horizontalNodes"] == 102 and back["domain"]["verticalNodes"] == 42
    assert back["physical"]["relaxationTime"] == 0.75
    ext = pa.roundtrip(pa.settings(storage="f32", seed=5))
    assert ext["mi355x"]["storage"] == "f32" and ext["mi355x"]["seed"] == 5
    assert ext["mi355x"]["rounding"] == "nearest"
    sr = pa.roundtrip(pa.settings(storage="f16-shifted", rounding="stochastic"))
    assert sr["mi355x"]["rounding"] == "stochastic"
    assert "rounding: stochastic" in pa.properties_string(pa.settings(storage="f16-shifted", rounding="stochastic"))
    assert pa.desc_from_settings(pa.settings(rounding="stochastic")).rounding == 1


def test_validate_rejects_unknown_rounding():
    with pytest.raises(pa.LbmError) as e:
        pa.validate(pa.settings(rounding="upward"))
    assert e.value.code == 1


def test_tile_kernel_option():
    assert pa.roundtrip(pa.settings("npol", tileKernel="block"))["mi355x"]["tileKernel"] == "block"
    assert pa.roundtrip(pa.settings("nsol", storage="f32"))["mi355x"]["tileKernel"] == "auto"
    assert pa.desc_from_settings(pa.settings("npol", tileKernel="sweep")).tile_kernel == 1
    assert pa.desc_from_settings(pa.settings("npol", tileKernel="staged")).tile_kernel == 3
    assert pa.roundtrip(pa.settings("nsol", tileKernel="staged"))["mi355x"]["tileKernel"] == "staged"
    assert pa.desc_from_settings(pa.settings("npol")).tile_kernel == 0
    with pytest.raises(pa.LbmError) as e:
        pa.validate(pa.settings("npol", tileKernel="warp"))
    assert e.value.code == 1


@pytest.mark.parametrize("field,value", [("algorithm", "xptl"), ("dataLayout", "soa"), ("timeSteps", 0),
                                         ("workGroupSize", 0), ("frameUpdateInterval", 0)])
def test_validate_rejects(field, value):
    doc = pa.settings()
    doc["algorithmic"][field] = value
    with pytest.raises(pa.LbmError) as e:
        pa.validate(doc)
    assert e.value.code == 1  # PropertyArgumentException


def test_validate_rejects_physics_and_scenario():
    for path, value in [(("physical", "inletDensity"), 0.0), (("physical", "outletDensity"), -1.0),
                        (("physical", "relaxationTime"), 0.0), (("domain", "scenario"), "moon")]:
        doc = pa.settings()
        doc[path[0]][path[1]] = value
        with pytest.raises(pa.LbmError):
            pa.validate(doc)
    with pytest.raises(pa.LbmError) as e:
        pa.validate('{"algorithmic": {}}')
    assert e.value.code == 2  # MissingPropertyException
    with pytest.raises(pa.LbmError) as e:
        pa.validate('{"algorithmic": ')
    assert e.value.code == 3  # UnknownPropertyException (syntax)


def test_nsol_minimum_work_group():
    with pytest.raises(pa.LbmError) as e:
        pa.domain_extents(pa.settings("nsol", work_group_size=5))
    assert e.value.code == 1
    with pytest.raises(pa.LbmError):
        pa.domain_extents(pa.settings("nsol", work_group_size=20))  # zero-width swap tile


def test_domain_extents_match_oracle():
    rng = np.random.default_rng(11)
    for _ in range(200):
        alg = ["nptl", "lptl", "npol", "nsol"][rng.integers(4)]
        w, h, wg = int(rng.integers(2, 400)), int(rng.integers(2, 200)), int(rng.integers(6, 1025))
        o = ob.domain(alg, w + 2, h + 2, wg)
        doc = pa.settings(alg, inner_h=w, inner_v=h, work_group_size=wg)
        if o is None:
            with pytest.raises(pa.LbmError):
                pa.domain_extents(doc)
            continue
        e = pa.domain_extents(doc)
        assert {k: e[k] for k in o} == o, (alg, w, h, wg)


SCENARIOS = ["Hagen-Poiseuille", "walls", "circle", "square", "plate", "skyscraper", "wing", "porous-circles",
             "cylinder"]


@pytest.mark.parametrize("alg", ["nptl", "lptl", "npol", "nsol"])
@pytest.mark.parametrize("scenario", SCENARIOS)
def test_phase_tables_bit_exact(alg, scenario):
    doc = pa.settings(alg, inner_h=203, inner_v=77, work_group_size=64, scenario=scenario, porosity=0.8,
                      obstacleMargin=12, cylinderX=50, cylinderRadius=20.5)
    s = ob.Simulation(doc)
    assert np.array_equal(pa.build_phase(doc), s.phase())
    s.close()


@pytest.mark.parametrize("alg", ["nptl", "npol"])
def test_porous_table_bit_exact_from_seed_one(alg):
    """porous draws the process-wide std::rand() stream; from a fresh stream
    (srand(1), glibc's initial state) the table equals the oracle's."""
    doc = pa.settings(alg, inner_h=150, inner_v=60, work_group_size=64, scenario="porous")
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    host = pa.build_phase(doc)
    s = ob.Simulation(doc)
    assert np.array_equal(host, s.phase())
    s.close()
    # and a second table in the same process continues the stream (reference behaviour)
    assert not np.array_equal(pa.build_phase(doc), host)


def test_porous_circles_full_size_table():
    """The benchmark's 8192 x 2048 table: bit-exact, porosity 0.7 reached."""
    doc = pa.settings("nptl", inner_h=8192, inner_v=2048, work_group_size=256, scenario="porous-circles",
                      seed=20261202, porosity=0.7)
    host = pa.build_phase(doc)
    s = ob.Simulation(doc)
    assert np.array_equal(host, s.phase())
    s.close()
    fluid = (host.reshape(2050, 8194)[2:2048, 1:8193] == 0).mean()
    assert 0.695 < fluid <= 0.7005


def test_properties_string():
    text = pa.properties_string(pa.settings("nsol", "collision"))
    assert "Algorithm: nsol" in text and "Data layout: collision" in text
