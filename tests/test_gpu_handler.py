"""Handler-level and C-ABI behaviour on the GPU: the reference's own smoke and
regression tests through create_handler, lifecycle (frames, pause/resume,
re-initialisation, waiting), error behaviour, NaN protection, reductions and
the reduced-precision storage forms."""
import ctypes
import json
import subprocess
import time
from pathlib import Path

import numpy as np
import pytest

import portlbm_amd as pa
from oracle import binding as ob
from tests.parity import FP64_RTOL, rel_err

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def test_reference_smoke_case():
    """tests/smoke_test.cpp:41-96: lptl/stream, HP 126 x 30, tau 0.6, inlet 1.3, 100 steps."""
    doc = pa.settings("lptl", "stream", inner_h=126, inner_v=30, work_group_size=64, time_steps=100,
                      frame_update_interval=100, tau=0.6, rho_in=1.3, rho_out=1.0)
    pa.validate(doc)
    with pa.Handler(doc) as h:
        h.start()
        h.block_until_finished()
        rho, ux, uy = h.get_densities(), h.get_x_velocities(), h.get_y_velocities()
    assert rho.size == 126 * 30 and ux.size == rho.size and uy.size == rho.size
    assert not (np.isnan(rho).any() or np.isnan(ux).any() or np.isnan(uy).any())
    assert rho.mean() > 0.5
    assert (np.abs(ux) > 1e-12).any()


@pytest.mark.parametrize("alg", ["lptl", "nptl", "npol", "nsol"])
def test_reference_poiseuille_regression(alg):
    """tests/test_regression.cpp:67-182 (lptl there; every algorithm here):
    HP 30 x 14, tau 1.0, inlet 1.01, 8000 steps, parabolic profile."""
    doc = pa.settings(alg, "stream", inner_h=30, inner_v=14, work_group_size=16, time_steps=8000,
                      frame_update_interval=8000, tau=1.0, rho_in=1.01, rho_out=1.0)
    with pa.Handler(doc) as h:
        h.start()
        h.block_until_finished()
        inner_h, inner_v = h.get_horizontal_nodes() - 2, h.get_vertical_nodes() - 2
        ux = h.get_x_velocities()
    assert ux.size == inner_h * inner_v
    prof = ux.reshape(inner_v, inner_h)[:, inner_h // 2]
    fl = prof[1:inner_v - 1]
    assert (fl > 0).all()
    umax = fl.max()
    for i in range(1, inner_v // 2 + 1):
        assert abs(prof[i] - prof[inner_v - 1 - i]) / umax < 0.01
    peak = int(np.argmax(fl)) + 1
    assert abs(peak - (1 + inner_v - 2) // 2) <= 1
    assert 1.35 < umax / fl.mean() < 1.55


def _oracle_fields(doc, steps):
    s = ob.Simulation(doc)
    s.run(steps)
    r = s.results()
    s.close()
    return r


def test_handler_fields_equal_oracle_and_progress():
    doc = pa.settings("nsol", "bundle", inner_h=90, inner_v=40, work_group_size=64, time_steps=300,
                      frame_update_interval=70, scenario="circle", tau=0.66, rho_in=1.04)
    with pa.Handler(doc) as h:
        h.start()
        h.block_until_finished()
        assert h.is_finished() and h.get_current_iteration() == 300 and h.get_progress() == 1.0
        assert 0 < h.get_last_frametime() < 1e4
        got = [h.field(k) for k in ("density", "x_velocity", "y_velocity", "absolute_velocity")]
    for a, b in zip(got, _oracle_fields(doc, 300)):
        assert rel_err(a, b) <= FP64_RTOL


def test_pause_resume_and_reinitialise():
    steps = 60000
    doc = pa.settings("nptl", "collision", inner_h=64, inner_v=32, work_group_size=64, time_steps=steps,
                      frame_update_interval=1000, scenario="square", tau=0.8, rho_in=1.02)
    ref = _oracle_fields(doc, steps)
    with pa.Handler(doc) as h:
        h.start()
        deadline = time.time() + 10
        while h.get_current_iteration() == 0 and time.time() < deadline:
            time.sleep(0.001)
        h.pause()
        it = h.get_current_iteration()
        assert 0 < it <= steps
        if it < steps:
            assert not h.is_finished()
            assert h.get_current_iteration() == it  # paused: no further steps
        h.start()
        h.block_until_finished()
        assert h.get_current_iteration() == steps
        for a, b in zip([h.field("density"), h.field("x_velocity")], ref[:2]):
            assert rel_err(a, b) <= FP64_RTOL
        # a second wait has nothing to wait for (WaitException)
        with pytest.raises(pa.LbmError) as e:
            h.block_until_finished()
        assert e.value.code == 4
        h.initialize()
        assert h.get_current_iteration() == 0
        assert (h.field("density") == -1).all()
        h.start()
        h.block_until_finished()
        assert rel_err(h.field("x_velocity"), ref[1]) <= FP64_RTOL


def test_work_group_above_device_maximum_is_rejected():
    assert pa.device_max_work_group_size() == 1024
    with pytest.raises(pa.LbmError) as e:
        pa.Handler(pa.settings("nptl", work_group_size=2048, inner_h=32, inner_v=16))
    assert e.value.code == 1


def test_nsol_small_work_group_rewrites_settings_file(tmp_path):
    """Domain::setup_for_swap (simulation.cpp:336-353): work-group below 6 ->
    the settings file is rewritten with 6 and PropertyArgumentException."""
    doc = pa.settings("nsol", "stream", inner_h=40, inner_v=20, work_group_size=4)
    path = tmp_path / "settings.json"
    path.write_text(json.dumps(doc))
    with pytest.raises(pa.LbmError) as e:
        pa.Handler(settings_path=path)
    assert e.value.code == 1
    fixed = json.loads(path.read_text())
    assert fixed["algorithmic"]["workGroupSize"] == 6
    assert fixed["domain"]["horizontalNodes"] == 40 and fixed["domain"]["verticalNodes"] == 20
    with pa.Handler(settings_path=path) as h:  # restart works
        h.start()
        h.block_until_finished()


def test_cli_runs_settings_file(tmp_path):
    doc = pa.settings("npol", "bundle", inner_h=128, inner_v=32, work_group_size=64, time_steps=200)
    path = tmp_path / "s.json"
    path.write_text(json.dumps(doc))
    out = subprocess.run([str(pa.LIB_DIR.parent / "bin" / "portlbm"), str(path)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert "steps: 200" in out.stdout


def test_c_abi_errors():
    lib = pa.hip_library()
    d = pa.desc_from_settings(pa.settings("nptl", inner_h=20, inner_v=10))
    ctx = ctypes.c_void_p()
    assert lib.lbm_create(ctypes.byref(d), ctypes.byref(ctx)) == 0
    assert lib.lbm_run(ctx, 1, 0) == 4  # LBM_ERR_STATE: not initialised
    assert b"lbm_init_populations" in lib.lbm_last_error()
    phase = np.zeros(3, dtype=np.int8)
    assert lib.lbm_set_phase(ctx, ctypes.c_void_p(phase.ctypes.data), 3) == 1
    lib.lbm_destroy(ctx)
    bad = pa.desc_from_settings(pa.settings("npol", inner_h=20, inner_v=10))
    bad.expanded_horizontal_nodes += 1
    assert lib.lbm_create(ctypes.byref(bad), ctypes.byref(ctx)) == 1


@pytest.mark.parametrize("nan_protection", [True, False])
def test_nan_protection_matches_oracle(nan_protection):
    """WITH_NAN_PROTECTION clamps NaN outputs to 0 (nptl_kernels.hpp:413-432);
    the outlet then reads the clamped uy.  A NaN planted in one fluid node
    spreads identically in the GPU library and the oracle."""
    doc = pa.settings("nptl", "stream", inner_h=40, inner_v=16, work_group_size=64, time_steps=30,
                      nanProtection=nan_protection, rho_in=1.05)
    lat = pa.Lattice.from_settings(doc)
    s = ob.Simulation(doc)
    lat.run(5)
    s.run(5)
    pop = s.populations()
    node = 20 + 8 * lat.desc.expanded_horizontal_nodes
    pop[lat.N * 4 + node] = np.nan  # stream layout: direction 4 of one fluid node
    lat.set_populations(pop)
    s.set_populations(pop)
    lat.run(20)
    s.run(20)
    g, o = lat.macros(), s.results()
    for a, b in zip(g, o):
        assert rel_err(a, b) <= FP64_RTOL
    assert (np.isnan(g[0]).any()) != nan_protection
    lat.close()
    s.close()


def test_reductions_match_oracle():
    doc = pa.settings("npol", "collision", inner_h=150, inner_v=60, work_group_size=64, time_steps=200,
                      scenario="cylinder", cylinderX=40, cylinderRadius=10.0, rho_in=1.03)
    lat = pa.Lattice.from_settings(doc)
    s = ob.Simulation(doc)
    lat.run(200)
    s.run(200)
    g, o = lat.reduce(), s.reduce()
    assert g["fluid"] == o["fluid"]
    for k in ("mass", "jx", "jy", "fx", "fy"):
        assert abs(g[k] - o[k]) <= 1e-11 * max(1.0, abs(o[k])), (k, g[k], o[k])
    assert g["fx"] > 0  # the flow pushes the cylinder downstream
    r1 = lat.residual()
    assert r1 == pytest.approx(1.0)
    lat.run(100)
    r2 = lat.residual()
    lat.run(100)
    r3 = lat.residual()
    assert 0 < r3 < r2 < 1
    lat.close()
    s.close()


@pytest.mark.parametrize("storage,bound", [("f32", 1e-4), ("f16-shifted", 1e-3), ("bf16-shifted", 1e-2)])
def test_precision_storage_error(storage, bound):
    """Reduced-precision storage against the FP64 path on the 256 x 64
    channel after 1000 steps: relative L2 error of the velocity field."""
    base = dict(inner_h=256, inner_v=64, work_group_size=256, time_steps=1000, tau=0.6, rho_in=1.2,
                scenario="circle")
    ref = pa.Lattice.from_settings(pa.settings("nptl", "stream", **base))
    red = pa.Lattice.from_settings(pa.settings("nptl", "stream", storage=storage, **base))
    ref.run(1000)
    red.run(1000)
    _, jx0, jy0, _ = ref.macros()
    _, jx1, jy1, _ = red.macros()
    err = np.sqrt(((jx1 - jx0) ** 2 + (jy1 - jy0) ** 2).sum() / (jx0 ** 2 + jy0 ** 2).sum())
    print(storage, "velocity L2 error", err)
    assert err < bound
    m0, m1 = ref.reduce()["mass"], red.reduce()["mass"]
    assert abs(m1 / m0 - 1) < 10 * bound
    ref.close()
    red.close()
