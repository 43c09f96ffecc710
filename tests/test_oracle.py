"""Pins the CPU oracle before it is trusted as the checker of the GPU library.

The reference ships no numeric fixtures for this path (its tests are unit
known-answer tests and analytic invariants), so every oracle function is
pinned by one of: the reference's own known-answer tests, a library routine
that fixes the same result (glibc rand), closed forms of the equilibrium /
outlet algebra, the Poiseuille regression of the reference, cross-algorithm
equality, and brute-force recomputation on small lattices.
"""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

import portlbm_amd as pa
from oracle import binding as ob

W = np.array([1 / 36, 1 / 9, 1 / 36, 1 / 9, 4 / 9, 1 / 9, 1 / 36, 1 / 9, 1 / 36])
CX = np.array([d % 3 - 1 for d in range(9)])
CY = np.array([d // 3 - 1 for d in range(9)])


def test_glibc_rand_restatement_matches_libc():
    """porous_medium_mask draws std::rand() from seed 1 (domain_initialization.hpp:277-278)."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    ref = np.array([libc.rand() for _ in range(5000)], dtype=np.int64)
    assert np.array_equal(ob.glibc_rand_seed1(5000).astype(np.int64), ref)
    assert ref[0] == 1804289383  # the well-known first output of glibc rand() with seed 1


# Domain extents quoted for the benchmark configurations (from simulation.cpp:223-382).
EXTENTS = [
    ("lptl", 256, 64, 256, (258, 66, 17028)),
    ("nptl", 256, 64, 256, (258, 66, 17028)),
    ("nptl", 256, 64, 1024, (258, 66, 17028)),
    ("npol", 256, 64, 256, (273, 69, 18837)),
    ("nsol", 256, 64, 256, (286, 81, 23166)),
    ("lptl", 8192, 2048, 1024, (8194, 2050, 16797700)),
    ("nptl", 8192, 2048, 1024, (8194, 2050, 16797700)),
    ("npol", 8192, 2048, 1024, (8449, 2113, 17852737)),
    ("nsol", 8192, 2048, 1024, (8495, 2145, 18221775)),
]


@pytest.mark.parametrize("alg,w,h,wg,expect", EXTENTS)
def test_domain_extents_known_answers(alg, w, h, wg, expect):
    e = ob.domain(alg, w + 2, h + 2, wg)
    assert (e["Hx"], e["Vx"], e["N"]) == expect


GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_kats.json").read_text())
DOMAIN_FIELDS = {"horizontal_nodes": "Hx", "vertical_nodes": "Vx", "total_node_count": "N",
                 "subdomain_horizontal_nodes": "sub_h", "subdomain_vertical_nodes": "sub_v",
                 "subdomain_count_horizontal": "cnt_h", "subdomain_count_vertical": "cnt_v"}


@pytest.mark.parametrize("case", GOLDEN["domain"], ids=lambda c: f"{c['algorithm']}-wg{c['wg']}-{c['source']}")
def test_domain_golden_cases_oracle_and_host_library(case):
    """core::Domain known answers of the reference's tests/test_domain.cpp, for the
    oracle and for the host library's Domain (the lattice extents the C ABI is given)."""
    H, V = case["inner_h"] + 2, case["inner_v"] + 2
    doc = pa.settings(case["algorithm"], "stream", inner_h=case["inner_h"], inner_v=case["inner_v"],
                      work_group_size=case["wg"])
    host = pa.domain_extents(doc)
    for impl in (ob.domain(case["algorithm"], H, V, case["wg"]), host):
        assert impl["N"] == impl["Hx"] * impl["Vx"]
        for field, value in case.get("equal", {}).items():
            assert impl[DOMAIN_FIELDS[field]] == value, (field, impl)
        for field, value in case.get("at_least", {}).items():
            assert impl[DOMAIN_FIELDS[field]] >= value, (field, impl)
        if "subdomain_product" in case:
            assert impl["sub_h"] * impl["sub_v"] == case["subdomain_product"]


def test_domain_closed_forms_and_reference_unit_cases():
    # tests/test_domain.cpp: nptl wg 4 -> 2x2, wg 8 -> sub_h*sub_v = 8, wg 7 -> 1x7; npol wg 16 -> 4x4
    assert (ob.domain("nptl", 12, 12, 4)["sub_h"], ob.domain("nptl", 12, 12, 4)["sub_v"]) == (2, 2)
    e = ob.domain("nptl", 12, 12, 8)
    assert e["sub_h"] * e["sub_v"] == 8
    e = ob.domain("nptl", 12, 12, 7)
    assert (e["sub_v"], e["sub_h"]) == (1, 7)
    e = ob.domain("npol", 22, 22, 16)
    assert (e["sub_h"], e["sub_v"]) == (4, 4)
    assert ob.domain("nsol", 12, 12, 5) is None  # below the swap minimum of 6
    # which_power_of_4(20) == 1 (simulation.hpp:401-418): 2 x 2 tiles for nptl,
    # a zero-width swap tile for nsol
    e = ob.domain("nptl", 12, 12, 20)
    assert (e["sub_h"], e["sub_v"]) == (2, 2)
    assert ob.domain("nsol", 12, 12, 20) is None
    rng = np.random.default_rng(7)
    for _ in range(300):
        alg = ["nptl", "npol", "nsol"][rng.integers(3)]
        w, h = int(rng.integers(2, 300)), int(rng.integers(2, 200))
        wg = int(rng.integers(6, 1025))
        e = ob.domain(alg, w + 2, h + 2, wg)
        if e is None:
            assert alg == "nsol"  # a zero-width swap tile (see test below)
            continue
        cnt_h, cnt_v = -(-w // e["sub_h"]), -(-h // e["sub_v"])
        assert (e["cnt_h"], e["cnt_v"]) == (cnt_h, cnt_v)
        if alg == "nptl":
            assert (e["Hx"], e["Vx"]) == (e["sub_h"] * cnt_h + 2, e["sub_v"] * cnt_v + 2)
        else:
            assert (e["Hx"], e["Vx"]) == ((e["sub_h"] + 1) * cnt_h + 1, (e["sub_v"] + 1) * cnt_v + 1)
        assert e["N"] == e["Hx"] * e["Vx"]


def _doc(alg="nptl", layout="stream", **kw):
    base = dict(inner_h=40, inner_v=18, work_group_size=64, time_steps=1, tau=0.7, rho_in=1.0, rho_out=1.0)
    base.update(kw)
    return pa.settings(alg, layout, **base)


def test_initial_state_is_rest_equilibrium_with_inlet_outlet_ghosts():
    doc = _doc("lptl", rho_in=1.2, rho_out=0.9, inner_h=12, inner_v=6)
    s = ob.Simulation(doc)
    f = s.populations().reshape(9, -1)  # stream layout: one row per direction
    H, V = 14, 8
    g = f.reshape(9, V, H)
    # rest everywhere except the ghost columns of rows 1..V-2 (feq(rho, 0) = w rho)
    assert np.array_equal(g[:, 0, 5], W)
    assert np.allclose(g[:, 3, 0], W * 1.2, rtol=0, atol=1e-16)
    assert np.allclose(g[:, 3, H - 1], W * 0.9, rtol=0, atol=1e-16)
    assert np.array_equal(g[:, 0, 0], W) and np.array_equal(g[:, V - 1, H - 1], W)
    s.close()


@pytest.mark.parametrize("alg", ["nptl", "lptl", "npol", "nsol"])
def test_rest_state_is_a_fixed_point(alg):
    """rho_in = rho_out = 1, u = 0: the lattice stays at w_d (the BGK rest
    equilibrium is a fixed point of collision, streaming and bounce-back)."""
    doc = _doc(alg, scenario="circle")
    s = ob.Simulation(doc)
    s.run(50)
    rho, jx, jy, _ = s.results()
    fluid = rho > -1
    assert np.abs(rho[fluid] - 1).max() < 1e-14
    assert np.abs(jx[fluid]).max() < 1e-14 and np.abs(jy[fluid]).max() < 1e-14
    s.close()


# ---- the BGK collision, pinned by what the mathematics fixes ----
#
# A spatially uniform state is left unchanged by streaming, so after one step
# every node away from walls, inlet and outlet holds collide(g) of the uniform
# populations g.  The nine moments M = sum_d f_d phi_d(c_d) of D2Q9 (1, cx,
# cy, cx^2, cy^2, cx cy, cx^2 cy, cx cy^2, cx^2 cy^2) determine f; BGK
# relaxes each towards its equilibrium value with the same factor,
# M' = (1 - 1/tau) M + (1/tau) M_eq, and the lattice's weights fix M_eq of the
# incompressible equilibrium in closed form (sums of w_d c^k over the nine
# velocities): rho, jx, jy, rho/3 + jx^2, rho/3 + jy^2, jx jy, jy/3, jx/3,
# rho/9 + |j|^2/3.  None of this restates feq_d: the check holds for any
# collision that is BGK towards that equilibrium and fails otherwise.


def _moment_basis():
    return np.stack([np.ones(9), CX, CY, CX ** 2, CY ** 2, CX * CY, CX ** 2 * CY, CX * CY ** 2,
                     CX ** 2 * CY ** 2]).astype(np.float64)


def _moments_eq(rho, jx, jy):
    return np.array([rho, jx, jy, rho / 3 + jx * jx, rho / 3 + jy * jy, jx * jy, jy / 3, jx / 3,
                     rho / 9 + (jx * jx + jy * jy) / 3])


def _collision_moment_error(tau, real="f64", lib_path=None, seed=0, alg="nptl", layout="stream"):
    """max |M'(node) - ((1 - 1/tau) M + M_eq / tau)| over the interior nodes
    after one step from a uniform random state; also the error of the
    reported (pre-collision) macroscopic values."""
    inner_h, inner_v = 24, 16
    doc = pa.settings(alg, layout, inner_h=inner_h, inner_v=inner_v, work_group_size=64, time_steps=1, tau=tau,
                      rho_in=1.0, rho_out=1.0)
    s = ob.Simulation(doc, real=real, lib_path=lib_path)
    rng = np.random.default_rng(seed)
    g = W * (1 + 0.6 * (rng.random(9) - 0.5))  # |j| ~ 0.05-0.1
    g = g.astype(s.dtype).astype(np.float64)
    N = s.N
    vals = np.empty(9 * N)
    nodes = np.arange(N)
    for d in range(9):
        vals[ob_layout_index(layout, nodes, d, N)] = g[d]
    s.set_populations(vals)
    s.run(1)
    f = s.populations().astype(np.float64)
    rho_r, jx_r, jy_r, jabs_r = s.results()
    Hx = s.extents["Hx"]
    H, V = inner_h + 2, inner_v + 2
    s.close()
    B = _moment_basis()
    m0 = B @ g
    rho, jx, jy = m0[0], m0[1], m0[2]
    expect = (1 - 1 / tau) * m0 + _moments_eq(rho, jx, jy) / tau
    err = 0.0
    # walls at rows 1 and V-2 (bounce-back next to them), outlet at x = H-2
    for y in range(3, V - 3):
        for x in range(1, H - 3):
            n = x + y * Hx  # nptl/lptl: expanded == unexpanded coordinates
            fn = np.array([f[ob_layout_index(layout, n, d, N)] for d in range(9)])
            err = max(err, float(np.abs(B @ fn - expect).max()))
            r = (x - 1) + (y - 1) * (H - 2)
            err = max(err, abs(rho_r[r] - rho), abs(jx_r[r] - jx), abs(jy_r[r] - jy),
                      abs(jabs_r[r] - np.hypot(jx, jy)))
    return err


def ob_layout_index(layout, n, d, N):
    """The reference's population index (access.hpp:56-60, 81-85, 106-110)."""
    if layout == "collision":
        return 9 * n + d
    if layout == "stream":
        return N * d + n
    return 3 * (d // 3) * N + (d % 3) + 3 * n


@pytest.mark.parametrize("tau", [0.6, 0.8, 1.3])
@pytest.mark.parametrize("alg,layout", [("nptl", "stream"), ("lptl", "collision"), ("nptl", "bundle")])
def test_collision_relaxes_all_nine_moments_fp64(tau, alg, layout):
    for seed in range(3):
        assert _collision_moment_error(tau, "f64", seed=seed, alg=alg, layout=layout) < 1e-15


@pytest.mark.parametrize("tau", [0.52, 0.6, 0.8, 1.3])
def test_collision_relaxes_all_nine_moments_fp32_instance(tau):
    """The USE_FLOAT instance: the same law within float rounding."""
    for seed in range(3):
        assert _collision_moment_error(tau, "f32", seed=seed) < 5e-7


def _poiseuille_peak_ratio(tau, real="f64", lib_path=None):
    """Hagen-Poiseuille channel to steady state: centre-line momentum over
    G W^2 / (8 nu), nu = (tau - 1/2) / 3, G = -dp/dx with p = rho / 3 read off
    the middle third of the channel; W = 12 fluid rows between half-way walls."""
    inner_h, inner_v, steps = 60, 14, 40000
    doc = pa.settings("nptl", "stream", inner_h=inner_h, inner_v=inner_v, work_group_size=64, time_steps=steps,
                      tau=tau, rho_in=1.01, rho_out=1.0)
    s = ob.Simulation(doc, real=real, lib_path=lib_path)
    s.run(steps)
    rho, jx, _, _ = s.results()
    s.close()
    rho, jx = rho.reshape(inner_v, inner_h), jx.reshape(inner_v, inner_h)
    xs = np.arange(inner_h // 3, 2 * inner_h // 3)
    G = -np.polyfit(xs, rho[1:-1, xs].mean(axis=0), 1)[0] / 3
    nu = (tau - 0.5) / 3
    return float(jx[1:-1, inner_h // 2].max() / (G * (inner_v - 2) ** 2 / (8 * nu)))


@pytest.mark.parametrize("tau", [0.6, 0.8])
def test_poiseuille_peak_follows_the_bgk_viscosity(tau):
    """Measured 0.986 (tau 0.6) and 0.989 (tau 0.8): the rest is the known
    slip of half-way bounce-back at W = 12."""
    assert abs(_poiseuille_peak_ratio(tau) - 1) < 0.03


@pytest.fixture(scope="module")
def planted_oracles(tmp_path_factory):
    """oracle.c built with each deliberate collision mistake (OR_PLANT)."""
    import subprocess

    src = Path(ob.__file__).resolve().parent / "oracle.c"
    out = {}
    d = tmp_path_factory.mktemp("planted")
    for k in (1, 2, 3, 4):
        so = d / f"liboracle_plant{k}.so"
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c11", f"-DOR_PLANT={k}",
                        "-Wl,-Bsymbolic", "-o", str(so), str(src), "-lm"], check=True)
        out[k] = so
    return out


@pytest.mark.parametrize("plant", [1, 2, 3, 4])
def test_planted_collision_mistakes_are_caught(planted_oracles, plant):
    """omega = tau, 3.0 for 9/2, 1.0 for 3/2, a dropped 3 c.j: the moment
    check fails each by more than 1e-4 (measured 3e-3 to 8e-2; correct: 6e-17)."""
    assert _collision_moment_error(0.6, lib_path=planted_oracles[plant]) > 1e-4


def test_planted_relaxation_mistake_fails_the_poiseuille_check(planted_oracles):
    assert abs(_poiseuille_peak_ratio(0.8, lib_path=planted_oracles[1]) - 1) > 0.5


def test_reduction_sums_and_momentum_exchange_independently():
    """oracle_reduce: mass and momentum equal the sums of the reported
    (pre-collision) fields, which collision conserves; the momentum-exchange
    force equals 2 f_d c_d summed over every fluid-to-solid link found from
    the phase table's geometry (not from the node codes)."""
    doc = pa.settings("nptl", "stream", inner_h=48, inner_v=20, work_group_size=64, scenario="circle",
                      time_steps=300, tau=0.7, rho_in=1.05, rho_out=1.0)
    s = ob.Simulation(doc)
    s.run(300)
    red = s.reduce()
    rho, jx, jy, _ = s.results()
    fluid = rho > -1
    assert red["fluid"] == int(fluid.sum())
    assert abs(red["mass"] - rho[fluid].sum()) < 1e-11
    assert abs(red["jx"] - jx[fluid].sum()) < 1e-13 and abs(red["jy"] - jy[fluid].sum()) < 1e-13
    phase, f, N, Hx = s.phase(), s.populations(), s.N, s.extents["Hx"]
    fx = fy = 0.0
    for n in np.flatnonzero(phase == 0):
        for d in range(9):
            if d != 4 and phase[n + CX[d] + CY[d] * Hx] == 1:
                fx += 2 * f[N * d + n] * CX[d]
                fy += 2 * f[N * d + n] * CY[d]
    s.close()
    assert abs(red["fx"] - fx) < 1e-13 and abs(red["fy"] - fy) < 1e-13
    assert red["fx"] > 0  # the flow pushes the disc and the walls downstream


def test_outlet_closed_form_after_one_step():
    """After the Zou-He update the border node's post-stream density equals
    rho_out and its post-stream jy equals the previous uy of column H-3
    (algebra of non_buffered.hpp:112-129).  The first step's collision at the
    border conserves both, so the reported macroscopic values show them."""
    doc = _doc("nptl", "stream", rho_in=1.1, rho_out=0.95, inner_h=30, inner_v=12)
    s = ob.Simulation(doc)
    for step in range(1, 6):
        _, _, jy_prev, _ = s.results()
        s.run(1)
        rho, jx, jy, _ = s.results()
        H, V = 32, 14
        for y in range(2, V - 2):
            r_border = (H - 3) + (y - 1) * (H - 2)
            r_src = (H - 4) + (y - 1) * (H - 2)
            assert abs(rho[r_border] - 0.95) < 1e-14, (step, y)
            prev = jy_prev[r_src] if step > 1 else 0.0
            assert abs(jy[r_border] - prev) < 1e-15, (step, y)
    s.close()


def test_planted_outlet_sign_error_is_caught():
    """The outlet check above fails for a plausible mistake: a wrong sign of
    the uy term (planted here in a numpy restatement of the outlet)."""
    f1, f4, f7, f2, f5, f8, uy, rho_out = 0.11, 0.44, 0.1, 0.03, 0.12, 0.027, 0.004, 0.95
    jx = -rho_out + (f1 + f4 + f7 + 2 * (f8 + f5 + f2))
    good0 = f8 - 0.5 * (f1 - f7) - jx / 6 - 0.5 * uy
    good6 = f2 + 0.5 * (f1 - f7) - jx / 6 + 0.5 * uy
    bad0 = f8 - 0.5 * (f1 - f7) - jx / 6 + 0.5 * uy
    f3 = f5 - 2 / 3 * jx
    post = lambda f0, f6: np.array([f0, f1, f2, f3, f4, f5, f6, f7, f8])
    good, bad = post(good0, good6), post(bad0, good6)
    assert abs(good.sum() - rho_out) < 1e-15 and abs((good * CY).sum() - uy) < 1e-15
    assert abs((bad * CY).sum() - uy) > 1e-3


@pytest.mark.parametrize("layout", ["collision", "stream", "bundle"])
def test_all_algorithms_agree_on_fluid_fields(layout):
    """The four algorithms are different schedules of one update: on shapes
    where the buffered outlet reaches the ghost column they agree exactly."""
    fields = []
    for alg in ["nptl", "lptl", "npol", "nsol"]:
        s = ob.Simulation(_doc(alg, layout, inner_h=60, inner_v=26, scenario="circle", rho_in=1.05, time_steps=200))
        s.run(200)
        fields.append(np.stack(s.results()))
        s.close()
    for f in fields[1:]:
        assert np.array_equal(f, fields[0])


def test_buffered_outlet_quirk_keeps_the_ghost_constant():
    """When a buffer column separates nodes H-3 and H-2 (H-3 a multiple of
    sub_h), the buffered outlet (buffered.hpp:285-288) addresses the buffer
    and the outlet ghost keeps feq(rho_out): the one-lattice schemes then
    differ from the two-lattice ones near the outlet, and only there."""
    doc = dict(inner_h=33, inner_v=18, work_group_size=16, scenario="circle", rho_in=1.05, time_steps=300)
    res = {}
    for alg in ["nptl", "npol"]:
        s = ob.Simulation(_doc(alg, "stream", **doc))
        s.run(300)
        res[alg] = s.results()[1].reshape(18, 33)
        s.close()
    diff = np.abs(res["nptl"] - res["npol"])
    assert diff.max() > 1e-6            # the quirk changes the flow
    assert diff[:, :2].max() < diff.max()  # and the change is largest at the outlet


def test_poiseuille_regression_invariants():
    """tests/test_regression.cpp:67-182: lptl/stream, 30 x 14 inner, tau 1.0,
    inlet 1.01, 8000 steps: parabolic centre-column profile."""
    doc = pa.settings("lptl", "stream", inner_h=30, inner_v=14, work_group_size=16, time_steps=8000, tau=1.0,
                      rho_in=1.01, rho_out=1.0)
    s = ob.Simulation(doc)
    s.run(8000)
    ux = s.results()[1].reshape(14, 30)
    s.close()
    prof = ux[:, 15]
    fl = prof[1:13]
    assert (fl > 0).all()
    umax = fl.max()
    for i in range(1, 8):
        assert abs(prof[i] - prof[13 - i]) / umax < 0.01
    assert abs(int(np.argmax(fl)) + 1 - 6) <= 1
    assert 1.35 < umax / fl.mean() < 1.55


def test_smoke_invariants():
    """tests/smoke_test.cpp:41-96: lptl/stream HP 126 x 30, tau 0.6, inlet 1.3, 100 steps."""
    doc = pa.settings("lptl", "stream", inner_h=126, inner_v=30, work_group_size=64, time_steps=100, tau=0.6,
                      rho_in=1.3, rho_out=1.0)
    s = ob.Simulation(doc)
    s.run(100)
    rho, jx, jy, _ = s.results()
    s.close()
    assert not np.isnan(rho).any() and not np.isnan(jx).any() and not np.isnan(jy).any()
    assert rho.mean() > 0.5
    assert (np.abs(jx) > 1e-12).any()


def _brute_codes(phase, H, V, Hx, buffered, sub_h, sub_v, outlet_active):
    """Node codes from unexpanded coordinates (independent of the oracle's
    buffer-following derivation)."""
    def mx(u, last, sub):
        if not buffered:
            return u
        if u == 0:
            return 0
        if u == last:
            return (last - 1) + (last - 2) // sub + 1
        return u + (u - 1) // sub
    codes = np.zeros(phase.size, dtype=np.uint16)
    for n in range(phase.size):
        if phase[n] == 1:
            codes[n] = 0x1000
        elif phase[n] == 0:
            xe, ye = n % Hx, n // Hx
            x = xe - xe // (sub_h + 1) if buffered else xe
            y = ye - ye // (sub_v + 1) if buffered else ye
            c = 0x200
            for d in range(9):
                if d == 4:
                    continue
                m = mx(x - CX[d], H - 1, sub_h) + mx(y - CY[d], V - 1, sub_v) * Hx
                if phase[m] == 1:
                    c |= 1 << d
            if outlet_active and x == H - 2 and 2 <= y <= V - 3:
                c |= 0x400
            if x == H - 3 and 2 <= y <= V - 3:
                c |= 0x800
            codes[n] = c
    return codes


@pytest.mark.parametrize("alg,w,h,wg", [("nptl", 37, 21, 64), ("lptl", 30, 15, 64), ("npol", 37, 21, 16),
                                        ("nsol", 37, 21, 64), ("npol", 33, 18, 16)])
def test_node_codes_brute_force(alg, w, h, wg):
    doc = pa.settings(alg, "stream", inner_h=w, inner_v=h, work_group_size=wg, scenario="walls")
    s = ob.Simulation(doc)
    e = s.extents
    buffered = alg in ("npol", "nsol")
    active = not (buffered and (w + 2 - 3) % e["sub_h"] == 0)
    expect = _brute_codes(s.phase(), w + 2, h + 2, e["Hx"], buffered, e["sub_h"], e["sub_v"], active)
    assert np.array_equal(s.node_codes(), expect)
    s.close()


def test_walls_stencil_terms_never_overlap():
    """The walls stencil sums two conditions (domain_initialization.hpp:64-71),
    but their row ranges y < V/2 and y > (V-1)/2 are disjoint for every V, so
    the sum is at most 1 (no node of phase 2) whatever the grid width."""
    for w in range(2, 30):
        for h in (2, 3, 4, 7, 20):
            s = ob.Simulation(pa.settings("lptl", "stream", inner_h=w, inner_v=h, scenario="walls"))
            assert s.phase().max() <= 1
            s.close()
