"""Split (debug) steps -- stream, outlet, moments, collision as separate
launches over an FP64 pass buffer (the reference's debug kernels,
nptl_kernels.hpp:48-295) -- equal the fused steps bit for bit on every fluid
node and in the results, and each pass does what it says:
checked against host arithmetic on the pass buffer."""
import numpy as np
import pytest

import portlbm_amd as pa

pytestmark = pytest.mark.gpu

CX = np.array([d % 3 - 1 for d in range(9)], dtype=float)
CY = np.array([d // 3 - 1 for d in range(9)], dtype=float)
W = np.array([1 / 36, 1 / 9, 1 / 36, 1 / 9, 4 / 9, 1 / 9, 1 / 36, 1 / 9, 1 / 36])


def _doc(alg, layout, storage="f64", rounding="nearest", scenario="circle", tile="auto"):
    wg = 64 if alg in ("npol", "nsol") else 128
    return pa.settings(alg, layout, inner_h=80, inner_v=36, scenario=scenario, work_group_size=wg, time_steps=1,
                       tau=0.7, rho_in=1.05, storage=storage, rounding=rounding, tileKernel=tile)


SPLIT_CASES = [
    ("nptl", "stream", "f64", "nearest"), ("nptl", "collision", "f32", "nearest"),
    ("lptl", "bundle", "f64", "nearest"), ("npol", "stream", "f64", "nearest"),
    ("nsol", "collision", "f64", "nearest"), ("nsol", "bundle", "f16-shifted", "stochastic"),
    ("lptl", "stream", "bf16-shifted", "nearest"), ("nptl", "collision", "f64", "nearest"),
    ("nptl", "collision", "f16-shifted", "stochastic"), ("lptl", "collision", "bf16-shifted", "nearest"),
    ("nptl", "bundle", "f32", "nearest"), ("lptl", "bundle", "f16-shifted", "nearest"),
    ("nptl", "stream", "f16-shifted", "stochastic"), ("npol", "bundle", "f32", "nearest"),
    ("nsol", "stream", "f32", "nearest")]


# npol/nsol with both step kernels of narrow tiles (the sweep and the block kernel)
@pytest.mark.parametrize("alg,layout,storage,rounding,tile",
                         [c + (t,) for c in SPLIT_CASES
                          for t in (("sweep", "block") if c[0] in ("npol", "nsol") else ("auto",))])
def test_split_steps_equal_fused_steps(alg, layout, storage, rounding, tile):
    doc = _doc(alg, layout, storage, rounding, tile=tile)
    with pa.Lattice.from_settings(doc) as fused, pa.Lattice.from_settings(doc) as split:
        # fluid slots: solid slots (and the npol/nsol buffer copies of them) are never read, and
        # the fused kernels zero some of them by wave segments
        keep = _fluid_slots(fused)
        fused.run(37)
        split.run_split(37)
        assert split.iteration() == 37
        np.testing.assert_array_equal(split.populations()[keep], fused.populations()[keep])
        for a, b in zip(split.macros(), fused.macros()):
            np.testing.assert_array_equal(a, b)
        # and the two keep agreeing when the fused run continues a split one
        split.run(5)
        fused.run_split(5)
        np.testing.assert_array_equal(split.populations()[keep], fused.populations()[keep])


# The staged kernels (LDS row staging, flat or range-by-range copy, 8/4/2 node
# rows per block) on shapes that leave partial row blocks and partial 64-node
# segments, against the split passes, which pull straight from HBM.
@pytest.mark.parametrize("alg,layout,storage", [
    ("nptl", "collision", "f64"), ("nptl", "collision", "f32"), ("lptl", "collision", "f16-shifted"),
    ("nptl", "collision", "bf16-shifted"), ("nptl", "bundle", "f64"), ("lptl", "bundle", "f32")])
@pytest.mark.parametrize("w,h", [(37, 21), (130, 43), (3, 3), (65, 9)])
def test_staged_kernels_equal_split_steps_on_ragged_shapes(alg, layout, storage, w, h):
    doc = pa.settings(alg, layout, inner_h=w, inner_v=h, scenario="circle" if min(w, h) > 8 else "walls",
                      work_group_size=128, time_steps=1, tau=0.7, rho_in=1.05, storage=storage)
    with pa.Lattice.from_settings(doc) as fused, pa.Lattice.from_settings(doc) as split:
        keep = _fluid_slots(fused)
        fused.run(23)
        split.run_split(23)
        np.testing.assert_array_equal(split.populations()[keep], fused.populations()[keep])
        for a, b in zip(split.macros(), fused.macros()):
            np.testing.assert_array_equal(a, b)


# npol/nsol tiles with several row groups, so that the tile sweep's hand-down
# between groups, its side arrays and (FP32) two nodes per lane are checked in
# the reduced-precision storages too: 32 x 32 tiles (FP32: groups of 4 rows),
# 30 x 31 (last group partial), 16 x 16 with the buffered-outlet quirk.
@pytest.mark.parametrize("alg,w,h,wg,layout,storage,rounding", [
    ("npol", 100, 70, 1024, "stream", "f32", "nearest"),
    ("nsol", 150, 50, 1024, "bundle", "f32", "nearest"),
    ("nsol", 150, 50, 1024, "stream", "f16-shifted", "stochastic"),
    ("npol", 97, 40, 256, "collision", "bf16-shifted", "nearest"),
    ("npol", 97, 40, 256, "stream", "f16-shifted", "nearest")])
def test_tile_sweep_equals_split_steps_on_multi_group_tiles(alg, w, h, wg, layout, storage, rounding):
    doc = pa.settings(alg, layout, inner_h=w, inner_v=h, scenario="circle", work_group_size=wg, time_steps=1,
                      tau=0.7, rho_in=1.05, storage=storage, rounding=rounding, tileKernel="sweep")
    with pa.Lattice.from_settings(doc) as fused, pa.Lattice.from_settings(doc) as split:
        keep = _fluid_slots(fused)
        fused.run(41)
        split.run_split(41)
        np.testing.assert_array_equal(split.populations()[keep], fused.populations()[keep])
        for a, b in zip(split.macros(), fused.macros()):
            np.testing.assert_array_equal(a, b)


def _fluid_slots(lat):
    """Population indices (layout order of lbm_get_populations) of the fluid nodes."""
    N = lat.N
    ok = lat.phase() == 0
    n = np.arange(N)
    layout = lat.desc.layout
    idx = []
    for d in range(9):
        if layout == pa.LAYOUTS["collision"]:
            idx.append(9 * n + d)
        elif layout == pa.LAYOUTS["stream"]:
            idx.append(d * N + n)
        else:
            idx.append(3 * (d // 3) * N + d % 3 + 3 * n)
    mask = np.zeros(9 * N, bool)
    for d in range(9):
        mask[idx[d][ok]] = True
    return mask


def test_each_pass_matches_host_arithmetic():
    doc = _doc("nptl", "stream")
    with pa.Lattice.from_settings(doc) as lat:
        lat.run(20)
        H, Hx, N = lat.H, lat.desc.expanded_horizontal_nodes, lat.N
        phase = lat.phase()
        codes = lat.node_codes()
        before = lat.populations().reshape(9, N)  # stream layout: direction-major
        seen = {}

        def cb(p, it):
            seen[pa.PASSES[p]] = (lat.pass_populations().copy(), it)
            if p == 2:
                seen["macros"] = lat.macros()
            if p == 3:
                seen["after"] = lat.populations().reshape(9, N).copy()

        lat.run_split(1, cb)
        assert [seen[k][1] for k in pa.PASSES] == [20, 20, 20, 21]
        fluid = np.nonzero(phase == 0)[0]
        # stream: f_d(n) = f_d(n - c_d), or the node's own f_{8-d} where that upstream node is solid
        streamed = seen["stream"][0]
        for d in range(9):
            up = fluid - int(CY[d]) * Hx - int(CX[d])
            bounce = (codes[fluid] >> d) & 1 == 1
            expect = np.where(bounce, before[8 - d, fluid], before[d, up])
            np.testing.assert_array_equal(streamed[d, fluid], expect, err_msg=f"d={d}")
        # outlet: only the three populations from the ghost column of the x = H-2 nodes change
        outl = seen["outlet"][0]
        is_outlet = (codes[fluid] & pa.CODE_OUTLET) != 0
        assert is_outlet.any()
        for d in range(9):
            changed = outl[d, fluid] != streamed[d, fluid]
            assert not changed[~is_outlet].any()
            if d not in (0, 3, 6):
                assert not changed.any()
        # moments of the pass buffer are the results (ρ, j; velocity is momentum)
        f = outl[:, fluid]
        rho = f.sum(axis=0)
        jx = (f * CX[:, None]).sum(axis=0)
        jy = (f * CY[:, None]).sum(axis=0)
        x, y = fluid % Hx, fluid // Hx
        r = (x - 1) + (y - 1) * (H - 2)
        np.testing.assert_allclose(seen["macros"][0][r], rho, rtol=1e-15)
        np.testing.assert_allclose(seen["macros"][1][r], jx, rtol=1e-13, atol=1e-15)  # sums of O(0.1) terms in another order
        np.testing.assert_allclose(seen["macros"][2][r], jy, rtol=1e-13, atol=1e-15)  # sums of O(0.1) terms in another order
        # collision: f' = -(1/tau)(f - feq) + f, incompressible feq
        cu = CX[:, None] * jx + CY[:, None] * jy
        feq = W[:, None] * (rho + 3 * cu + 4.5 * cu * cu - 1.5 * (jx * jx + jy * jy))
        np.testing.assert_allclose(seen["after"][:, fluid], -(1 / 0.7) * (f - feq) + f, rtol=1e-13, atol=1e-15)


def test_handler_debug_mode_records_every_iteration():
    doc = pa.settings("nptl", "stream", inner_h=16, inner_v=8, scenario="circle", work_group_size=64, time_steps=6,
                      debug_mode=True)
    plain = dict(doc, algorithmic=dict(doc["algorithmic"], debugMode=False))
    with pa.Handler(doc) as dbg, pa.Handler(plain) as ref:
        dbg.start()
        dbg.block_until_finished()
        ref.start()
        ref.block_until_finished()
        hist = dbg.debug_history("density").reshape(6, 16 * 8)
        np.testing.assert_array_equal(hist[-1], ref.field("density"))
        np.testing.assert_array_equal(dbg.field("x_velocity"), ref.field("x_velocity"))
        assert not np.array_equal(hist[0], hist[-1])
        assert dbg.debug_history("y_velocity").size == 6 * 16 * 8
        assert ref.debug_history("density").size == 0
