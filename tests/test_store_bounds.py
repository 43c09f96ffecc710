"""Host-side check of the device store ranges of the two staged step kernels
(k_step_nptl_staged and k_step_tile_staged in portlbm_amd/csrc/lbm_runtime.hip).

Both kernels write their results through LDS as 16-byte vectors whose
addresses they compute from the lattice placement (row pitch P, row offset
X0 = 63, direction stride Sd), plus the side-array slots of the tile kernel.
A store outside the block's own nodes would corrupt another block's nodes, or
land outside the allocation (in the worst case in host memory mapped for the
GPU).  This file restates the kernels' address arithmetic on the CPU and
checks, for every block of every lattice below, that each stored element is
a population of one of the block's own nodes (nptl: the block's row segment;
npol/nsol: the block's tile) and that each side-array store lands in the
array the kernel writes that step.  The arithmetic is the kernels' own
(vectors of VE = 16 / sizeof(storage) elements, partial vectors written
element by element); the lattice geometry comes from the host library.
"""
import numpy as np
import pytest

import portlbm_amd as pa

ELEM = {"f64": 8, "f32": 4, "f16-shifted": 2, "bf16-shifted": 2}
K_OF = {"collision": 9, "stream": 1, "bundle": 3}


def _ru(n, a):
    return (n + a - 1) // a * a


def _placement(ext, elem):
    """P, X0, Sd and the allocation's element count as lbm_create sets them."""
    X0 = 63
    P = _ru(X0 + ext["Hx"], 64)
    M = P * ext["Vx"]
    align_nodes = ((2 << 20) if M * elem >= (32 << 20) else 4096) // elem
    Sd = _ru(M, align_nodes) + 512 // elem
    return P, X0, Sd


def _node_of(e, layout, Sd):
    """(node place, direction) of lattice element e."""
    if layout == "collision":
        return e // 9, e % 9
    if layout == "stream":
        return e % Sd, e // Sd
    a, r = e // (3 * Sd), e % (3 * Sd)
    return r // 3, 3 * a + r % 3


def _staged_shape(layout, storage):
    """Nodes per row segment and node rows per block of k_step_nptl_staged
    (nptl_staged_width, nptl_staged_rows)."""
    if layout != "collision":
        return 64, 1
    return {"f64": (32, 8), "f32": (64, 4)}.get(storage, (32, 4))


def _check_nptl_staged(w, h, layout, storage, row_skew=0):
    ext = pa.domain_extents(pa.settings("nptl", layout, inner_h=w, inner_v=h, work_group_size=256))
    elem = ELEM[storage]
    VE = 16 // elem
    P, X0, Sd = _placement(ext, elem)
    Hx, Vx = ext["Hx"], ext["Vx"]
    B, R = _staged_shape(layout, storage)
    K = 9 if layout == "collision" else 3
    bad = 0
    for by in range((Vx - 2 + R - 1) // R):
        for bx in range((Hx - 2 + B - 1) // B):
            x0, y0 = 1 + bx * B, 1 + by * R
            length, rows = min(B, Hx - 1 - x0), min(R, Vx - 1 - y0)
            b0 = X0 + x0 + y0 * P
            own = {X0 + x + y * P for y in range(y0, y0 + rows) for x in range(x0, x0 + length)}
            arrays = rows if layout == "collision" else 3
            per = K * length
            for a in range(arrays):
                g0 = 9 * (b0 + (a + row_skew) * P) if layout == "collision" else 3 * a * Sd + 3 * (b0 + row_skew * P)
                assert g0 % VE == 0
                nvec = per // VE
                e = np.concatenate([g0 + np.arange(nvec * VE), g0 + np.arange(nvec * VE, per)])
                for x in e:
                    node, d = _node_of(int(x), layout, Sd)
                    bad += node not in own or not (0 <= d < 9) or (layout == "bundle" and d // 3 != a)
    return bad


def _check_tile_staged(alg, w, h, wg, layout, storage, lead=None):
    ext = pa.domain_extents(pa.settings(alg, layout, inner_h=w, inner_v=h, work_group_size=wg))
    sh, sv = ext["sub_h"], ext["sub_v"]
    if not 8 <= sh <= 64:
        return None
    elem = ELEM[storage]
    VE = 16 // elem
    P, X0, Sd = _placement(ext, elem)
    Hx, Vx, cnt_h, cnt_v = ext["Hx"], ext["Vx"], ext["cnt_h"], ext["cnt_v"]
    K = K_OF[layout]
    lead = K if lead is None else lead  # elements from a row's left buffer place to its first tile node
    NRG = 9 // K
    rsl = {"collision": 0, "stream": Sd, "bundle": 3 * Sd}[layout]
    CR = 256 // sh
    al = 256 // elem
    rows_len, cols_len = _ru((cnt_v - 1) * 6 * Hx, al), _ru((cnt_h - 1) * 6 * Vx, al)
    first = _ru(9 * Sd + 64 // elem, al)
    rows_out, cols_out = (first, first + rows_len), (first + 2 * rows_len, first + 2 * rows_len + cols_len)
    bad = 0
    for by in range(cnt_v):
        for bx in range(cnt_h):
            x0, y00 = bx * (sh + 1) + 1, by * (sv + 1) + 1
            own = {X0 + x + y * P for y in range(y00, y00 + sv) for x in range(x0, x0 + sh)}
            e00 = K * (X0 + x0 - 1 + (y00 - 1) * P)
            s1 = (e00 + lead) % VE
            nvo = (s1 + K * sh + VE - 1) // VE
            for k in range((sv + CR - 1) // CR):
                for r in range(min(CR, sv - k * CR)):
                    for ra in range(NRG):
                        g0 = ra * rsl + e00 + lead + K * (1 + k * CR + r) * P
                        v = np.arange(nvo)[:, None] * VE + np.arange(VE)[None, :]
                        keep = (v >= s1) & (v < s1 + K * sh)
                        ga = g0 - s1 + v[keep]
                        bad += int(np.any((g0 - s1 + np.arange(nvo) * VE) % VE))
                        for x in ga:
                            node, d = _node_of(int(x), layout, Sd)
                            bad += node not in own or d // K != ra
            # side-array stores of the tile's edge nodes, in the arrays written this step
            for y in range(y00, y00 + sv):
                for q in range(3):
                    if bx + 1 < cnt_h:
                        i = cols_out[1] + (bx * 6 + q) * Vx + y
                        bad += not cols_out[1] <= i < cols_out[1] + cols_len
                    if bx > 0:
                        i = cols_out[1] + ((bx - 1) * 6 + 3 + q) * Vx + y
                        bad += not cols_out[1] <= i < cols_out[1] + cols_len
            for x in range(x0, x0 + sh):
                for q in range(3):
                    if by + 1 < cnt_v:
                        i = rows_out[1] + (by * 6 + 3 + q) * Hx + x
                        bad += not rows_out[1] <= i < rows_out[1] + rows_len
                    if by > 0:
                        i = rows_out[1] + ((by - 1) * 6 + q) * Hx + x
                        bad += not rows_out[1] <= i < rows_out[1] + rows_len
    return bad


@pytest.mark.parametrize("storage", ["f64", "f32", "bf16-shifted"])
@pytest.mark.parametrize("layout", ["collision", "bundle"])
@pytest.mark.parametrize("w,h", [(96, 40), (130, 43), (37, 21), (2, 2)])
def test_nptl_staged_stores_stay_in_their_rows(w, h, layout, storage):
    assert _check_nptl_staged(w, h, layout, storage) == 0


TILES = [("npol", 96, 40, 256), ("nsol", 96, 40, 256), ("npol", 100, 70, 1024), ("nsol", 150, 50, 1024),
         ("npol", 250, 70, 64), ("npol", 150, 40, 96), ("npol", 97, 40, 256), ("nsol", 85, 40, 256)]


@pytest.mark.parametrize("storage", ["f64", "f32", "f16-shifted"])
@pytest.mark.parametrize("layout", ["collision", "stream", "bundle"])
@pytest.mark.parametrize("alg,w,h,wg", TILES)
def test_tile_staged_stores_stay_in_their_tile(alg, w, h, wg, layout, storage):
    bad = _check_tile_staged(alg, w, h, wg, layout, storage)
    assert bad is not None, "tile shape outside the staged kernel's range"
    assert bad == 0


def test_the_check_sees_misplaced_stores():
    """A row pitch off by one row (nptl) or output rows that start on the
    left buffer place instead of the first tile node (tile kernel) fail."""
    assert _check_nptl_staged(96, 40, "collision", "bf16-shifted", row_skew=1) > 0
    assert _check_nptl_staged(96, 40, "bundle", "f32", row_skew=1) > 0
    assert _check_tile_staged("npol", 96, 40, 256, "collision", "f64", lead=0) > 0
    assert _check_tile_staged("nsol", 96, 40, 256, "stream", "f16-shifted", lead=0) > 0
