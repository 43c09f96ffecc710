"""Host side of the learned warm start (portlbm_amd/surrogate.py): input and
target encoding, decoding back to lattice fields, trapped-momentum nodes,
model files.  No GPU needed."""
import numpy as np
import pytest

from portlbm_amd import surrogate as sg


def test_features_layout():
    mask = np.ones((16, 32), bool)
    mask[4:8, 10:14] = False
    x = sg.features(mask, 0.725, 1.03, 1.0)
    assert x.shape == (4, 16, 32) and x.dtype == np.float32
    np.testing.assert_array_equal(x[0], mask)
    np.testing.assert_allclose(x[1, 0], (np.arange(32) + 0.5) / 32 * 2 - 1, rtol=1e-6)
    assert np.all(x[2] == 0) and np.all(np.abs(x[3]) < 1e-6)  # centre of the tau and drho ranges


def test_encode_decode_round_trip():
    r = np.random.default_rng(0)
    mask = r.random((16, 16)) < 0.8
    tau, rin, rout = 0.8, 1.04, 1.0
    rho = 1.0 + 0.04 * r.random(256)
    jx, jy = 0.01 * r.normal(size=256), 0.01 * r.normal(size=256)
    t = sg.encode_targets(rho, jx, jy, mask, tau, rin, rout, j_scale=30.0, rho_scale=4.0)
    assert np.all(t[:, ~mask] == 0)
    r2, jx2, jy2 = sg.decode(t, mask, tau, rin, rout, j_scale=30.0, rho_scale=4.0)
    m = mask.ravel()
    keep_x = m & ~sg.trapped(mask, 1).ravel()
    keep_y = m & ~sg.trapped(mask, 0).ravel()
    np.testing.assert_allclose(r2[m], rho[m], rtol=1e-6)
    np.testing.assert_allclose(jx2[keep_x], jx[keep_x], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(jy2[keep_y], jy[keep_y], rtol=1e-5, atol=1e-9)
    assert np.all(r2[~m] == rout) and np.all(jx2[~m] == 0) and np.all(jy2[~m] == 0)
    # density only, and the blend halves the momentum
    _, jx0, jy0 = sg.decode(t, mask, tau, rin, rout, 30.0, 4.0, j_blend=0.0)
    assert not jx0.any() and not jy0.any()
    _, jxh, _ = sg.decode(t, mask, tau, rin, rout, 30.0, 4.0, j_blend=0.5)
    np.testing.assert_allclose(jxh[keep_x], 0.5 * jx[keep_x], rtol=1e-5, atol=1e-9)


def test_trapped_nodes():
    # a one-node-wide vertical gap between two solid columns, and a horizontal one
    mask = np.ones((9, 9), bool)
    mask[:, 3] = False
    mask[:, 5] = False
    tx = sg.trapped(mask, 1)
    assert tx[2:7, 4].all()                 # column 4 between solid columns 3 and 5
    assert not tx[:, [0, 1, 2, 6, 7, 8]].any()
    # rows beyond the top and bottom count as solid, columns beyond in/outlet do not
    m2 = np.zeros((3, 5), bool)
    m2[1, :] = True
    ty = sg.trapped(m2, 0)
    assert ty[1, 1:-1].all() and not ty[1, [0, -1]].any()  # the end columns pull from the in/outlet ghosts
    assert not sg.trapped(m2, 1).any()


def test_masked_smoothing_keeps_constants():
    mask = np.random.default_rng(1).random((32, 32)) < 0.7
    a = np.full((32, 32), 1.02)
    rho, _, _ = sg.decode(np.stack([np.full((32, 32), 0.5), np.zeros((32, 32)), np.zeros((32, 32))]), mask,
                          0.7, 1.04, 1.0, 1.0, 1.0, 1.0, rho_sigma=2.0)
    np.testing.assert_allclose(rho[mask.ravel()], a[mask], rtol=1e-12)


def test_model_file_round_trip(tmp_path):
    p = np.arange(10, dtype=np.float32)
    m = sg.Model(p, 256, 256, 39.5, 4.0, 0.5, 2.0, {"samples": 3}, rho_profile=8.0)
    path = tmp_path / "m.npz"
    m.save(path)
    m2 = sg.Model.load(path)
    np.testing.assert_array_equal(m2.params, p)
    assert (m2.height, m2.width, m2.j_scale, m2.rho_scale, m2.j_blend, m2.rho_sigma, m2.rho_profile) == \
        (256, 256, 39.5, 4.0, 0.5, 2.0, 8.0)
    assert m2.meta == {"samples": 3}
    with np.load(path, allow_pickle=False) as z:  # nothing pickled
        assert set(z.files) == {"params", "shape", "encoding", "meta"}
    # a file written before the rho_profile option loads with it off
    np.savez(path, params=p, shape=np.array([256, 256]), encoding=np.array([39.5, 4.0, 0.5, 2.0]),
             meta=np.array("{}"))
    assert sg.Model.load(path).rho_profile == 0.0


def _channel(nx=40, ny=12, rin=1.04, rout=1.0):
    """All-fluid channel, column means falling linearly, one flux."""
    x = np.arange(nx)
    mean = rin - (rin - rout) * (x + 0.5) / nx
    return np.ones((ny, nx), bool), np.tile(mean, (ny, 1)), np.full((ny, nx), 0.01)


def test_darcy_profile_keeps_deviations_and_spans_the_drop():
    mask, rho, jx = _channel()
    dev = 0.001 * np.sin(np.arange(12))[:, None] * np.cos(np.arange(40))[None, :]
    dev = dev - dev.mean(axis=0)
    bad = 1.0 + 0.8 * (rho - 1.0) + 0.003  # the right shape, 80 % of the drop, offset
    out = sg.darcy_profile(bad + dev, jx, mask, 1.04, 1.0, 4.0)
    mean = out.mean(axis=0)
    np.testing.assert_allclose(out - mean[None, :], dev, atol=1e-12)
    assert np.all(np.diff(mean) < 0) and 1.0 < mean[-1] < mean[0] < 1.04
    # about half a column's fall between each boundary density and its column
    assert 0.25 < (1.04 - mean[0]) / (mean[0] - mean[1]) < 1.0
    assert 0.25 < (mean[-1] - 1.0) / (mean[-2] - mean[-1]) < 1.0


def test_darcy_profile_mirror_symmetry():
    r = np.random.default_rng(5)
    mask = r.random((16, 48)) < 0.8
    rho = 1.0 + 0.03 * r.random((16, 48))
    jx = 0.01 + 0.005 * r.random((16, 48))
    a = sg.darcy_profile(rho, jx, mask, 1.03, 1.0, 3.0)
    b = sg.darcy_profile((2.03 - rho)[:, ::-1], jx[:, ::-1], mask[:, ::-1], 1.03, 1.0, 3.0)
    np.testing.assert_allclose(np.where(mask, a, 0), np.where(mask, 2.03 - b[:, ::-1], 0), atol=1e-12)


def test_darcy_profile_follows_the_flux():
    """Where the predicted flux is twice as large for the same fall the column
    resistance halves, so the rebuilt profile falls half as steeply there; a
    column without fluid takes interpolated values and stays solid."""
    mask, rho, jx = _channel()
    jx[:, 20:] *= 2.0
    mask[:, 5] = False
    out = sg.darcy_profile(rho, jx, mask, 1.04, 1.0, 1e-3)
    mean = out[0]
    steep = np.diff(mean[2:18]).mean()
    flat = np.diff(mean[22:38]).mean()
    assert flat == pytest.approx(steep / 2.0, rel=1e-6)
    assert np.all(out[:, 5] == rho[:, 5])
    assert np.all(np.diff(mean[np.r_[0:5, 6:40]]) < 0)


def test_porous_docs_ranges():
    docs = sg.porous_docs(50, 3, 256)
    for d in docs:
        tau, rin, rout = sg._physical(d)
        assert 0.55 <= tau <= 0.9 and 1.01 <= rin <= 1.05 and rout == 1.0
        assert 0.6 <= d["mi355x"]["porosity"] <= 0.9
        assert d["domain"]["horizontalNodes"] == d["domain"]["verticalNodes"] == 256
    assert sg.porous_docs(5, 3)[0] == docs[0]
