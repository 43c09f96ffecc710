"""Batched domains (lbm_batch_*): every domain of a batch equals the same
domain run alone -- against the oracle in FP64, bit-exactly against the
single-lattice library in FP32 -- in all three launch modes, across repeated
graph replays and after autotuning."""
import json

import numpy as np
import pytest

import portlbm_amd as pa
from oracle import binding as ob
from portlbm_amd import workloads
from tests.parity import FP64_RTOL, rel_err

pytestmark = pytest.mark.gpu


def _small_batch(count=10, seed=7, storage="f32"):
    return workloads.batch_channels(count, seed=seed, storage=storage, nx_range=(30, 300), ny_range=(12, 80))


def test_batch_fp64_matches_oracle():
    docs = _small_batch(storage="f64")
    with pa.Batch(docs, storage="f64") as b:
        b.run(150, "persistent")
        for i, doc in enumerate(docs):
            s = ob.Simulation(doc)
            s.run(150)
            for a, r in zip(b.macros(i), s.results()):
                assert rel_err(a, r) <= FP64_RTOL, i
            s.close()


def test_batch_fp32_equals_single_lattice_bit_for_bit():
    docs = _small_batch(storage="f32")
    with pa.Batch(docs, storage="f32") as b:
        b.run(200, "graph")
        for i, doc in enumerate(docs):
            with pa.Lattice.from_settings(doc) as lat:
                lat.run(200)
                for a, r in zip(b.macros(i), lat.macros()):
                    assert np.array_equal(a, r), i


def test_modes_agree_and_graph_replays_are_stable():
    docs = _small_batch(count=16, seed=3)
    results = {}
    for mode in ("per-domain", "persistent"):
        with pa.Batch(docs) as b:
            b.run(90, mode)
            results[mode] = [b.macros(i) for i in range(len(docs))]
    # graph: the same 30-step graph replayed three times, with host reads between replays
    with pa.Batch(docs) as b:
        for _ in range(3):
            b.run(30, "graph")
            _ = b.macros(0)
        results["graph"] = [b.macros(i) for i in range(len(docs))]
    for mode in ("persistent", "graph"):
        for i in range(len(docs)):
            for a, r in zip(results[mode][i], results["per-domain"][i]):
                assert np.array_equal(a, r), (mode, i)


def test_autotune_keeps_results():
    docs = _small_batch(count=12, seed=5, storage="f32")
    with pa.Batch(docs) as ref:
        ref.run(60, "persistent")
        expect = [ref.macros(i) for i in range(len(docs))]
    with pa.Batch(docs) as b:
        best = b.autotune(steps_per_trial=6)
        assert best > 0
        assert all(w in (64, 128, 256) for w in b.tiles())
        b.run(60, "graph")
        for i in range(len(docs)):
            for a, r in zip(b.macros(i), expect[i]):
                assert np.array_equal(a, r), i


def test_set_tiles_rejects_bad_widths():
    with pa.Batch(_small_batch(count=2)) as b:
        with pytest.raises(pa.LbmError):
            b.set_tiles([64, 96, 128, 256])
        b.set_tiles([64, 64, 128, 256])
        assert b.tiles() == [64, 64, 128, 256]


def _fluid_mask(doc):
    lin = json.loads(json.dumps(doc))
    lin["algorithmic"]["algorithm"] = "lptl"
    phase = pa.build_phase(lin)
    H, V = doc["domain"]["horizontalNodes"] + 2, doc["domain"]["verticalNodes"] + 2
    return (phase.reshape(V, H)[1:-1, 1:-1] == 0).ravel()


def test_residuals_match_host_momentum_difference():
    docs = _small_batch(count=6, seed=11, storage="f64")
    with pa.Batch(docs, storage="f64") as b:
        b.run(120, "persistent")
        first = b.residuals()
        assert np.all(first == 1.0)  # compared against the zero snapshot
        m1 = [b.macros(i) for i in range(len(docs))]
        b.run(40, "persistent")
        r = b.residuals()
        m2 = [b.macros(i) for i in range(len(docs))]
        again = b.residuals()  # no step in between: the snapshot is current
        assert np.all(again == 0.0)
    for i, doc in enumerate(docs):
        fl = _fluid_mask(doc)
        dx, dy = m2[i][1][fl] - m1[i][1][fl], m2[i][2][fl] - m1[i][2][fl]
        host = np.sqrt(np.sum(dx * dx + dy * dy) / np.sum(m2[i][1][fl] ** 2 + m2[i][2][fl] ** 2))
        assert r[i] > 0 and abs(r[i] - host) <= 1e-6 * host, (i, r[i], host)


def test_batch_equilibrium_warm_start_equals_single_lattice():
    docs = _small_batch(count=4, seed=13, storage="f32")
    with pa.Batch(docs, storage="f32") as b:
        for i, doc in enumerate(docs):
            H, V = b.shapes[i]
            yy, xx = np.mgrid[0:V - 2, 0:H - 2]
            rho = 1.0 + 0.01 * np.sin(xx / 7.0) * np.cos(yy / 5.0)
            jx = 0.02 * np.cos(yy / 3.0)
            jy = 0.005 * np.sin(xx / 4.0)
            b.set_equilibrium(i, rho, jx, jy)
            with pa.Lattice.from_settings(doc) as lat:
                lat.set_equilibrium_field(rho, jx, jy)
                lat.run(50)
                expect = lat.macros()
            b.run(50, "persistent")  # advances every domain; only domain i is compared at this point
            got = b.macros(i)
            b.init()
            for a, e in zip(got, expect):
                assert np.array_equal(a, e), i


def test_reset_domain_restarts_one_domain_from_rest():
    docs = _small_batch(count=5, seed=17, storage="f32")
    with pa.Batch(docs) as fresh:
        fresh.run(60, "persistent")
        expect = [fresh.macros(i) for i in range(len(docs))]
    with pa.Batch(docs) as b:
        b.run(101, "persistent")  # odd: the reset lands on the other lattice of the pair
        before = [b.macros(i) for i in range(len(docs))]
        b.reset_domain(2)
        b.run(60, "persistent")
        for a, e in zip(b.macros(2), expect[2]):
            assert np.array_equal(a, e)
        for i in (0, 1, 3, 4):  # untouched domains kept stepping
            assert not np.array_equal(b.macros(i)[1], before[i][1])
