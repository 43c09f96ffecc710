"""Batched domains (lbm_batch_*): every domain of a batch equals the same
domain run alone -- against the oracle in FP64, bit-exactly against the
single-lattice library in FP32 -- in all three launch modes, across repeated
graph replays and after autotuning."""
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
