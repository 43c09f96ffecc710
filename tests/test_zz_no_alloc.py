"""lbm_batch_residuals_noalloc (the entry point that allocates nothing, its
partial sums in a static device array) against lbm_batch_residuals, through
the C ABI: bit-equal residuals on two identical batches, over successive
calls, and then on a smaller batch -- the static array keeps the previous
call's contents, which must not leak into a smaller call."""
import numpy as np
import pytest

import portlbm_amd as pa
from portlbm_amd import workloads

pytestmark = pytest.mark.gpu


def _pair(count, seed, storage):
    docs = workloads.batch_channels(count, seed=seed, storage=storage, nx_range=(30, 300), ny_range=(12, 80))
    return pa.Batch(docs, storage=storage), pa.Batch(docs, storage=storage)


def _same(a, b):
    assert a.shape == b.shape
    assert np.array_equal(np.isnan(a), np.isnan(b))
    assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_noalloc_residuals_equal_the_allocating_entry_point(storage):
    big_a, big_b = _pair(24, 11, storage)
    small_a, small_b = _pair(5, 12, storage)
    try:
        for a, b in ((big_a, big_b), (small_a, small_b)):  # the second batch is the smaller input
            for steps in (150, 100, 100):
                a.run(steps, "persistent")
                b.run(steps, "persistent")
                ra, rb = a.residuals(), b.residuals_noalloc()
                print(storage, len(ra), "domains after", steps, "steps:", float(np.nanmax(ra)), flush=True)
                _same(ra, rb)
            assert np.all(ra[np.isfinite(ra)] < 1.0)
    finally:
        for x in (big_a, big_b, small_a, small_b):
            x.close()


def test_noalloc_reports_a_diverged_domain_as_nan():
    docs = workloads.batch_channels(3, seed=5, storage="f32", nx_range=(40, 60), ny_range=(12, 20))
    with pa.Batch(docs) as a, pa.Batch(docs) as b:
        H, V = a.shapes[1]
        bad = np.full((V - 2) * (H - 2), np.nan)
        for x in (a, b):
            x.run(20, "persistent")
            x.set_equilibrium(1, bad, bad, bad)
            x.run(5, "persistent")
        ra, rb = a.residuals(), b.residuals_noalloc()
        assert np.isnan(ra[1]) and np.isnan(rb[1])
        _same(ra, rb)
