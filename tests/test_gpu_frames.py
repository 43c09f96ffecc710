"""lbm_run_frames copies the results of every frame step to pinned host
memory on a second stream while the steps go on (the reference's
non-blocking frame copies, nptl.hpp:71-75).  The run itself must equal
lbm_run bit for bit, and the frame handed out must be the fields of the last
frame step."""
import numpy as np
import pytest

import portlbm_amd as pa

pytestmark = pytest.mark.gpu


def _doc(alg, layout, storage, rounding="nearest"):
    wg = 64 if alg in ("npol", "nsol") else 128
    return pa.settings(alg, layout, inner_h=150, inner_v=50, scenario="porous-circles", work_group_size=wg,
                       time_steps=1, tau=0.7, rho_in=1.05, storage=storage, rounding=rounding, radiusMin=2.0,
                       radiusMax=6.0, obstacleMargin=4)


@pytest.mark.parametrize("alg,layout,storage,rounding", [
    ("nptl", "stream", "f64", "nearest"), ("npol", "collision", "f32", "nearest"),
    ("nsol", "bundle", "f16-shifted", "stochastic"), ("lptl", "stream", "bf16-shifted", "nearest")])
def test_frames_equal_the_fields_of_the_frame_step(alg, layout, storage, rounding):
    doc = _doc(alg, layout, storage, rounding)
    with pa.Lattice.from_settings(doc) as framed, pa.Lattice.from_settings(doc) as plain, \
            pa.Lattice.from_settings(doc) as at_frame:
        with pytest.raises(pa.LbmError, match="no frame"):
            framed.frame()
        framed.run_frames(7, 10)  # iterations 0..6: no frame step yet
        with pytest.raises(pa.LbmError, match="no frame"):
            framed.frame()
        framed.run_frames(88, 10)  # iterations 7..94: frame steps 9, 19, ..., 89
        plain.run(95, 10)
        np.testing.assert_array_equal(framed.populations(), plain.populations())
        for a, b in zip(framed.macros(), plain.macros()):
            np.testing.assert_array_equal(a, b)
        it, fields = framed.frame()
        assert it == 89
        at_frame.run(90)
        for a, b in zip(fields, at_frame.macros()):
            np.testing.assert_array_equal(a, b)
        # a frame on the last step of a call, then a call without frames keeps it
        framed.run_frames(5, 10)  # iterations 95..99, frame step 99
        it, fields = framed.frame()
        assert it == 99
        for a, b in zip(fields, framed.macros()):
            np.testing.assert_array_equal(a, b)
        framed.run_frames(3, 0)
        assert framed.frame()[0] == 99


def test_handler_frames_reach_the_host_vectors():
    """The handler with a frame every 10 steps: after the run the host vectors
    hold the final fields, equal to the kernel-level run."""
    doc = _doc("nptl", "stream", "f64")
    doc["algorithmic"]["timeSteps"] = 1000
    doc["algorithmic"]["frameUpdateInterval"] = 10
    with pa.Handler(doc) as h, pa.Lattice.from_settings(doc) as lat:
        h.initialize()
        h.start()
        h.block_until_finished()
        lat.run(1000, 10)
        rho, jx, jy, jabs = lat.macros()
        np.testing.assert_array_equal(h.get_densities(), rho)
        np.testing.assert_array_equal(h.get_x_velocities(), jx)
        np.testing.assert_array_equal(h.get_y_velocities(), jy)
