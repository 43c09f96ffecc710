"""Writes tests/golden/scenario_masks.npz: the obstacle masks of the
reference's own scenario screenshots (README.md:9-11, images/{circle,wing,
walls,skyscraper,hagen-poiseuille}.png), rendered by the reference's GUI.

Solid nodes are drawn black there; a pixel is solid when R + G + B < 40
(the colour map's darkest blue is far above that).  Each mask is stored as
np.packbits of the 767 x 2324 boolean image, row 0 at the top.

The screenshots show the inner nodes of a domain at 2 px per node; of the
sizes tried, horizontalNodes 1164 and verticalNodes 384 (H 1166, V 386
with ghosts) overlap the stencils best, and the test uses that size.

Run from the repository root, in a container that has the reference:
    python tests/golden/make_scenario_masks.py /root/reference
"""
import sys
from pathlib import Path

import numpy as np
from PIL import Image

SCENARIOS = ("circle", "wing", "walls", "skyscraper", "hagen-poiseuille")


def main(reference: str) -> None:
    out = {}
    for name in SCENARIOS:
        rgb = np.asarray(Image.open(Path(reference) / "images" / f"{name}.png").convert("RGB")).astype(np.int32)
        mask = rgb.sum(axis=2) < 40
        out[name] = np.packbits(mask, axis=None)
        out[name + ":shape"] = np.array(mask.shape, dtype=np.int64)
    np.savez_compressed(Path(__file__).with_name("scenario_masks.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
