"""Settings documents of the measured workloads.

* porous_channel: 8192 x 2048 channel of seeded non-overlapping discs
  (radii U[4, 16), porosity 0.7), tau 0.6, inlet 1.05 / outlet 1.0.
* cylinder_channel: 8192 x 2048 channel with one disc of radius 64 at
  x = 1024, centred, inlet 1.02 / outlet 1.0.
* batch_channels: uneven porous channels, nx ~ U{64..1024},
  ny ~ U{32..256}, porosity ~ U[0.5, 0.9], tau ~ U[0.52, 1.0],
  inlet ~ U[1.005, 1.05], outlet 1.0, disc radii U[2, 8), obstacle margin 8,
  at most 20 000 disc candidates per domain.
"""
from __future__ import annotations

import numpy as np

from . import settings


def porous_channel(algorithm: str = "nptl", layout: str = "stream", storage: str = "f64", steps: int = 5000,
                   scenario: str = "porous-circles", inner_h: int = 8192, inner_v: int = 2048,
                   tau: float = 0.6) -> dict:
    wg = 1024 if algorithm in ("npol", "nsol") else 256
    return settings(algorithm, layout, inner_h=inner_h, inner_v=inner_v, scenario=scenario, work_group_size=wg,
                    time_steps=steps, tau=tau, rho_in=1.05, rho_out=1.0, storage=storage, seed=20261202,
                    porosity=0.7, radiusMin=4.0, radiusMax=16.0, obstacleMargin=32)


def cylinder_channel(tau: float, storage: str = "f64", steps: int = 20000, inner_h: int = 8192,
                     inner_v: int = 2048) -> dict:
    return settings("nptl", "stream", inner_h=inner_h, inner_v=inner_v, scenario="cylinder", work_group_size=256,
                    time_steps=steps, tau=tau, rho_in=1.02, rho_out=1.0, storage=storage, cylinderX=1024,
                    cylinderRadius=64.0)


def batch_channels(count: int = 512, seed: int = 4242, steps: int = 2000, storage: str = "f32",
                   nx_range=(64, 1024), ny_range=(32, 256)) -> list[dict]:
    rng = np.random.default_rng(seed)
    docs = []
    for i in range(count):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        docs.append(settings("lptl", "stream", inner_h=nx, inner_v=ny, scenario="porous-circles",
                             work_group_size=256, time_steps=steps, tau=float(rng.uniform(0.52, 1.0)),
                             rho_in=float(rng.uniform(1.005, 1.05)), rho_out=1.0, storage=storage,
                             seed=int(rng.integers(1, 2**62)), porosity=float(rng.uniform(0.5, 0.9)),
                             radiusMin=2.0, radiusMax=8.0, obstacleMargin=8, maxAttempts=20000))
    return docs
