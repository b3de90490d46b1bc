"""Loads tests/golden/fixtures.npz (written by make_golden.py) into oracle
plane structs and capi plane descriptions."""
from __future__ import annotations

from pathlib import Path

import numpy as np

PATH = Path(__file__).resolve().parent / "fixtures.npz"


def load():
    with np.load(PATH, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def plane(O, fx, prefix):
    """The oracle FgoPlane of fixture plane `prefix` (c1, c2, c3)."""
    u = fx[f"{prefix}_plane"]
    f = fx[f"{prefix}_floats"]
    p = O.FgoPlane()
    p.in_w, p.in_h, p.out_w, p.out_h, p.n_samples, p.dist_kind = (int(v) for v in u)
    p.seed = int(fx[f"{prefix}_seed"][0])
    p.zoom, p.delta, p.rm, p.radius_mean, p.log_mu, p.log_sigma = (float(v) for v in f)
    return p


def desc(fx, prefix):
    from film_grain_amd import capi
    u = fx[f"{prefix}_plane"]
    f = fx[f"{prefix}_floats"]
    return capi.make_desc(int(u[0]), int(u[1]), int(u[2]), int(u[3]), int(u[4]), int(u[5]),
                          int(fx[f"{prefix}_seed"][0]), *(float(v) for v in f))
