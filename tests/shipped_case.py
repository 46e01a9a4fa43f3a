"""FreeLB as its benchmarks build it (GCC's default floating-point contraction,
-march=native) on the halfway 100^3 cavity after 1000 steps, against rho / u
of another run on the fixture's sampled cells. The bit-exact fixtures come
from a contraction-off build; this one states how far the shipped build's
rounding moves the fields and holds the bar of a relative error of at most
1e-5 on rho and on u (relative to the largest |u|)."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PATH = os.path.join(GOLDEN, "cavity_d3q19_hw_100x100x100_1000_shipped.npz")
BAR = 1e-5


def errors(rhou: np.ndarray) -> dict:
    """rhou: [4][102^3] (rho, u) of a run of the cavity. Returns the maximum
    relative errors over the fixture's cells: rho against |rho|, u against
    the largest |u| of the shipped run, and u pointwise against
    max(|u|, 1e-3 max |u|)."""
    fx = np.load(PATH)
    cells = fx["cells"]
    ref = fx["rhou"].astype(np.float64)
    got = np.asarray(rhou)[:, cells].astype(np.float64)
    stats = json.loads(str(fx["stats_vs_contraction_off"]))
    umax = stats["u_max"]
    du = np.abs(got[1:] - ref[1:])
    return {"rho_max_rel": float((np.abs(got[0] - ref[0]) / np.abs(ref[0])).max()),
            "u_max_over_umax": float(du.max() / umax),
            "u_pointwise_floor1e-3": float((du / np.maximum(np.abs(ref[1:]), 1e-3 * umax)).max()),
            "cells": int(cells.size), "full_field_shipped_vs_contraction_off": stats}
