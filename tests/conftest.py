"""Shared test setup: markers, repo on sys.path, golden-vector loaders."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


def load_pp_case(name: str) -> dict:
    z = np.load(GOLDEN / f"{name}.npz")
    shape = tuple(int(s) for s in z["rhoc_shape"])
    rhoc = np.zeros(int(np.prod(shape)), np.int32)
    rhoc[z["rhoc_idx"]] = z["rhoc_val"]
    vp_out = z["vp_in"].copy()
    vp_out[z["vp_out_idx"]] = z["vp_out_val"]
    return dict(rhoc=rhoc.reshape(shape), idx_b_r=z["idx_b_r"], xp=z["xp"], vp_in=z["vp_in"],
                mass=float(z["mass"]), a_mid=float(z["a_mid"]), dt=float(z["dt"]),
                vp_out=vp_out, f2max=float(z["f2max"][0]), vmax=z["vmax"],
                ptotal_tile=z["ptotal_tile"])


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle
