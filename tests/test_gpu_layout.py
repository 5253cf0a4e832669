"""The device layout builder (cube2_layout_build_dev) against the host one
(cube2_amd/layout.py: build_layout): rhoc, idx_b_r, xp and vp bit-identical,
including positions on cell boundaries, outside [0, ng) and an empty set."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available()
    from cube2_amd import _lib
    _lib.lib()
    return _lib


def _compare(pos, vel, geom):
    from cube2_amd.layout import build_layout, build_layout_device
    host = build_layout(pos, vel, geom)
    rhoc, idx, xp, vp, n_used = build_layout_device(pos, vel, geom)
    assert n_used == host.n_used
    assert np.array_equal(rhoc.cpu().numpy(), host.rhoc)
    assert np.array_equal(idx.cpu().numpy(), host.idx_b_r)
    assert np.array_equal(xp.cpu().numpy(), host.xp)
    assert np.array_equal(vp.cpu().numpy().view(np.uint32), host.vp.view(np.uint32))
    return n_used


def test_layout_c0_lattice(lib):
    from cube2_amd.geometry import C0
    from cube2_amd.synth import config0
    pos, _ = config0(seed=0)
    vel = np.random.default_rng(1).standard_normal(pos.shape).astype(np.float32)
    assert _compare(pos, vel, C0) > pos.shape[0]


def test_layout_multi_tile_clumps_and_edges(lib):
    from cube2_amd.geometry import Geometry
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, app=0.05)
    rng = np.random.default_rng(5)
    pos = np.concatenate([rng.uniform(0, 64, (20000, 3)),
                          np.array([31.5, 30.0, 33.0]) + rng.normal(0, 0.7, (3000, 3)),
                          # cell and tile boundaries, the box edge, outside the box
                          np.array([[0.0, 0.0, 0.0], [4.0, 8.0, 32.0], [63.99999999, 64.0, -0.0],
                                    [-1e-17, 1e-300, 128.0], [-4.0, 67.5, -64.0],
                                    [31.999999999999996, 32.0, 32.000000000000007]])])
    vel = rng.standard_normal(pos.shape).astype(np.float32)
    _compare(pos, vel, g)
    _compare(pos, None, g)


def test_layout_empty(lib):
    from cube2_amd.geometry import Geometry
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, app=0.05)
    assert _compare(np.zeros((0, 3)), None, g) == 0
