"""Green's table cache files on the device (initialize.f90:102-175): a
written table is the reference's stream file of real(4) Gk(n/2+1,n,n) in
Fortran order (checked against the oracle's Green_3D read that way), a kick
with tables read back from the files equals one with freshly built tables,
and a short file is refused."""
from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    import torch
    assert torch.cuda.is_available()
    from cube2_amd import _lib
    from cube2_amd.geometry import Geometry
    from cube2_amd.layout import build_layout
    _lib.lib()
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, app=0.05)
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, g.ng, (20000, 3))
    for c in rng.uniform(0, g.ng, (3, 3)):
        pos = np.concatenate([pos, (c + rng.normal(0, 0.8, (4000, 3))) % g.ng])
    vel = rng.normal(0, 1, pos.shape).astype(np.float32)
    return g, build_layout(pos, vel, g)


def test_written_table_is_the_reference_stream_file(small, oracle_lib, tmp_path):
    from cube2_amd.geometry import APM3F, RATIO_SF
    from cube2_amd.kick import ShortRangeKick, green_file_name
    g, lay = small
    k = ShortRangeKick(lay)
    k.use_green_cache(str(tmp_path), read=False)
    L = 2
    rsf = int(RATIO_SF[L - 1])
    t = k.green(L).cpu().numpy()
    n, nh = g.nft(L), g.nft(L) // 2 + 1
    path = tmp_path / green_file_name(f"Gk3_{rsf}")
    print(f"Gk3 level {L}: n={n}, {path.name}, {os.path.getsize(path)} bytes", flush=True)
    assert os.path.getsize(path) == 4 * n * n * nh
    fort = np.fromfile(path, np.float32).reshape((nh, n, n), order="F")   # Gk(kx, ky, kz)
    want = oracle_lib.green_3d(n, float(APM3F), float(APM3F * rsf),
                               float(np.float32(1) / np.float32(rsf)))   # [kz][ky][kx]
    assert np.allclose(fort, want.transpose(2, 1, 0), rtol=2.5e-7, atol=0)
    assert np.array_equal(fort.transpose(2, 1, 0), t)


def test_kick_with_tables_read_back_equals_built_tables(small, tmp_path):
    from cube2_amd.kick import ShortRangeKick
    g, lay = small
    a = ShortRangeKick(lay)
    a.use_green_cache(str(tmp_path), read=False)
    ra = a.kick(1.0, 0.02, 1e-3, pm2=True)
    written = sorted(os.listdir(tmp_path))
    print("written:", written, flush=True)
    assert "Gk2_1.bin" in written and any(f.startswith("Gk3_") for f in written)
    b = ShortRangeKick(lay)
    b.use_green_cache(str(tmp_path), read=True)
    rb = b.kick(1.0, 0.02, 1e-3, pm2=True)
    for key in a.greens:
        assert np.array_equal(a.greens[key].cpu().numpy(), b.greens[key].cpu().numpy()), key
    assert np.array_equal(a.vp_host(), b.vp_host())
    assert (ra.f2max_pm3, ra.f2max_pm2) == (rb.f2max_pm3, rb.f2max_pm2)


def test_short_cache_file_is_refused(small, tmp_path):
    from cube2_amd import _lib
    from cube2_amd.kick import ShortRangeKick, green_file_name
    g, lay = small
    np.zeros(10, np.float32).tofile(tmp_path / green_file_name("Gk2"))
    k = ShortRangeKick(lay)
    k.use_green_cache(str(tmp_path), read=True)
    with pytest.raises(_lib.Cube2Error) as e:
        k.green2()
    assert e.value.code == _lib.ERR_ARG
