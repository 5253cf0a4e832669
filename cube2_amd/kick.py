"""Host-side mirror of the kick driver's short-range part (kick.f90:76-236).

``ShortRangeKick`` holds one image's layout on the device and runs
level selection -> PM3 -> PP through the library's C ABI, the way
kick.f90 calls c_pm3_force_kernel then c_pp_force_kernel, and derives the
time-step limits dt_pm3, dt_pp (kick.f90:180,227) and dt_vmax
(kick.f90:238) in FP64.  PM2 (kick.f90:76-105, a CPU loop over tiles in the
reference) runs on the same mesh kernels when asked for.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .geometry import APM3, APM3F, RATIO_SF, Geometry, ixyz3
from .layout import Layout

G_GRID = 1.0 / 6.0 / math.pi   # variables.f90:11
VBUF = 0.9                     # parameters.f90:83


def _np_ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def select_levels(geom: Geometry, rhoc: np.ndarray, ix: np.ndarray | None = None):
    """ires, isort (1-based), napm of kick.f90:107-131."""
    _lib.set_params(geom)
    ix = ixyz3(geom) if ix is None else np.ascontiguousarray(ix, np.int32)
    ires = np.zeros(geom.ns3, np.int32)
    isort = np.zeros(geom.ns3, np.int32)
    napm = np.zeros(7, np.int32)
    rhoc = np.ascontiguousarray(rhoc, np.int32)
    _lib.check(_lib.lib().cube2_select_levels(_np_ptr(rhoc), _np_ptr(ix), _np_ptr(ires),
                                               _np_ptr(isort), _np_ptr(napm)))
    return ires, isort, napm


@dataclass
class KickResult:
    f2max_pm2: float = 0.0
    f2max_pm3: float = 0.0
    f2max_pp: float = 0.0
    vmax: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    ptotal_pm2: np.ndarray | None = None
    ptotal_pm3: np.ndarray | None = None
    ptotal_pp: np.ndarray | None = None
    ttotal_pp: np.ndarray | None = None
    p_cand: int = 0
    p_acc: int = 0
    dt_native: dict | None = None   # dt limits as the library's kick driver computed them
    recorded: bool = False          # kick_graph: this call recorded the graph

    def dt_limits(self, a_mid: float) -> dict:
        """dt_pm2, dt_pm3, dt_pp, dt_vmax in FP64 (kick.f90:97,180,227,238)."""
        def lim(f2):
            return 0.5 * math.sqrt(1.0 / (math.sqrt(f2) * a_mid * G_GRID)) if f2 > 0 else math.inf
        vm = float(np.max(self.vmax))
        return dict(dt_pm2=lim(float(self.f2max_pm2)), dt_pm3=lim(float(self.f2max_pm3)), dt_pp=lim(float(self.f2max_pp)),
                    dt_vmax=VBUF * 14 / vm if vm > 0 else math.inf)


def green_file_name(kind: str, image: int = 1) -> str:
    """Cache file name of a Green's table (initialize.f90:102-175): kind 'Gk2'
    or 'Gk3_<ratio_sf>', then output_suffix() = '_<image>.bin'
    (basic_functions.f08:76-81)."""
    return f"{kind}_{int(image)}.bin"


class ShortRangeKick:
    """One image resident in HBM; each ``kick()`` runs PM3 then PP.

    Green's tables are built on the device on first use.  With
    ``use_green_cache(dir, read)`` they follow the reference's read_Gks switch
    (initialize.f90:102-175): read from ``dir`` when ``read``, else built and
    written there, one stream file per table."""

    green_dir: str | None = None
    read_greens: bool = False
    image: int = 1

    def __init__(self, layout: Layout, device: str = "cuda:0", greens: dict | None = None):
        import torch
        self.torch = torch
        self.geom = layout.geom
        self.layout = layout
        self.dev = torch.device(device)
        _lib.set_params(self.geom)
        self.d_rhoc = torch.from_numpy(np.ascontiguousarray(layout.rhoc)).to(self.dev)
        self.d_idx = torch.from_numpy(np.ascontiguousarray(layout.idx_b_r)).to(self.dev)
        self.d_xp = torch.from_numpy(np.ascontiguousarray(layout.xp)).to(self.dev)
        self.d_vp = torch.from_numpy(np.ascontiguousarray(layout.vp)).to(self.dev)
        self.n_used = layout.n_used
        self.ix = ixyz3(self.geom)
        self.ires, self.isort, self.napm = select_levels(self.geom, layout.rhoc, self.ix)
        self.greens = {} if greens is None else dict(greens)
        torch.cuda.synchronize()

    @classmethod
    def from_device(cls, geom: Geometry, rhoc, idx_b_r, xp, vp, n_used: int,
                    greens: dict | None = None) -> "ShortRangeKick":
        """A kick over an image already in HBM (e.g. from the device layout
        builder), without a host copy of the particle arrays."""
        import torch
        self = cls.__new__(cls)
        self.torch = torch
        self.geom = geom
        self.layout = None
        self.dev = rhoc.device
        _lib.set_params(geom)
        self.d_rhoc, self.d_idx, self.d_xp, self.d_vp = rhoc, idx_b_r, xp, vp
        self.n_used = int(n_used)
        self.ix = ixyz3(geom)
        self.ires, self.isort, self.napm = select_levels(geom, rhoc.cpu().numpy(), self.ix)
        self.greens = {} if greens is None else dict(greens)
        torch.cuda.synchronize()
        return self

    def set_levels(self, ires: np.ndarray) -> None:
        self.ires = np.ascontiguousarray(ires, np.int32)
        self.isort = np.argsort(-self.ires, kind="stable").astype(np.int32) + 1

    def use_green_cache(self, directory: str | None, read: bool = False, image: int = 1) -> None:
        """Tables not yet built are read from (read=True) or written to
        ``directory`` (None: neither), as files named by green_file_name."""
        self.green_dir, self.read_greens, self.image = directory, bool(read), int(image)

    def _green_table(self, kind: str, n: int, soft: float, rng: float, gridsize: float):
        import os
        t = self.torch.empty((n, n, n // 2 + 1), dtype=self.torch.float32, device=self.dev)
        path = None
        if self.green_dir is not None:
            path = os.path.join(self.green_dir, green_file_name(kind, self.image)).encode()
        if path is not None and self.read_greens:
            _lib.check(_lib.lib().cube2_green_file_read_dev(t.data_ptr(), n, path))
            return t
        _lib.check(_lib.lib().cube2_green_3d_dev(t.data_ptr(), n, soft, rng, gridsize))
        if path is not None:
            _lib.check(_lib.lib().cube2_green_file_write_dev(t.data_ptr(), n, path))
        return t

    def green(self, level: int):
        """Device Gk3 table of a level (Gk3_<ratio_sf>, initialize.f90:142-168),
        built once with the FP64 Green kernel (or read from the cache)."""
        if level not in self.greens:
            rsf = int(RATIO_SF[level - 1])
            self.greens[level] = self._green_table(
                f"Gk3_{rsf}", self.geom.nft(level), float(APM3F), float(APM3F * rsf),
                float(np.float32(1.0) / np.float32(rsf)))
        return self.greens[level]

    def green2(self):
        """Device Gk2 table (initialize.f90:135): Green_3D(ngt, apm2 = 3.5,
        apm1 = 3.5 ratio_cs, gridsize 1), built once with the FP64 kernel (or
        read from the cache)."""
        if "pm2" not in self.greens:
            g = self.geom
            self.greens["pm2"] = self._green_table("Gk2", g.ngt, float(APM3F),
                                                   float(APM3F * g.ratio_cs), 1.0)
        return self.greens["pm2"]

    def pm2(self, mass: float, a_mid: float, dt: float, res: KickResult) -> None:
        """PM2 over every tile (kick.f90:76-105); its Delta-v stays pending
        like PM3's until PP or the velocity apply."""
        L = _lib.lib()
        _lib.set_params(self.geom)
        f2 = np.zeros(1, np.float32)
        pt = np.zeros(self.geom.ntile, np.float64)
        _lib.check(L.cube2_pm2_force_dev(
            self.d_rhoc.data_ptr(), self.d_idx.data_ptr(), self.d_xp.data_ptr(), self.n_used,
            self.green2().data_ptr(), float(mass), float(a_mid), float(dt), _np_ptr(f2),
            _np_ptr(pt)))
        res.f2max_pm2 = float(f2[0])
        res.ptotal_pm2 = pt

    def pm3(self, mass: float, a_mid: float, dt: float, res: KickResult) -> None:
        L = _lib.lib()
        _lib.set_params(self.geom)
        used = sorted(set(int(x) for x in self.ires))
        arr = (C.c_void_p * 5)()
        for lev in used:
            arr[lev - 2] = self.green(lev).data_ptr()
        f2 = np.zeros(1, np.float32)
        pt = np.zeros(self.geom.ns3, np.float64)
        _lib.check(L.cube2_pm3_force_dev(
            _np_ptr(self.isort), _np_ptr(self.ires), _np_ptr(self.ix), _np_ptr(RATIO_SF),
            self.d_rhoc.data_ptr(), self.d_idx.data_ptr(), self.d_xp.data_ptr(), self.n_used,
            arr, float(mass), float(a_mid), float(dt), _np_ptr(f2), _np_ptr(pt)))
        res.f2max_pm3 = float(f2[0])
        res.ptotal_pm3 = pt

    def pp(self, mass: float, a_mid: float, dt: float, res: KickResult) -> None:
        L = _lib.lib()
        _lib.set_params(self.geom)
        f2 = np.zeros(1, np.float32)
        vm = np.zeros(3, np.float32)
        pt = np.zeros(self.geom.ns3, np.float64)
        tt = np.zeros(self.geom.ns3, np.float64)
        cnt = np.zeros(2, np.uint64)
        _lib.check(L.cube2_pp_force_dev(
            _np_ptr(self.ires), _np_ptr(self.ix), _np_ptr(APM3), _np_ptr(RATIO_SF),
            self.d_rhoc.data_ptr(), self.d_idx.data_ptr(), self.d_xp.data_ptr(),
            self.d_vp.data_ptr(), self.n_used, float(mass), float(a_mid), float(dt),
            _np_ptr(f2), _np_ptr(vm), _np_ptr(pt), _np_ptr(tt), _np_ptr(cnt)))
        res.f2max_pp = float(f2[0])
        res.vmax = vm
        res.ptotal_pp = pt
        res.ttotal_pp = tt
        res.p_cand, res.p_acc = int(cnt[0]), int(cnt[1])

    def _green_ptrs(self, pm2: bool = False):
        arr = (C.c_void_p * 6)()
        for lev in sorted(set(int(x) for x in self.ires)):
            arr[lev - 2] = self.green(lev).data_ptr()
        if pm2:
            arr[5] = self.green2().data_ptr()
        return arr

    def kick(self, mass: float, a_mid: float, dt: float, pm3: bool = True, pp: bool = True,
             select: bool = False, pm2: bool = False, learned: bool = False,
             apm3: np.ndarray | None = None) -> KickResult:
        """One short-range kick through the library's kick driver
        (cube2_kick_dev).  select=True re-selects levels on the device from
        rhoc (kick.f90:107-131); otherwise the current ires/isort are used.
        pm2=True runs PM2 (kick.f90:76-105) first.  learned=True lets the
        library's loaded network stand in for PP in sparse level-2 cells.
        apm3: the PP ranges per level (default: the reference's table)."""
        apm3 = APM3 if apm3 is None else np.ascontiguousarray(apm3, np.float32)
        L = _lib.lib()
        _lib.set_params(self.geom)
        flags = ((_lib.KICK_PM3 if pm3 else 0) | (_lib.KICK_PP if pp else 0) |
                 (_lib.KICK_PM2 if pm2 else 0) | (_lib.KICK_LEARNED if learned else 0))
        if select:
            self.ires = np.zeros(self.geom.ns3, np.int32)
            self.isort = np.zeros(self.geom.ns3, np.int32)
            # Green tables are bound after selection; select first, then kick
            out = _lib.KickResultC()
            _lib.check(L.cube2_kick_dev(
                0, _np_ptr(self.ix), _np_ptr(APM3), _np_ptr(RATIO_SF), self.d_rhoc.data_ptr(),
                self.d_idx.data_ptr(), self.d_xp.data_ptr(), self.d_vp.data_ptr(), self.n_used,
                None, 1.0, 1.0, 0.0, _np_ptr(self.ires), _np_ptr(self.isort), None, None, None,
                C.byref(out)))
            self.napm = np.array(out.napm[:], np.int32)
        ns3 = self.geom.ns3
        pt3, ptp, ttp = np.zeros(ns3), np.zeros(ns3), np.zeros(ns3)
        out = _lib.KickResultC()
        _lib.check(L.cube2_kick_dev(
            flags | _lib.KICK_KEEP_LEVELS, _np_ptr(self.ix), _np_ptr(apm3), _np_ptr(RATIO_SF),
            self.d_rhoc.data_ptr(), self.d_idx.data_ptr(), self.d_xp.data_ptr(),
            self.d_vp.data_ptr(), self.n_used, self._green_ptrs(pm2), float(mass), float(a_mid),
            float(dt), _np_ptr(self.ires), _np_ptr(self.isort), _np_ptr(pt3), _np_ptr(ptp),
            _np_ptr(ttp), C.byref(out)))
        res = KickResult(f2max_pm2=out.f2max_pm2, f2max_pm3=out.f2max_pm3, f2max_pp=out.f2max_pp,
                         vmax=np.array(out.vmax[:], np.float32),
                         ptotal_pm3=pt3 if pm3 else None, ptotal_pp=ptp if pp else None,
                         ttotal_pp=ttp if pp else None, p_cand=int(out.pair_counts[0]),
                         p_acc=int(out.pair_counts[1]))
        res.dt_native = dict(dt_pm2=out.dt_pm2, dt_pm3=out.dt_pm3, dt_pp=out.dt_pp,
                             dt_vmax=out.dt_vmax)
        return res

    def kick_graph(self, mass: float, a_mid: float, dt: float, pm3: bool = True,
                   learned: bool = False) -> KickResult:
        """The same kick through cube2_kick_graph_dev: recorded into a HIP
        graph on the first call (res.recorded = True) and replayed while the
        levels, arrays and sizes stay the same."""
        L = _lib.lib()
        _lib.set_params(self.geom)
        flags = ((_lib.KICK_PM3 if pm3 else 0) | _lib.KICK_PP |
                 (_lib.KICK_LEARNED if learned else 0))
        ns3 = self.geom.ns3
        pt3, ptp, ttp = np.zeros(ns3), np.zeros(ns3), np.zeros(ns3)
        out = _lib.KickResultC()
        rec = C.c_int(0)
        _lib.check(L.cube2_kick_graph_dev(
            flags, _np_ptr(self.ix), _np_ptr(APM3), _np_ptr(RATIO_SF), self.d_rhoc.data_ptr(),
            self.d_idx.data_ptr(), self.d_xp.data_ptr(), self.d_vp.data_ptr(), self.n_used,
            self._green_ptrs() if pm3 else None, float(mass), float(a_mid), float(dt),
            _np_ptr(self.ires), _np_ptr(self.isort), _np_ptr(pt3), _np_ptr(ptp), _np_ptr(ttp),
            C.byref(out), C.addressof(rec)))
        res = KickResult(f2max_pm3=out.f2max_pm3, f2max_pp=out.f2max_pp,
                         vmax=np.array(out.vmax[:], np.float32),
                         ptotal_pm3=pt3 if pm3 else None, ptotal_pp=ptp, ttotal_pp=ttp,
                         p_cand=int(out.pair_counts[0]), p_acc=int(out.pair_counts[1]))
        res.dt_native = dict(dt_pm3=out.dt_pm3, dt_pp=out.dt_pp, dt_vmax=out.dt_vmax)
        res.recorded = bool(rec.value)
        return res

    def apply_pending(self) -> np.ndarray:
        """Adds a pending mesh Delta-v (PM2/PM3 without PP) to the velocities
        of the physical slots; returns max |vp| per component."""
        vm = np.zeros(3, np.float32)
        _lib.check(_lib.lib().cube2_velocity_apply_dev(
            self.d_rhoc.data_ptr(), self.d_idx.data_ptr(), self.d_vp.data_ptr(), self.n_used,
            _np_ptr(vm)))
        return vm

    def vp_host(self) -> np.ndarray:
        return self.d_vp.cpu().numpy()

    def celllist(self, it: int) -> dict:
        L = _lib.lib()
        rcp, S, N = C.c_int(), C.c_int(), C.c_int()
        _lib.check(L.cube2_pp_debug_celllist(it, C.addressof(rcp), C.addressof(S), C.addressof(N),
                                             None, None, None))
        n, s = N.value, S.value
        keys = np.zeros(max(n, 1), np.uint32)
        slot = np.zeros(max(n, 1), np.uint32)
        start = np.zeros(s ** 3 + 1, np.int32)
        _lib.check(L.cube2_pp_debug_celllist(it, None, None, None, _np_ptr(keys), _np_ptr(slot),
                                             _np_ptr(start)))
        return dict(rcp=rcp.value, S=s, N=n, keys=keys[:n], slot=slot[:n], cell_start=start)
