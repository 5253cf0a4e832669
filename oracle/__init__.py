"""ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded only by tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of bench.py, always as the checker.  The product library never imports it.
``ref()`` exposes the reference's own pp_force_kernel.c, compiled from
/root/reference by ``make -C oracle ref`` into oracle/_ref (container only).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None
_REF = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
i16p = np.ctypeslib.ndpointer(np.int16, flags="C")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")


class OrGeom(C.Structure):
    _fields_ = [("ng", C.c_int), ("nnt", C.c_int), ("nns", C.c_int), ("ratio_cs", C.c_int),
                ("ngb", C.c_int), ("nteam", C.c_int), ("nnest", C.c_int), ("app", C.c_float)]


def build(force: bool = False) -> None:
    if force or not (HERE / "liboracle.so").exists():
        subprocess.run(["make", "-C", str(HERE), "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        L.or_F_ra.restype = C.c_float
        L.or_F_ra.argtypes = [C.c_float, C.c_float]
        L.or_pair_forces.argtypes = [f32p, f32p, C.c_float, C.c_int, f32p]
        L.or_stable_sort_u32_pair.argtypes = [u32p, u32p, C.c_int, C.c_uint32]
        L.or_pp_tile_sizes.argtypes = [C.POINTER(OrGeom), i32p, i32p, i32p, i32p, C.c_int,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_pp_tile_celllist.argtypes = [C.POINTER(OrGeom), i32p, i32p, f32p, i32p, i32p, i64p,
                                          i16p, C.c_int, u32p, u32p, u32p, i32p, f32p, f32p]
        L.or_pp_force.restype = C.c_int
        L.or_pp_force.argtypes = [C.POINTER(OrGeom), i32p, i32p, f32p, i32p, i32p, i64p, i16p,
                                  f32p, C.c_void_p, C.c_float, C.c_float, C.c_float, f32p, f32p,
                                  f64p, f64p, u64p, C.c_int, C.c_int]
        L.or_select_levels.argtypes = [C.POINTER(OrGeom), i32p, i32p, i32p, i32p, i32p, f32p]
        L.or_indexed_sort.argtypes = [C.c_int, f32p, i32p]
        L.or_fft3d_r2c.argtypes = [f32p, f64p, C.c_int]
        L.or_fft3d_c2r.argtypes = [f64p, f64p, C.c_int]
        L.or_fft1d.argtypes = [f64p, C.c_int, C.c_int]
        L.or_potential_to_force.argtypes = [f64p, C.c_int, C.c_int, C.c_int, f64p,
                                            C.POINTER(C.c_double)]
        L.or_green_3d.argtypes = [f32p, C.c_int, C.c_float, C.c_float, C.c_float]
        L.or_pm3_job.restype = C.c_int
        L.or_pm3_job.argtypes = [C.POINTER(OrGeom), C.c_int, i32p, i32p, i32p, i32p, i64p, i16p,
                                 f32p, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.POINTER(C.c_double)]
        L.or_pm3_force.restype = C.c_int
        L.or_pm3_force.argtypes = [C.POINTER(OrGeom), i32p, i32p, i32p, i32p, i32p, i64p, i16p,
                                   C.POINTER(C.c_void_p), C.c_float, C.c_float, C.c_float, f64p,
                                   f32p, f64p]
        L.or_pm2_force.restype = C.c_int
        L.or_pm2_force.argtypes = [C.POINTER(OrGeom), i32p, i64p, i16p, f32p, C.c_float, C.c_float,
                                   C.c_float, f64p, f32p, f64p, C.c_int, C.c_int]
        L.or_drift.restype = C.c_int
        L.or_drift.argtypes = [C.POINTER(OrGeom), i32p, i64p, i16p, f32p, C.c_float, C.c_float,
                               i32p, i64p, i16p, f32p, C.c_int64, C.POINTER(C.c_int64)]
        L.or_velocity_apply.argtypes = [C.POINTER(OrGeom), i32p, i64p, f32p, f64p, f32p]
        L.or_direct_pp_f64_cells.argtypes = [f64p, C.c_int, C.c_double, i32p, C.c_int, f64p,
                                             C.c_double, C.c_int, f64p]
        L.or_direct_pp_f64.argtypes = [f64p, C.c_int, C.c_double, i32p, C.c_int, f64p, C.c_double,
                                       C.c_int, f64p]
        _LIB = L
    return _LIB


def geom_struct(geom, nteam: int = 4, nnest: int | None = None) -> OrGeom:
    if nnest is None:
        nnest = max(1, (os.cpu_count() or 8) // nteam)
    return OrGeom(geom.ng, geom.nnt, geom.nns, geom.ratio_cs, geom.ngb, nteam, nnest,
                  float(geom.app))


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def F_ra(r, apm):
    L = lib()
    r = np.asarray(r, dtype=np.float32).reshape(-1)
    a = np.broadcast_to(np.asarray(apm, dtype=np.float32), r.shape)
    return np.array([L.or_F_ra(float(x), float(y)) for x, y in zip(r, a)], dtype=np.float32)


def pair_forces(d, ranges, app):
    """(F_ra(r,app) - F_ra(r,range)) d / r per pair, the reference's FP32 rounding."""
    d = _c(d, np.float32).reshape(-1, 3)
    rg = _c(ranges, np.float32).reshape(-1)
    out = np.zeros_like(d)
    lib().or_pair_forces(d, rg, float(app), d.shape[0], out)
    return out


def stable_sort(keys, vals, bound):
    k = _c(keys, np.uint32).copy()
    v = _c(vals, np.uint32).copy()
    lib().or_stable_sort_u32_pair(k, v, k.size, int(bound))
    return k, v


def pp_force(geom, ires, ixyz3, apm3, ratio_sf, rhoc, idx_b_r, xp, vp, mass, a_mid, dt,
             dv64=None, tiles=None, nteam=4, nnest=None):
    """Returns dict(vp, f2max, vmax, ptotal, ttotal, p_cand, p_acc); vp updated copy."""
    L = lib()
    g = geom_struct(geom, nteam, nnest)
    vp2 = _c(vp, np.float32).copy()
    f2 = np.zeros(1, np.float32)
    vm = np.zeros(3, np.float32)
    pt = np.zeros(geom.ns3, np.float64)
    tt = np.zeros(geom.ns3, np.float64)
    cnt = np.zeros(2, np.uint64)
    t0, tn = (0, -1) if tiles is None else tiles
    dvp = None
    if dv64 is not None:
        dv64 = _c(dv64, np.float64)
        dvp = dv64.ctypes.data
    rc = L.or_pp_force(C.byref(g), _c(ires, np.int32), _c(ixyz3, np.int32), _c(apm3, np.float32),
                       _c(ratio_sf, np.int32), _c(rhoc, np.int32), _c(idx_b_r, np.int64),
                       _c(xp, np.int16), vp2, dvp, float(mass), float(a_mid), float(dt), f2, vm,
                       pt, tt, cnt, int(t0), int(tn))
    if rc != 0:
        raise ValueError("oracle rejected the PP inputs")
    return dict(vp=vp2, f2max=float(f2[0]), vmax=vm, ptotal=pt, ttotal=tt,
                p_cand=int(cnt[0]), p_acc=int(cnt[1]))


def pp_tile_celllist(geom, ires, ixyz3, apm3, ratio_sf, rhoc, idx_b_r, xp, it):
    L = lib()
    g = geom_struct(geom)
    rcp, S, N = C.c_int(), C.c_int(), C.c_int()
    ires = _c(ires, np.int32)
    ixyz3 = _c(ixyz3, np.int32)
    ratio_sf = _c(ratio_sf, np.int32)
    rhoc = _c(rhoc, np.int32)
    L.or_pp_tile_sizes(C.byref(g), ires, ixyz3, ratio_sf, rhoc, int(it), C.byref(rcp),
                       C.byref(S), C.byref(N))
    n, s = N.value, S.value
    keys = np.zeros(max(n, 1), np.uint32)
    perm = np.zeros(max(n, 1), np.uint32)
    slot = np.zeros(max(n, 1), np.uint32)
    start = np.zeros(s ** 3 + 1, np.int32)
    xf = np.zeros(max(3 * n, 1), np.float32)
    rg = np.zeros(max(n, 1), np.float32)
    L.or_pp_tile_celllist(C.byref(g), ires, ixyz3, _c(apm3, np.float32), ratio_sf, rhoc,
                          _c(idx_b_r, np.int64), _c(xp, np.int16), int(it), keys, perm, slot,
                          start, xf, rg)
    return dict(rcp=rcp.value, S=s, N=n, keys=keys[:n], perm=perm[:n], slot=slot[:n],
                cell_start=start, xf=xf[:3 * n].reshape(3, n), range=rg[:n])


def select_levels(geom, rhoc, ixyz3):
    L = lib()
    g = geom_struct(geom)
    ires = np.zeros(geom.ns3, np.int32)
    isort = np.zeros(geom.ns3, np.int32)
    napm = np.zeros(7, np.int32)
    npc = np.zeros(geom.ns3, np.float32)
    L.or_select_levels(C.byref(g), _c(rhoc, np.int32), _c(ixyz3, np.int32), ires, isort, napm, npc)
    return ires, isort, napm, npc


def indexed_sort(arr):
    a = _c(arr, np.float32)
    idx = np.zeros(a.size, np.int32)
    lib().or_indexed_sort(a.size, a, idx)
    return idx


def fft3d_r2c(x):
    n = x.shape[0]
    out = np.zeros((n, n, n // 2 + 1), np.complex128)
    lib().or_fft3d_r2c(_c(x, np.float32), out.view(np.float64), n)
    return out


def fft3d_c2r(c):
    n = c.shape[0]
    out = np.zeros((n, n, n), np.float64)
    lib().or_fft3d_c2r(_c(c, np.complex128).view(np.float64), out, n)
    return out


def fft1d(z, sign=-1):
    a = _c(z, np.complex128).copy()
    lib().or_fft1d(a.view(np.float64), a.size, int(sign))
    return a


def green_3d(n, apm_soft, apm_range, gridsize):
    g = np.zeros((n, n, n // 2 + 1), np.float32)
    lib().or_green_3d(g, int(n), float(apm_soft), float(apm_range), float(gridsize))
    return g


def potential_to_force(phi, nfb, nfp):
    """potential_to_force of an (n, n, n) potential [z][y][x]: returns
    (F[3][nfp+2]^3, f2max over the interior)."""
    phi = _c(phi, np.float64)
    n = phi.shape[0]
    nf = nfp + 2
    F = np.zeros((3, nf, nf, nf), np.float64)
    f2 = C.c_double()
    lib().or_potential_to_force(phi.reshape(-1), n, int(nfb), int(nfp), F.reshape(-1), C.byref(f2))
    return F, f2.value


def pm3_job(geom, it, ires, ixyz3, ratio_sf, rhoc, idx_b_r, xp, green, mass,
            want_rho=True, want_phi=True, want_force=True):
    L = lib()
    g = geom_struct(geom)
    level = int(ires[it])
    n = geom.nft(level)
    nf = geom.nfp(level) + 2
    rho = np.zeros((n, n, n), np.float32) if want_rho else None
    phi = np.zeros((n, n, n), np.float64) if want_phi else None
    force = np.zeros((3, nf, nf, nf), np.float64) if want_force else None
    f2 = C.c_double()
    rc = L.or_pm3_job(C.byref(g), int(it), _c(ires, np.int32), _c(ixyz3, np.int32),
                      _c(ratio_sf, np.int32), _c(rhoc, np.int32), _c(idx_b_r, np.int64),
                      _c(xp, np.int16), _c(green, np.float32), float(mass),
                      None if rho is None else rho.ctypes.data,
                      None if phi is None else phi.ctypes.data,
                      None if force is None else force.ctypes.data, C.byref(f2))
    if rc:
        raise ValueError("oracle rejected the PM3 job")
    return dict(rho=rho, phi=phi, force=force, f2max=f2.value)


def pm3_force(geom, isort, ires, ixyz3, ratio_sf, rhoc, idx_b_r, xp, greens, mass, amid, dt):
    """greens: dict level->float32 (n,n,n/2+1).  Returns dict(dv64, f2max, ptotal)."""
    L = lib()
    g = geom_struct(geom)
    n_used = int(np.asarray(idx_b_r).reshape(-1)[-1])
    dv = np.zeros((n_used, 3), np.float64)
    f2 = np.zeros(1, np.float32)
    pt = np.zeros(geom.ns3, np.float64)
    keep = {L_: _c(v, np.float32) for L_, v in greens.items()}
    arr = (C.c_void_p * 7)()
    for L_, v in keep.items():
        arr[L_] = v.ctypes.data
    rc = L.or_pm3_force(C.byref(g), _c(isort, np.int32), _c(ires, np.int32), _c(ixyz3, np.int32),
                        _c(ratio_sf, np.int32), _c(rhoc, np.int32), _c(idx_b_r, np.int64),
                        _c(xp, np.int16), arr, float(mass), float(amid), float(dt), dv, f2, pt)
    if rc:
        raise ValueError("oracle rejected the PM3 inputs")
    return dict(dv64=dv, f2max=float(f2[0]), ptotal=pt)


def pm2_force(geom, rhoc, idx_b_r, xp, green2, mass, amid, dt, tiles=None):
    """green2: float32 (n,n,n/2+1), n = nt*ratio_cs + 2*ngb; tiles: (first,
    count) or None for all.  Returns dict(dv64, f2max, ptotal)."""
    L = lib()
    g = geom_struct(geom)
    n_used = int(np.asarray(idx_b_r).reshape(-1)[-1])
    dv = np.zeros((n_used, 3), np.float64)
    f2 = np.zeros(1, np.float32)
    pt = np.zeros(geom.nnt ** 3, np.float64)
    rc = L.or_pm2_force(C.byref(g), _c(rhoc, np.int32), _c(idx_b_r, np.int64), _c(xp, np.int16),
                        _c(green2, np.float32), float(mass), float(amid), float(dt), dv, f2, pt,
                        *((0, 0) if tiles is None else (int(tiles[0]), int(tiles[1]))))
    if rc:
        raise ValueError("oracle rejected the PM2 inputs")
    return dict(dv64=dv, f2max=float(f2[0]), ptotal=pt)


def drift(geom, rhoc, idx_b_r, xp, vp, dt_mid, vz_max):
    """Returns dict(rhoc, idx_b_r, xp, vp, n) of the drifted physical-only image."""
    L = lib()
    g = geom_struct(geom)
    rhoc = _c(rhoc, np.int32)
    n_in = int(np.asarray(idx_b_r).reshape(-1)[-1])
    ro = np.zeros_like(rhoc)
    io = np.zeros(np.asarray(idx_b_r).shape, np.int64)
    xo = np.zeros((max(n_in, 1), 3), np.int16)
    vo = np.zeros((max(n_in, 1), 3), np.float32)
    n = C.c_int64(0)
    rc = L.or_drift(C.byref(g), rhoc, _c(idx_b_r, np.int64), _c(xp, np.int16), _c(vp, np.float32),
                    float(dt_mid), float(vz_max), ro, io, xo, vo, n_in, C.byref(n))
    if rc:
        raise ValueError(f"oracle drift failed ({rc})")
    return dict(rhoc=ro, idx_b_r=io, xp=xo[:n.value], vp=vo[:n.value], n=n.value)


def velocity_apply(geom, rhoc, idx_b_r, vp, dv64):
    vp2 = _c(vp, np.float32).copy()
    vm = np.zeros(3, np.float32)
    lib().or_velocity_apply(C.byref(geom_struct(geom)), _c(rhoc, np.int32),
                            _c(idx_b_r, np.int64), vp2, _c(dv64, np.float64), vm)
    return vp2, vm


def direct_pp_f64(pos, box, targets, ranges, app, cells=False, law="s2"):
    """FP64 direct sum for the targets; cells=True bins the particles so each
    target scans its 27 neighbour cells (same pairs, other summation order).
    law "s2": F_ra(r, app) - F_ra(r, range) (pp_force_kernel.c:308-318);
    "plummer": r / (r^2 + app^2)^(3/2) - F_ra(r, range)."""
    pos = _c(pos, np.float64)
    targets = _c(targets, np.int32)
    acc = np.zeros((targets.size, 3), np.float64)
    f = lib().or_direct_pp_f64_cells if cells else lib().or_direct_pp_f64
    f(pos, pos.shape[0], float(box), targets, targets.size, _c(ranges, np.float64), float(app),
      {"s2": 0, "plummer": 1}[law], acc)
    return acc


# ---------------------------------------------------------------- reference
def ref_available() -> bool:
    return (HERE / "_ref" / "libppref.so").exists()


def ref():
    """The reference pp_force_kernel.c compiled as shipped (fixed geometry
    ng=512, nnt=4, nns=4, ratio_cs=4, ngb=16, nteam=4, nnest=8, app=0.06)."""
    global _REF
    if _REF is None:
        path = HERE / "_ref" / "libppref.so"
        if not path.exists():
            raise FileNotFoundError("oracle/_ref/libppref.so is not built (make -C oracle ref)")
        R = C.CDLL(str(path))
        R.F_ra.restype = C.c_float
        R.F_ra.argtypes = [C.c_float, C.c_float]
        R.radix_sort_u32_pair.argtypes = [u32p, u32p, C.c_int]
        R.c_pp_force_kernel.argtypes = [i32p, i32p, i32p, f32p, i32p, i32p, i64p, i16p, f32p,
                                        C.c_float, C.c_float, C.c_float, f32p, f32p, f64p, f64p]
        _REF = R
    return _REF
