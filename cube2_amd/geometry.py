"""Runtime geometry of one CUBE2 image.

The reference bakes these numbers into Fortran ``parameter`` statements
(parameters.f90:19-28,42,57,77-78) and into the generated C++ header
``cuda_parameters_generated.h`` (tools/gen_cuda_parameters.py:9-21 lists the
eleven names).  Here they are a runtime value handed to the native library
through ``cube2_set_params`` so that one build serves every configuration.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, replace, asdict
from pathlib import Path

import numpy as np

# parameters.f90:22 -- standard cells per fine cell at each PM3/PP level.
RATIO_SF = np.array([1, 2, 4, 6, 8, 12, 16], dtype=np.int32)
# parameters.f90:39 -- PP range per level, real(4) 3.5/ratio_sf.
APM3 = (np.float32(3.5) / RATIO_SF.astype(np.float32)).astype(np.float32)
# parameters.f90:38 -- PM3 softening (fine cells); apm2f = 3.5*ratio_sf is the PM3 range.
APM3F = np.float32(3.5)


@dataclass(frozen=True)
class Geometry:
    """The eleven generator names plus derived sizes (variables.f90, parameters.f90)."""

    ng: int = 512
    nnt: int = 4
    nns: int = 4
    ratio_cs: int = 4
    np_nc: int = 4
    ngb: int = 16
    ncore: int = 32
    body_centered_cubic: bool = False
    image_buffer: float = 2.0
    tile_buffer: float = 3.0
    app: float = 0.06

    # ---- derived (parameters.f90:53-78, variables.f90:32) ----
    @property
    def nc(self) -> int:
        return self.ng // self.ratio_cs

    @property
    def nt(self) -> int:
        return self.nc // self.nnt

    @property
    def ntt(self) -> int:
        return self.nt // self.nns

    @property
    def ncb(self) -> int:
        return self.ngb // self.ratio_cs

    @property
    def nte(self) -> int:
        return self.nt + 2 * self.ncb

    @property
    def nsub(self) -> int:
        return self.nnt * self.nns

    @property
    def ns3(self) -> int:
        return self.nsub ** 3

    @property
    def ntile(self) -> int:
        return self.nnt ** 3

    def nfp(self, level: int) -> int:
        """Physical fine cells per subtile edge at level (2..6)."""
        return self.ntt * self.ratio_cs * int(RATIO_SF[level - 1])

    def nfb(self, level: int) -> int:
        return self.ratio_cs * int(RATIO_SF[level - 1])

    def nft(self, level: int) -> int:
        """PM3 FFT size n = nfp + 2 nfb (parameters.f90:62-65)."""
        return self.nfp(level) + 2 * self.nfb(level)

    @property
    def ngp(self) -> int:
        """Standard cells per tile edge (parameters.f90:54)."""
        return self.nt * self.ratio_cs

    @property
    def ngt(self) -> int:
        """PM2 FFT size ngp + 2 ngb (parameters.f90:58)."""
        return self.ngp + 2 * self.ngb

    def validate(self) -> None:
        if self.ng % self.ratio_cs or self.nc % self.nnt or self.nt % self.nns:
            raise ValueError(f"inconsistent geometry {self}")
        if self.ngb % self.ratio_cs or self.ncb < 1:
            raise ValueError("ngb must be a positive multiple of ratio_cs")

    def with_(self, **kw) -> "Geometry":
        return replace(self, **kw)

    def as_dict(self) -> dict:
        return asdict(self)


def ixyz3(geom: Geometry) -> np.ndarray:
    """Subtile table of initialize.f90:207-224: rows (sx,sy,sz,tx,ty,tz), 1-based."""
    t = np.arange(geom.nnt)
    s = np.arange(geom.nns)
    tz, ty, tx, sz, sy, sx = np.meshgrid(t, t, t, s, s, s, indexing="ij")
    out = np.stack([sx, sy, sz, tx, ty, tz], axis=-1).reshape(-1, 6) + 1
    return np.ascontiguousarray(out.astype(np.int32))


def parse_parameters_f90(path: str | Path) -> Geometry:
    """Read the eleven names from a parameters.f90 (same contract as the
    reference generator, tools/gen_cuda_parameters.py:9-21)."""
    names = ("ng", "nnt", "nns", "ratio_cs", "np_nc", "ngb", "ncore",
             "body_centered_cubic", "image_buffer", "tile_buffer", "app")
    raw: dict[str, str] = {}
    for line in Path(path).read_text().splitlines():
        line = line.split("!", 1)[0]
        if "parameter" not in line.lower():
            continue
        for name in names:
            m = re.search(rf"\b{name}\s*=\s*([^,\s]+)", line, re.IGNORECASE)
            if m:
                raw[name] = m.group(1)
    missing = [n for n in names if n not in raw]
    if missing:
        raise ValueError(f"missing parameter(s): {', '.join(missing)}")

    def num(v: str):
        v = v.strip()
        if v.lower() in (".true.", ".false."):
            return v.lower() == ".true."
        if v in raw:
            return num(raw[v])
        v = re.sub(r"_\w+$", "", v).replace("d", "e").replace("D", "e")
        return float(v) if any(c in v for c in ".e") else int(v)

    vals = {n: num(raw[n]) for n in names}
    return Geometry(**{k: (type(getattr(Geometry, k, v))(v) if k != "body_centered_cubic" else bool(v))
                       for k, v in vals.items()})


# ---- the configurations the path is measured on (seeded synthetic inputs) ----
# C0: 64^3 particles, one tile, fine mesh 128^3 at level 2.
C0 = Geometry(ng=64, nnt=1, nns=1, ratio_cs=4, np_nc=4, ngb=16, app=0.05)
# C1: 256^3 particles, fine mesh 512^3 as 8^3 subtiles of 64^3 fine cells.
C1 = Geometry(ng=256, nnt=4, nns=2, ratio_cs=4, np_nc=4, ngb=24, app=0.05)
# C2 (batching): 16^3 = 4096 tiles of 32^3 fine cells at level 2 (16 standard
# cells, 4 coarse cells per tile edge), one subtile per tile, a one-coarse-
# cell tile buffer (all PM3 and PP read; PM2 is not in this kick).
C2 = Geometry(ng=256, nnt=16, nns=1, ratio_cs=4, np_nc=4, ngb=4, app=0.05)
# The reference's own compiled default (cuda_parameters_generated.h).
REFERENCE_DEFAULT = Geometry()
