"""Times the reference's own pp_force_kernel.c (oracle/_ref, compiled from
the reference sources where they lie) against the oracle's restatement of it
on identical inputs, at the reference's compiled geometry (ng=512, nnt=4,
nns=4, ratio_cs=4, ngb=16) with uniform ires = 2, and checks that both give
the same velocities.  Test infrastructure: run in the container that has the
reference (python oracle/ref_ratio.py [n_particles]).
"""
from __future__ import annotations

import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

import oracle  # noqa: E402
from cube2_amd.geometry import APM3, RATIO_SF, REFERENCE_DEFAULT as G, ixyz3  # noqa: E402
from cube2_amd.layout import build_layout  # noqa: E402


def main(n: int = 400_000) -> None:
    rng = np.random.default_rng(7)
    # Plummer-like clumps plus a uniform background
    centres = rng.uniform(0, 512, (40, 3))
    pos = np.concatenate([centres[rng.integers(0, 40, n // 2)] + rng.normal(0, 1.5, (n // 2, 3)),
                          rng.uniform(0, 512, (n - n // 2, 3))]) % 512
    lay = build_layout(pos, None, G)
    ix = ixyz3(G)
    ires = np.full(G.ns3, 2, np.int32)
    isort = np.arange(1, G.ns3 + 1, dtype=np.int32)
    mass, amid, dt = 1.0, 0.02, 1e-3
    R = oracle.ref()
    vp_ref = lay.vp.copy()
    f2 = np.zeros(1, np.float32)
    vm = np.zeros(3, np.float32)
    pt = np.zeros(G.ns3)
    tt = np.zeros(G.ns3)
    t0 = time.perf_counter()
    R.c_pp_force_kernel(isort, ires, ix, APM3, RATIO_SF, lay.rhoc, lay.idx_b_r, lay.xp, vp_ref,
                        mass, amid, dt, f2, vm, pt, tt)
    t_ref = time.perf_counter() - t0
    t0 = time.perf_counter()
    o = oracle.pp_force(G, ires, ix, APM3, RATIO_SF, lay.rhoc, lay.idx_b_r, lay.xp, lay.vp, mass,
                        amid, dt, nteam=4, nnest=8)
    t_or = time.perf_counter() - t0
    same = np.array_equal(o["vp"].view(np.uint32), vp_ref.view(np.uint32))
    print(f"particles {lay.n_particles}, slots {lay.n_used}, host cores {os.cpu_count()}, "
          f"OpenMP nteam=4 x nnest=8")
    print(f"reference pp_force_kernel.c: {t_ref:.3f} s; oracle restatement: {t_or:.3f} s; "
          f"ratio oracle/reference {t_or / t_ref:.3f}; velocities bit-identical: {same}; "
          f"P_acc {o['p_acc']}")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 400_000)
