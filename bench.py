"""Short-range kick benchmark on one MI355X.

Workload: 256^3 particles in a 300 Mpc/h box (60% Zel'dovich lattice at
z=49, 40% in 2000 Plummer spheres), fine mesh 512^3 as 8^3 subtiles of 64^3
fine cells (nnt=4, nns=2, 6-coarse-cell tile buffer), synthetic and seeded.
One step = one kick: PM3 over all 512 subtiles -> PP over all 64 parent
tiles with the velocity update, inputs resident in HBM.  Timed twice: with
eager dispatch (per-stage HIP-event profile) and as a recorded HIP graph
replay; `value` is the replayed kick.

Prints one JSON line: kick ms/step (value, lower is better) plus PP pair
interactions per second, the roofline of the dominant kernel from HIP events
on the library stream, and a CPU baseline measured on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 vector (packed)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pp_flops(p_cand: int, p_acc: int) -> float:
    """Algorithmic PP work: 8 flops per examined candidate (3 sub, 1 mul,
    2 FMA), 36 per accepted pair (rsqrt, r, 2x13 for F_ra, sub, scale, 3 FMA)."""
    return 8.0 * p_cand + 36.0 * p_acc


def cpu_baseline(lay, ires, isort, greens_host, p_full, mass, amid, dt, budget_s: float):
    """Oracle (C/OpenMP restatement, same two-level structure as the
    reference) on a bounded sample: PP over the lightest parent tiles until
    ~budget_s (extrapolated to the kick by PP flops), PM3 over up to 4
    subtiles of each level (extrapolated per level by subtile count; levels
    past the budget by the last timed level's time per n^3 log2 n)."""
    import oracle
    from cube2_amd.geometry import APM3, RATIO_SF, ixyz3
    g = lay.geom
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    nteam = 4
    nnest = max(1, ncpu // nteam)
    # PP: lightest tiles first (cost proxy: sum of coarse counts squared); the
    # dense tiles would take minutes each on the host
    proxy = (lay.rhoc.reshape(g.ntile, -1).astype(np.float64) ** 2).sum(1)
    order = np.argsort(proxy, kind="stable")
    t_pp, f_pp, tiles, px = 0.0, 0.0, [], 0.0
    for t in order:
        if tiles and t_pp / px * proxy[t] > 2 * (budget_s - t_pp):
            break  # this tile alone would overrun the sample budget
        t0 = time.perf_counter()
        o = oracle.pp_force(g, ires, ixyz3(g), APM3, RATIO_SF, lay.rhoc, lay.idx_b_r, lay.xp,
                            lay.vp, mass, amid, dt, tiles=(int(t), 1), nteam=1,
                            nnest=ncpu)
        t_pp += time.perf_counter() - t0
        f_pp += pp_flops(o["p_cand"], o["p_acc"])
        px += max(proxy[t], 1.0)
        tiles.append(int(t))
        if t_pp >= budget_s * 0.75:
            break
    pp_rate = f_pp / t_pp
    pp_ms = pp_flops(*p_full) / pp_rate * 1e3
    # PM3: up to 4 subtiles of every level present (spread over the level's
    # subtiles), the coarse levels first, within a third of the budget; each
    # level extrapolated by its subtile count
    levels = sorted(set(int(L) for L in ires))
    per_level = {}
    t0 = time.perf_counter()
    for L in levels:
        subs = [i for i in range(g.ns3) if ires[i] == L]
        picks = subs[:: max(1, len(subs) // 4)][:4]
        tl, nl = 0.0, 0
        for it in picks:
            t1 = time.perf_counter()
            oracle.pm3_job(g, it, ires, ixyz3(g), RATIO_SF, lay.rhoc, lay.idx_b_r, lay.xp,
                           greens_host[L], mass, want_rho=False, want_phi=False, want_force=True)
            tl += time.perf_counter() - t1
            nl += 1
            if time.perf_counter() - t0 > budget_s / 3:
                break
        per_level[L] = (tl / nl, len(subs), nl)
        if time.perf_counter() - t0 > budget_s / 3:
            break
    # levels not reached: the last timed level's time per n^3 log2 n
    Lt = max(per_level)
    unit = per_level[Lt][0] / (g.nft(Lt) ** 3 * np.log2(g.nft(Lt)))
    pm3_s = 0.0
    for L in levels:
        cnt = sum(1 for i in range(g.ns3) if ires[i] == L)
        pm3_s += cnt * (per_level[L][0] if L in per_level
                        else unit * g.nft(L) ** 3 * np.log2(g.nft(L)))
    t_pm3 = time.perf_counter() - t0
    pm3_ms = pm3_s * 1e3
    timed = ", ".join(f"L{L}: {v[2]} of {v[1]} ({1e3 * v[0]:.0f} ms each)"
                      for L, v in sorted(per_level.items()))
    return dict(value=round(pp_ms + pm3_ms, 1), unit="ms/step", cores=ncpu, kind="port",
                sample=(f"oracle PP on {len(tiles)} of {g.ntile} parent tiles "
                        f"({t_pp:.1f} s, {pp_rate / 1e9:.2f} GFLOP/s algorithmic, extrapolated "
                        f"by PP flops: {pp_ms:.0f} ms) + oracle PM3 jobs per level ({timed}; "
                        f"{t_pm3:.1f} s, extrapolated by subtile count: {pm3_ms:.0f} ms); "
                        f"OpenMP threads {ncpu}"))


def time_dropin(lay, k, greens_host, mass, amid, dt, steps, exact=False):
    """The Fortran host's two calls per kick through the drop-in C ABI on host
    arrays (kick.f90:144-146 then 202-205): c_pm3_force_kernel uploads the
    positions and leaves the mesh Delta-v pending, c_pp_force_kernel uploads
    the velocities, applies PM3 + PP and downloads them.  Wall ms per kick,
    PCIe transfers included (pageable host memory, as a Fortran caller
    hands it over).  Outside the exact-order mode only the physical
    particles' velocity rows cross PCIe (halo copies are never changed)."""
    import ctypes as C

    import torch

    from cube2_amd import _lib
    from cube2_amd.geometry import APM3, RATIO_SF, ixyz3
    g = lay.geom
    L = _lib.lib()
    _lib.set_params(g)
    ires = np.ascontiguousarray(k.ires, np.int32)
    isort = np.ascontiguousarray(k.isort, np.int32)
    ix = ixyz3(g)
    rhoc = np.ascontiguousarray(lay.rhoc, np.int32)
    idx = np.ascontiguousarray(lay.idx_b_r, np.int64)
    xp = np.ascontiguousarray(lay.xp, np.int16)
    vp = np.ascontiguousarray(lay.vp, np.float32).copy()
    gk = [np.ascontiguousarray(greens_host[Lv], np.float32) if Lv in greens_host
          else np.zeros(1, np.float32) for Lv in range(2, 7)]
    f2 = np.zeros(1, np.float32)
    vm = np.zeros(3, np.float32)
    pt = np.zeros(g.ns3)
    tt = np.zeros(g.ns3)
    a = lambda x: x.ctypes.data  # noqa: E731
    apm3 = np.ascontiguousarray(APM3, np.float32)
    rsf = np.ascontiguousarray(RATIO_SF, np.int32)
    times = []
    for s in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.c_pm3_force_kernel(a(isort), a(ires), a(ix), a(rsf), a(rhoc), a(idx), a(xp),
                             *[a(x) for x in gk], a(vp), mass, amid, dt, a(f2), a(vm), a(pt))
        L.c_pp_force_kernel(a(isort), a(ires), a(ix), a(apm3), a(rsf), a(rhoc), a(idx), a(xp),
                            a(vp), mass, amid, dt, a(f2), a(vm), a(pt), a(tt))
        times.append(1e3 * (time.perf_counter() - t0))
        log(f"drop-in kick {s}: {times[-1]:.1f} ms")
    L.c_pm3_fft_finalize()
    ph = slice(g.ncb, g.ncb + g.nt)
    vp_rows = vp.shape[0] if exact else int(rhoc.reshape(-1, g.nte, g.nte, g.nte)[:, ph, ph, ph].sum())
    return {"ms": round(min(times[1:]), 3), "first_call_ms": round(times[0], 3),
            "vp_rows_over_pcie": vp_rows,
            "bytes_over_pcie_per_kick": int(xp.nbytes + 2 * 12 * vp_rows + 2 * (rhoc.nbytes + idx.nbytes))}


def profiled_traffic(stage: str):
    """HBM bytes per launch of a stage's kernel at this workload, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (None if absent)."""
    f = ROOT / "profiles" / f"{stage}_c1_traffic.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def copy_probe(mib: int) -> int:
    """Three device-to-device copies of a known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_copy": n * 4}), flush=True)
    return 0


def bench_c2(args) -> int:
    """Batching workload C2: 4096 tiles of 32^3 fine cells, per-tile counts
    lognormal (median 4000, sigma 1.5, capped at 200k), laid out on the
    device; the kick recorded into a HIP graph and replayed (optionally after
    the autotuner).  One JSON line; not the headline metric."""
    import torch
    from cube2_amd import _lib
    from cube2_amd.geometry import C2
    from cube2_amd.kick import ShortRangeKick
    from cube2_amd.layout import build_layout_device
    from cube2_amd.synth import config2

    t0 = time.perf_counter()
    pos, vel, cnt = config2(seed=2)
    n = pos.shape[0]
    rhoc, idx, xp, vp, n_used = build_layout_device(pos, vel, C2)
    del pos, vel
    log(f"C2 inputs: {n} particles in {C2.ntile} tiles (median {int(np.median(cnt))}, "
        f"max {int(cnt.max())}), {n_used} slots ({time.perf_counter() - t0:.1f} s)")
    _lib.set_pp_mode({"ref": _lib.PP_REF, "exact": _lib.PP_EXACT, "fast": _lib.PP_FAST,
                     "precise": _lib.PP_PRECISE, "plummer": _lib.PP_PLUMMER}[args.pp_mode])
    k = ShortRangeKick.from_device(C2, rhoc, idx, xp, vp, n_used)
    levels = sorted(set(int(x) for x in k.ires))
    for L in levels:
        k.green(L)
    mass, amid, dt = 1.0, 0.02, 1e-3
    tuning = apply_tune_args(args)
    t1 = time.perf_counter()
    res = k.kick(mass, amid, dt)
    cold_ms = 1e3 * (time.perf_counter() - t1)
    for _ in range(args.warmup):
        k.kick(mass, amid, dt)
    if args.autotune:
        from cube2_amd.autotune import autotune
        tuning = autotune(k, mass, amid, dt, log=log)
        log(f"autotune chose {tuning['chosen']}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        k.kick(mass, amid, dt)
    torch.cuda.synchronize()
    eager_ms = 1e3 * (time.perf_counter() - t1) / args.steps
    _lib.profile(True)
    _lib.profile_reset()
    for _ in range(args.steps):
        res = k.kick(mass, amid, dt)
    torch.cuda.synchronize()
    stats = _lib.profile_stats()
    _lib.profile(False)
    k.kick_graph(mass, amid, dt)
    for _ in range(args.warmup):
        k.kick_graph(mass, amid, dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        rg = k.kick_graph(mass, amid, dt)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t1) / args.steps
    assert (rg.p_cand, rg.p_acc) == (res.p_cand, res.p_acc)
    pp_name = {"ref": "pp_pairs_ref", "exact": "pp_pairs_exact", "fast": "pp_pairs_fastq",
               "precise": "pp_pairs_precise", "plummer": "pp_pairs_plummer"}[args.pp_mode]
    pp_ms = stats[pp_name]["ms"] / args.steps
    stages = {nm: round(v["ms"] / args.steps, 3) for nm, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])}
    log(f"C2 kick {ms:.1f} ms (graph), {eager_ms:.1f} ms eager, first call {cold_ms:.1f} ms; "
        f"stages {stages}")
    out = {
        "metric": "C2 batching kick ms/step", "value": round(ms, 3), "unit": "ms/step",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 Green's tables, velocity accumulation, max reductions)",
        "data": "synthetic (seeded lognormal tile counts, uniform within tiles)",
        "config": {"workload": "C2: 4096 tiles of 32^3 fine cells (ng=256 nnt=16 nns=1 ngb=4), "
                               "lognormal counts median 4000 sigma 1.5 cap 200k",
                   "particles": int(n), "n_used_slots": int(n_used), "pp_mode": args.pp_mode,
                   "levels": {int(a): int(b) for a, b in zip(*np.unique(k.ires, return_counts=True))}},
        "kick_ms_eager": round(eager_ms, 3),
        "kick_ms_first_call": round(cold_ms, 3), "pp_ms_per_step": round(pp_ms, 3),
        "pp_pair_interactions_per_s": float(f"{res.p_acc / (pp_ms * 1e-3):.4g}"),
        "p_cand": int(res.p_cand), "p_acc": int(res.p_acc), "stage_ms_per_step": stages,
        "tuning": tuning, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def apply_tune_args(args):
    """Sets the tunables given by --tune NAME=VALUE; returns them, or None."""
    if not args.tune:
        return None
    from cube2_amd import _lib
    chosen = {}
    for item in args.tune:
        name, _, value = item.partition("=")
        _lib.set_tuning(_lib.TUNE_NAMES[name], int(value))
        chosen[name] = int(value)
    log(f"tunables set: {chosen}")
    return {"chosen": chosen, "source": "--tune"}


def bench_c4(args) -> int:
    """Learned stand-in C4: 2M training samples (the particles the network
    stands in for) from four C1-geometry boxes at z = 10, 3, 1, 0 kicked by
    the plain PP; Adam training on the matrix-core kernels (timed); then on
    the 256^3 headline set the plain kick and the learned kick (graph
    replays, timed) and the learned kick's error against the plain one: on
    the substituted particles' PP acceleration and on every particle's
    short-range (PM3 + PP) Delta-v; and the substituted error on a held-out
    box (seed 99, z = 0).  One JSON line; not the headline metric."""
    import torch
    from cube2_amd import _lib
    from cube2_amd.geometry import C1
    from cube2_amd.kick import ShortRangeKick
    from cube2_amd.layout import build_layout_device
    from cube2_amd.learned import CHANNELS, PATCH, LearnedPP, TrainingSet, box_samples, c1_style_box, train
    from cube2_amd.synth import config1

    t0 = time.perf_counter()
    sets = []
    for seed, z in ((10, 10.0), (11, 3.0), (12, 1.0), (13, 0.0)):
        ts, k = box_samples(C1, c1_style_box(seed, z), max_samples=args.c4_samples // 4, seed=seed)
        del k
        sets.append(ts)
        log(f"C4 box seed {seed} z {z}: {ts.n} samples ({time.perf_counter() - t0:.0f} s)")
    ts = TrainingSet.concat(sets)
    del sets
    torch.cuda.empty_cache()
    net = LearnedPP(seed=0)
    steps, batch = args.c4_steps, 256
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hist = train(net, ts, steps=steps, batch=batch, log=log)
    torch.cuda.synchronize()
    train_ms = 1e3 * (time.perf_counter() - t1) / steps
    del ts
    torch.cuda.empty_cache()
    # flops of a training step: forward convolutions, backward dW and dX
    fwd, bwd, D = 0, 0, PATCH
    for l in range(4):
        P = D - 2
        f = P ** 3 * 27 * CHANNELS[l] * CHANNELS[l + 1] * 2
        fwd += f
        bwd += f * (2 if l else 1)
        D = P
    tflops = batch * (fwd + bwd) / (train_ms * 1e-3) / 1e12
    log(f"C4 training: {train_ms:.2f} ms/step, {tflops:.2f} TFLOP/s")

    # held-out box: substituted particles' error against plain PP
    hold, kh = box_samples(C1, c1_style_box(99, 0.0), max_samples=200000, seed=99)
    pred = net.predict(hold).double()
    tgt = hold.target.double()
    held_err = float(torch.sqrt(((pred - tgt) ** 2).sum() / (tgt ** 2).sum()))
    del hold, kh
    torch.cuda.empty_cache()

    # the headline 256^3 set: plain and learned kicks
    pos, vel = config1(seed=args.seed)
    rhoc, idx, xp, vp, n_used = build_layout_device(pos, vel, C1)
    del pos, vel
    k = ShortRangeKick.from_device(C1, rhoc, idx, xp, vp, n_used)
    mass, amid, dt = 1.0, 0.02, 1e-3
    vp0 = k.d_vp.clone()
    levels = sorted(set(int(x) for x in k.ires))
    for L in levels:
        k.green(L)
    net.load()

    def timed(learned):
        k.d_vp.copy_(vp0)
        k.kick_graph(mass, amid, dt, learned=learned)
        ms = []
        for _ in range(args.steps):
            k.d_vp.copy_(vp0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            r = k.kick_graph(mass, amid, dt, learned=learned)
            torch.cuda.synchronize()
            ms.append(1e3 * (time.perf_counter() - t))
        return min(ms), r, k.d_vp.clone()

    plain_ms, rp, vp_plain = timed(False)
    learned_ms, rl, vp_learned = timed(True)
    # stage profile of one eager learned kick and one plain kick, with PM3 and PP
    # serial so that each stage's events time it alone (pair kernel, network)
    _lib.set_tuning(_lib.TUNE_KICK_OVERLAP, 2)
    k.d_vp.copy_(vp0)
    torch.cuda.synchronize()
    _lib.profile(True)
    _lib.profile_reset()
    k.kick(mass, amid, dt, learned=True)
    torch.cuda.synchronize()
    st_l = _lib.profile_stats()
    _lib.profile_reset()
    k.d_vp.copy_(vp0)
    torch.cuda.synchronize()
    k.kick(mass, amid, dt)
    torch.cuda.synchronize()
    st_p = _lib.profile_stats()
    _lib.profile(False)
    _lib.set_tuning(_lib.TUNE_KICK_OVERLAP, 0)
    _lib.lib().cube2_learned_clear()
    dv_p = (vp_plain - vp0).double()
    dv_l = (vp_learned - vp0).double()
    diff = (dv_l - dv_p)
    moved = (diff != 0).any(1)
    sr_err = float(torch.sqrt((diff ** 2).sum() / (dv_p ** 2).sum()))
    sub_err = float(torch.sqrt((diff[moved] ** 2).sum() / (dv_p[moved] ** 2).sum())) if moved.any() else 0.0
    pp_plain = st_p.get("pp_pairs_ref", {}).get("ms", float("nan"))
    pp_dense = st_l.get("pp_pairs_ref", {}).get("ms", float("nan"))
    net_ms = sum(v["ms"] for n_, v in st_l.items() if n_.startswith("learned_"))
    lay = st_l.get("learned_layers", {})
    conv_tflops = lay.get("flops", 0.0) / (lay.get("ms", float("nan")) * 1e-3) / 1e12
    log(f"C4 kick: plain {plain_ms:.1f} ms, learned {learned_ms:.1f} ms; pair kernel {pp_plain:.1f} -> "
        f"{pp_dense:.1f} ms, network {net_ms:.1f} ms ({conv_tflops:.1f} TFLOP/s in the layers); "
        f"short-range error {sr_err:.3e}, substituted {sub_err:.3f}, held-out {held_err:.3f}")
    out = {
        "metric": "C4 learned stand-in: learned vs plain kick ms at 256^3, short-range error",
        "value": round(learned_ms, 3), "unit": "ms/step", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": round(learned_ms, 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (FP32 accumulation; FP32 master weights)",
        "data": "synthetic (four C1-geometry boxes at z = 10, 3, 1, 0, seeds 10-13; held-out seed 99; "
                "kicks on the 256^3 headline set, seed 1)",
        "config": {"workload": "C4: 3-D CNN 4 x 3^3 convolutions, 4 -> 32 -> 32 -> 32 -> 3 channels, "
                               "fine-cell count + sub-cell offset input, CIC to particles; stands in "
                               "for PP in level-2 PP cells of < 64 particles",
                   "batch": batch, "training_samples": int(args.c4_samples), "steps": steps},
        "kick_ms_plain": round(plain_ms, 3), "kick_ms_learned": round(learned_ms, 3),
        "speedup": round(plain_ms / learned_ms, 4),
        "pair_kernel_ms_plain": round(pp_plain, 3), "pair_kernel_ms_learned": round(pp_dense, 3),
        "sparse_pp_time_share": round((pp_plain - pp_dense) / pp_plain, 4),
        "network_ms_in_kick": round(net_ms, 3), "network_layers_tflops": round(conv_tflops, 2),
        "p_acc_plain": int(rp.p_acc), "p_acc_learned": int(rl.p_acc),
        "short_range_rel_rms_error": sr_err, "substituted_pp_rel_rms_error": sub_err,
        "held_out_substituted_rel_rms_error": held_err, "without_network_substituted_error": 1.0,
        "train_ms_per_step": round(train_ms, 3), "train_tflops": round(tflops, 3),
        "roofline": {"kernel": "k_conv3 (learned layers in the kick)", "bound": "mfma",
                     "achieved": round(conv_tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(conv_tflops / 2500.0, 5), "traffic": None},
        "loss_first_last": [round(float(np.mean(hist[:50])), 4), round(float(np.mean(hist[-100:])), 4)],
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pp-mode", default="ref", choices=["ref", "exact", "fast", "precise", "plummer"])
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="append", default=[], metavar="NAME=VALUE",
                    help="set a library tunable (names as in cube2_amd._lib.TUNE_NAMES)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the launch-shape tunables on this workload before timing")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager dispatch only (value = eager kick), e.g. under counters")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-pm2", action="store_true", help="skip the PM2 measurement")
    ap.add_argument("--no-drift", action="store_true", help="skip the drift measurement")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the drop-in C ABI timing on host arrays")
    ap.add_argument("--workload", default="c1", choices=["c1", "c2", "c4"],
                    help="c1: the headline kick at 256^3; c2: the 4096-tile batching workload; "
                         "c4: the learned stand-in (training and held-out error)")
    ap.add_argument("--c4-samples", type=int, default=2_000_000, help="C4 training samples")
    ap.add_argument("--c4-steps", type=int, default=6000, help="C4 Adam steps")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (calibrates HBM counters)")
    args = ap.parse_args(argv)
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this path runs on a single MI355X; "
              "sharding across GPUs (by parent tile) is not implemented in this library.",
              file=sys.stderr)
        return 2

    if args.workload == "c2":
        return bench_c2(args)
    if args.workload == "c4":
        return bench_c4(args)

    import torch
    from cube2_amd import _lib
    from cube2_amd.geometry import C1
    from cube2_amd.kick import ShortRangeKick
    from cube2_amd.layout import build_layout
    from cube2_amd.synth import config1

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    t0 = time.perf_counter()
    pos, vel = config1(seed=args.seed)
    lay = build_layout(pos, vel, C1)
    log(f"inputs: {lay.n_particles} particles, {lay.n_used} slots ({time.perf_counter() - t0:.1f} s)")
    _lib.set_pp_mode({"ref": _lib.PP_REF, "exact": _lib.PP_EXACT, "fast": _lib.PP_FAST,
                     "precise": _lib.PP_PRECISE, "plummer": _lib.PP_PLUMMER}[args.pp_mode])
    k = ShortRangeKick(lay)
    levels = sorted(set(int(x) for x in k.ires))
    for L in levels:
        k.green(L)
    torch.cuda.synchronize()
    log(f"levels {dict(zip(*np.unique(k.ires, return_counts=True)))}; Green tables built")
    mass, amid, dt = 1.0, 0.02, 1e-3

    tuning = apply_tune_args(args)  # before the first kick, so that warmup runs the same code
    t1 = time.perf_counter()
    res = k.kick(mass, amid, dt)
    cold_ms = 1e3 * (time.perf_counter() - t1)
    log(f"first kick (allocations, plans): {cold_ms:.1f} ms "
        f"(P_cand {res.p_cand:.3e}, P_acc {res.p_acc:.3e})")
    for w in range(args.warmup):
        t1 = time.perf_counter()
        res = k.kick(mass, amid, dt)
        log(f"warmup {w}: {1e3 * (time.perf_counter() - t1):.1f} ms")
    if args.autotune:
        from cube2_amd.autotune import autotune
        tuning = autotune(k, mass, amid, dt, log=log)
        log(f"autotune chose {tuning['chosen']}")

    # eager dispatch, with per-stage HIP-event timing
    _lib.profile(True)
    _lib.profile_reset()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for s in range(args.steps):
        res = k.kick(mass, amid, dt)
        log(f"eager step {s} done")
    torch.cuda.synchronize()
    wall_eager = time.perf_counter() - t1
    stats = _lib.profile_stats()
    _lib.profile(False)
    eager_ms = 1e3 * wall_eager / args.steps

    # the same kick recorded once into a HIP graph, then replayed
    graph_ms = None
    ms_step = eager_ms
    if not args.no_graph:
        t1 = time.perf_counter()
        rg = k.kick_graph(mass, amid, dt)
        log(f"graph recorded ({rg.recorded}) and run: {1e3 * (time.perf_counter() - t1):.1f} ms")
        for w in range(args.warmup):
            k.kick_graph(mass, amid, dt)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for s in range(args.steps):
            rg = k.kick_graph(mass, amid, dt)
            log(f"graph step {s} done")
        torch.cuda.synchronize()
        wall_graph = time.perf_counter() - t1
        assert not rg.recorded and (rg.p_cand, rg.p_acc) == (res.p_cand, res.p_acc)
        graph_ms = 1e3 * wall_graph / args.steps
        ms_step = graph_ms

    # the Fortran host's path: the two drop-in calls on host arrays, transfers included
    dropin = None
    if not args.no_dropin:
        greens_host = {L: k.green(L).cpu().numpy() for L in levels}
        dropin = time_dropin(lay, k, greens_host, mass, amid, dt, max(1, min(args.steps, 3)),
                             exact=args.pp_mode == "exact")
        del greens_host
        log(f"drop-in C ABI kick (host arrays): {dropin['ms']:.1f} ms")

    # the same layout built on the device from the positions (not part of the kick)
    from cube2_amd.layout import build_layout_device
    d_pos = torch.from_numpy(pos).cuda()
    d_vel = torch.from_numpy(np.ascontiguousarray(vel, np.float32)).cuda()
    lay_ms = []
    for _ in range(3):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        d_rhoc, d_idx, d_xp, d_vp, d_n = build_layout_device(d_pos, d_vel, C1)
        torch.cuda.synchronize()
        lay_ms.append(1e3 * (time.perf_counter() - t1))
    assert d_n == lay.n_used and np.array_equal(d_idx.cpu().numpy(), lay.idx_b_r)
    assert torch.equal(d_xp, k.d_xp), "device layout differs from the host layout"
    del d_pos, d_vel, d_rhoc, d_idx, d_xp, d_vp, pos, vel
    log(f"device layout build: {min(lay_ms):.1f} ms (host: identical arrays)")

    # PM2 (kick.f90:76-105) on the same image: the 64 parent tiles on the
    # standard grid, mesh ngt^3 each; timed on its own, not part of the kick
    pm2 = None
    if not args.no_pm2:
        from cube2_amd.kick import KickResult
        r2 = KickResult()
        k.green2()
        t1 = time.perf_counter()
        k.pm2(mass, amid, dt, r2)
        pm2_cold = 1e3 * (time.perf_counter() - t1)
        _lib.profile(True)
        _lib.profile_reset()
        pm2_ms = []
        for s in range(args.steps):
            t1 = time.perf_counter()
            k.pm2(mass, amid, dt, r2)          # the pending Delta-v accumulates
            pm2_ms.append(1e3 * (time.perf_counter() - t1))
        st2 = _lib.profile_stats()
        _lib.profile(False)
        k.apply_pending()
        torch.cuda.synchronize()
        n2 = C1.ngt
        pm2_stages = {n: round(v["ms"] / args.steps, 3) for n, v in
                      sorted(st2.items(), key=lambda kv: -kv[1]["ms"]) if n.startswith("pm2_")}
        dn, dst = max(((n, v) for n, v in st2.items() if n.startswith("pm2_")),
                      key=lambda kv: kv[1]["ms"])
        gbs = dst["bytes"] / (dst["ms"] * 1e-3) / 1e9
        pm2 = {"ms": round(min(pm2_ms), 3), "first_call_ms": round(pm2_cold, 3),
               "mesh": n2, "tiles": C1.ntile, "f2max": r2.f2max_pm2, "stage_ms": pm2_stages,
               "roofline": dict(kernel=dn, bound="hbm", achieved=round(gbs, 1),
                                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 4),
                                avg_launch_ms=round(dst["ms"] / dst["launches"], 3))}
        log(f"PM2: {pm2['ms']:.1f} ms ({C1.ntile} tiles, mesh {n2}^3); stages {pm2_stages}")

    # drift (drift.f90) of the same image with a step that moves the fastest
    # particle one coarse cell, then the next kick's layout from the result
    drift = None
    if not args.no_drift:
        from cube2_amd.drift import drift_device, next_layout, refresh_halo_velocities
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        refresh_halo_velocities(k.d_rhoc, k.d_idx, k.d_vp, k.n_used, C1)   # buffer_v
        torch.cuda.synchronize()
        bv_ms = 1e3 * (time.perf_counter() - t1)
        vabs = k.d_vp[:k.n_used].abs()
        vz_max = float(vabs[:, 2].max())
        dt_d = float(np.float32(C1.ratio_cs / float(vabs.max())))
        dr_ms, nl_ms = [], []
        for _ in range(3):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            img = drift_device(k.d_rhoc, k.d_idx, k.d_xp, k.d_vp, k.n_used, dt_d, vz_max, C1)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            nxt = next_layout(*img, C1)
            torch.cuda.synchronize()
            dr_ms.append(1e3 * (t2 - t1))
            nl_ms.append(1e3 * (time.perf_counter() - t2))
        assert img[4] == lay.n_particles, "drift lost particles"
        drift = {"ms": round(min(dr_ms), 3), "next_layout_ms": round(min(nl_ms), 3),
                 "halo_velocities_ms": round(bv_ms, 3),
                 "dt_mid": dt_d, "vz_max": vz_max, "particles": int(img[4]),
                 "next_n_used": int(nxt[4])}
        del img, nxt
        log(f"drift: {drift['ms']:.1f} ms, next layout {drift['next_layout_ms']:.1f} ms")

    pp_name = {"ref": "pp_pairs_ref", "exact": "pp_pairs_exact", "fast": "pp_pairs_fastq",
               "precise": "pp_pairs_precise", "plummer": "pp_pairs_plummer"}[args.pp_mode]
    pp_ms = stats[pp_name]["ms"] / args.steps
    flops = pp_flops(res.p_cand, res.p_acc)
    pair_rate = res.p_acc / (pp_ms * 1e-3)
    # dominant kernel: the pair kernel, the largest GPU time of the kick (the
    # other stages' HIP-event spans in the overlapped kick include their
    # waits for CU slots beside it, so their spans do not rank them)
    name, st = pp_name, stats[pp_name]
    per_launch_ms = st["ms"] / st["launches"]
    if name.startswith("pp_pairs"):
        achieved = flops / (pp_ms * 1e-3) / 1e12
        roof = dict(kernel=name, bound="valu", achieved=round(achieved, 3),
                    peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(achieved / FP32_PEAK_TFLOPS, 4), traffic=profiled_traffic(name),
                    avg_launch_ms=round(per_launch_ms, 3))
    else:
        achieved = st["bytes"] / (st["ms"] * 1e-3) / 1e9
        roof = dict(kernel=name, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                    avg_launch_ms=round(per_launch_ms, 3))
    stages = {n: round(v["ms"] / args.steps, 3) for n, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])}
    log(f"per-step stage ms: {stages}")
    if name.startswith("pp_pairs") and _lib.get_tuning(_lib.TUNE_KICK_OVERLAP) == 1:
        # the pair kernel shares the GPU with PM3 in the timed kicks; one more
        # kick with PM3 and PP serial times it alone (reported beside the
        # roofline of the timed region, not in place of it)
        _lib.set_tuning(_lib.TUNE_KICK_OVERLAP, 2)
        _lib.profile(True)
        _lib.profile_reset()
        k.kick(mass, amid, dt)
        torch.cuda.synchronize()
        alone = _lib.profile_stats()[name]
        _lib.profile(False)
        _lib.set_tuning(_lib.TUNE_KICK_OVERLAP, 0)
        alone_ms = alone["ms"] / alone["launches"]
        roof["alone_avg_launch_ms"] = round(alone_ms, 3)
        roof["alone_frac"] = round(flops / (alone_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4)
        log(f"pair kernel alone (PM3 and PP serial): {alone_ms:.2f} ms, "
            f"frac {roof['alone_frac']}")

    cpu = None
    if not args.no_cpu:
        greens_host = {L: k.green(L).cpu().numpy() for L in levels}
        log("cpu baseline ...")
        cpu = cpu_baseline(lay, k.ires, k.isort, greens_host, (res.p_cand, res.p_acc), mass, amid,
                           dt, args.cpu_budget)
        import oracle
        if pm2 is not None:
            g2 = k.green2().cpu().numpy()
            t1 = time.perf_counter()
            nt2 = 8
            oracle.pm2_force(C1, lay.rhoc, lay.idx_b_r, lay.xp, g2, mass, amid, dt,
                             tiles=(0, nt2))
            t_s = time.perf_counter() - t1
            pm2["cpu_baseline"] = dict(value=round(t_s / nt2 * C1.ntile * 1e3, 1),
                                       unit="ms/call", cores=cpu["cores"], kind="port",
                                       sample=f"oracle PM2 on tiles 0-{nt2 - 1} of {C1.ntile} "
                                              f"({t_s:.2f} s), scaled by {C1.ntile}/{nt2}")
        if drift is not None:
            t1 = time.perf_counter()
            oracle.drift(C1, lay.rhoc, lay.idx_b_r, lay.xp, k.vp_host(), drift["dt_mid"],
                         drift["vz_max"])
            drift["cpu_baseline"] = dict(value=round(1e3 * (time.perf_counter() - t1), 1),
                                         unit="ms/call", cores=1, kind="port",
                                         sample="oracle drift of the whole image, one thread")

    out = {
        "metric": "kick ms/step + PP pair-int/s @256^3",
        "value": round(ms_step, 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 Green's tables, velocity accumulation, max reductions)",
        "data": "synthetic (seeded Zel'dovich lattice + Plummer spheres)",
        "config": {"workload": "C1: 256^3 particles, 300 Mpc/h, fine mesh 512^3 as 8^3 subtiles, "
                               "ng=256 nnt=4 nns=2 ngb=24, PM3+PP kick",
                   "pp_mode": args.pp_mode, "n_used_slots": int(lay.n_used),
                   "levels": {int(a): int(b) for a, b in zip(*np.unique(k.ires, return_counts=True))}},
        "kick_ms_graph_replay": None if graph_ms is None else round(graph_ms, 3),
        "kick_ms_eager": round(eager_ms, 3),
        "kick_ms_first_call": round(cold_ms, 3),
        "layout_build_ms": round(min(lay_ms), 3),
        "dropin_abi_kick": dropin,
        "pp_pair_interactions_per_s": float(f"{pair_rate:.4g}"),
        "pp_ms_per_step": round(pp_ms, 3),
        "p_cand": int(res.p_cand),
        "p_acc": int(res.p_acc),
        "stage_ms_per_step": stages,
        "pm2": pm2,
        "drift": drift,
        "roofline": roof,
        "tuning": tuning,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
