"""Learned stand-in for PP in sparse cells (no counterpart in the reference).

A 3-D CNN maps the fine-cell stencil around a fine cell of the level-2 mesh
(0.5 standard cells) to the PP acceleration at that cell; a particle's PP
acceleration is the cloud-in-cell interpolation of the cell values at its
position.  Four 3^3 valid convolutions, 4 -> 32 -> 32 -> 32 -> 3 channels
(ReLU after the first three).  Input channels per fine cell: log(1 + n) of
its particle count and the mean offset of its particles from the cell
centre (the sub-cell information a count alone lacks).

Training (this module): samples are the particles the network stands in for
(physical particles of level-2 subtiles in PP cells of fewer than 64
particles) of C1-style boxes kicked by the plain PP alone; a sample is the
10^3 x 4 input patch whose 2^3 outputs surround the particle, its CIC
fractions and its plain-PP acceleration.  The forward pass runs the
library's matrix-core convolution (cube2_cnn_conv3_dev); the backward pass
im2col and BF16 GEMMs (cube2_gemm_bf16_dev); the loss is the CIC
combination's MSE (cube2_cnn_cic_mse_dev); Adam updates FP32 master weights.

In the kick (cube2_kick_dev with CUBE2_KICK_LEARNED, after
``LearnedPP.load``), the library evaluates the same layers over each
level-2 subtile's grid and the pair kernel skips those particles.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib

CHANNELS = (4, 32, 32, 32, 3)
PATCH = 10          # training patch: 10^3 inputs -> 2^3 outputs
OCCUPANCY = 64      # PP cells of fewer particles are left to the network


def _p(t):
    return t.data_ptr()


class LearnedPP:
    def __init__(self, params=None, seed: int = 0, device: str = "cuda:0"):
        """params: [(W [27*c_in][c_out], b [c_out])] (numpy), or He-uniform
        weights from ``seed``."""
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        if params is None:
            rng = np.random.default_rng(seed)
            params = []
            for cin, cout in zip(CHANNELS[:-1], CHANNELS[1:]):
                lim = np.sqrt(6.0 / (27 * cin))
                params.append((rng.uniform(-lim, lim, (27 * cin, cout)), np.zeros(cout)))
        t = lambda a: torch.tensor(np.asarray(a, np.float32), device=self.dev).contiguous()  # noqa: E731
        self.W = [t(W) for W, _ in params]
        self.b = [t(b) for _, b in params]
        self.Wb = [torch.empty(w.shape, dtype=torch.int16, device=self.dev) for w in self.W]
        self.Wp = []
        for cin, cout in zip(CHANNELS[:-1], CHANNELS[1:]):
            taps, coutp = (27 if cin == 32 else 32), (16 if cout == 3 else 32)
            self.Wp.append(torch.empty(taps * coutp * cin, dtype=torch.int16, device=self.dev))
        self._refresh(bf16=True)
        z = lambda x: torch.zeros_like(x)  # noqa: E731
        self.mW, self.vW = [z(w) for w in self.W], [z(w) for w in self.W]
        self.mb, self.vb = [z(b) for b in self.b], [z(b) for b in self.b]
        self.t = 0
        self.y_scale = 1.0

    def _refresh(self, bf16: bool):
        """BF16 copies of the weights: Wb for the GEMMs (unless Adam wrote
        them), Wp packed for the convolution."""
        L = _lib.lib()
        for l, (cin, cout) in enumerate(zip(CHANNELS[:-1], CHANNELS[1:])):
            if bf16:
                _lib.check(L.cube2_to_bf16_dev(_p(self.W[l]), _p(self.Wb[l]), self.W[l].numel()))
            _lib.check(L.cube2_cnn_conv_pack_dev(_p(self.W[l]), _p(self.Wp[l]), cin, cout))

    def params_numpy(self):
        return [(w.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64))
                for w, b in zip(self.W, self.b)]

    def forward(self, x):
        """x [B][D][D][D][4] float32 on the device -> out [B][D-8][D-8][D-8][3]
        (the matrix-core convolution per layer, FP32 activations kept for
        the backward pass)."""
        torch, L = self.torch, _lib.lib()
        B, D = x.shape[0], x.shape[1]
        h = x.contiguous()
        self.acts = [h]
        for l in range(4):
            cin, cout = CHANNELS[l], CHANNELS[l + 1]
            P = D - 2
            y = torch.empty((B, P, P, P, cout), dtype=torch.float32, device=self.dev)
            _lib.check(L.cube2_cnn_conv3_dev(_p(h), 0, _p(self.Wp[l]), _p(self.b[l]), _p(y), 0, B,
                                             D, cin, cout, int(l < 3)))
            h = y
            self.acts.append(h)
            D = P
        return h

    def backward(self, dout):
        """Gradients [(dW, db)] of sum(dout * out) for the last forward."""
        torch, L = self.torch, _lib.lib()
        B = dout.shape[0]
        g = dout.contiguous().clone()
        grads = [None] * 4
        for l in range(3, -1, -1):
            cin, cout = CHANNELS[l], CHANNELS[l + 1]
            D = self.acts[l].shape[1]
            P, Kd = D - 2, 27 * cin
            rows = B * P ** 3
            y = self.acts[l + 1]
            dy = g.reshape(rows, cout).contiguous()
            dyb = torch.empty((rows, cout), dtype=torch.int16, device=self.dev)
            db = torch.empty(cout, dtype=torch.float32, device=self.dev)
            _lib.check(L.cube2_cnn_act_bwd_dev(_p(y), _p(dy), _p(dyb), _p(db), rows, cout,
                                               int(l < 3)))
            col = torch.empty((rows, Kd), dtype=torch.int16, device=self.dev)
            _lib.check(L.cube2_cnn_im2col_dev(_p(self.acts[l]), _p(col), B, D, cin))
            dW = torch.empty((Kd, cout), dtype=torch.float32, device=self.dev)
            # dW = col^T dY
            _lib.check(L.cube2_gemm_bf16_dev(_p(col), 1, Kd, _p(dyb), cout, 1, _p(dW),
                                             cout, Kd, cout, rows, 0))
            grads[l] = (dW, db)
            if l > 0:
                # dcol = dY W^T, then the adjoint of im2col
                dcol = torch.empty((rows, Kd), dtype=torch.float32, device=self.dev)
                _lib.check(L.cube2_gemm_bf16_dev(_p(dyb), cout, 1, _p(self.Wb[l]), 1, cout,
                                                 _p(dcol), Kd, rows, Kd, cout, 0))
                g = torch.empty((B, D, D, D, cin), dtype=torch.float32, device=self.dev)
                _lib.check(L.cube2_cnn_col2im_dev(_p(dcol), _p(g), B, D, cin))
        return grads

    def step(self, grads, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        L = _lib.lib()
        self.t += 1
        for l, (dW, db) in enumerate(grads):
            _lib.check(L.cube2_cnn_adam_dev(_p(self.W[l]), _p(dW), _p(self.mW[l]), _p(self.vW[l]),
                                            _p(self.Wb[l]), self.W[l].numel(), lr, b1, b2, eps,
                                            self.t))
            _lib.check(L.cube2_cnn_adam_dev(_p(self.b[l]), _p(db), _p(self.mb[l]), _p(self.vb[l]),
                                            None, self.b[l].numel(), lr, b1, b2, eps, self.t))
        self._refresh(bf16=False)

    def cic_mse(self, out, frac, target):
        """(loss, dloss/dout, predictions) of the CIC combination of out
        [n][2][2][2][3] with fractions [n][3] against target [n][3]."""
        torch = self.torch
        n = out.shape[0]
        out, frac, target = out.contiguous(), frac.contiguous(), target.contiguous()
        d = torch.empty_like(out)
        pred = torch.empty((n, 3), dtype=torch.float32, device=self.dev)
        loss = torch.empty(1, dtype=torch.float32, device=self.dev)
        _lib.check(_lib.lib().cube2_cnn_cic_mse_dev(_p(out), _p(frac), _p(target), n, _p(pred),
                                                    _p(d), _p(loss)))
        return float(loss.item()), d, pred

    def predict(self, ts, idx=None, batch: int = 8192):
        """Accelerations (physical units) of samples idx of a TrainingSet."""
        torch = self.torch
        n = ts.n if idx is None else int(idx.numel())
        out = []
        for s in range(0, n, batch):
            sel = (torch.arange(s, min(n, s + batch), device=self.dev) if idx is None
                   else idx[s:s + batch])
            x = ts.patches(sel)
            o = self.forward(x)
            _, _, p = self.cic_mse(o, ts.frac[sel], ts.target[sel])
            out.append(p * self.y_scale)
        return torch.cat(out) if out else torch.empty((0, 3), device=self.dev)

    def load(self, occupancy: int = OCCUPANCY) -> None:
        """Hands the weights to the library for learned kicks."""
        Ws = [np.ascontiguousarray(w.cpu().numpy(), np.float32) for w in self.W]
        bs = [np.ascontiguousarray(b.cpu().numpy(), np.float32) for b in self.b]
        Wp = (C.c_void_p * 4)(*[w.ctypes.data for w in Ws])
        bp = (C.c_void_p * 4)(*[b.ctypes.data for b in bs])
        _lib.check(_lib.lib().cube2_learned_set(Wp, bp, float(self.y_scale), int(occupancy)))


# --------------------------------------------------------------- data
def c1_style_box(seed: int, z: float, geom=None, npd: int = 216, n_total: int = 256 ** 3,
                 box_mpc: float = 300.0):
    """Config-1-style particles at redshift z: an npd^3 Zel'dovich lattice
    plus Plummer spheres holding the rest of n_total (40% at the C1 sizes),
    in the geometry ``geom`` (C1 by default)."""
    from .geometry import C1
    from .synth import plummer_spheres, zeldovich_lattice
    g = geom or C1
    p1, _ = zeldovich_lattice(npd, g, box_mpc, z, seed)
    n2 = n_total - npd ** 3
    p2, _, _ = plummer_spheres(n2, max(8, round(2000 * n2 / (256 ** 3 - 216 ** 3))), g, seed + 1000)
    return np.concatenate([p1, p2])


class TrainingSet:
    """Samples of one or more boxes: the boxes' input grids X
    [jobs][DI][DI][DI][4] (device), per sample its job, patch origin, CIC
    fractions and plain-PP acceleration (device tensors)."""

    def __init__(self, X, DI, job, base, frac, target, slot=None):
        self.X, self.DI = X, DI
        self.job, self.base, self.frac, self.target = job, base, frac, target
        self.slot = slot  # the samples' particle slots in their box (one box only)
        self.n = int(job.numel())

    def patches(self, sel):
        torch = __import__("torch")
        n = int(sel.numel())
        out = torch.empty((n, PATCH, PATCH, PATCH, 4), dtype=torch.float32, device=self.X.device)
        # held until the call returns: the library reads them through their addresses
        job, base = self.job[sel].contiguous(), self.base[sel].contiguous()
        _lib.check(_lib.lib().cube2_learned_gather_dev(_p(self.X), self.DI, _p(job), _p(base), n,
                                                       _p(out)))
        return out

    @staticmethod
    def concat(sets):
        torch = __import__("torch")
        offs, o = [], 0
        for s in sets:
            offs.append(o)
            o += s.X.shape[0]
        return TrainingSet(torch.cat([s.X for s in sets]), sets[0].DI,
                           torch.cat([s.job + off for s, off in zip(sets, offs)]),
                           torch.cat([s.base for s in sets]), torch.cat([s.frac for s in sets]),
                           torch.cat([s.target for s in sets]))


def box_samples(geom, pos, max_samples: int | None = None, seed: int = 0,
                occupancy: int = OCCUPANCY):
    """Kicks one box with the plain PP alone (PRECISE arithmetic, every
    subtile at level 2, zero initial velocities) and returns its TrainingSet
    (at most max_samples samples, drawn at random) and the kick."""
    import torch

    from .kick import ShortRangeKick
    from .layout import build_layout_device
    vel = np.zeros_like(pos, dtype=np.float32)
    rhoc, idx, xp, vp, n_used = build_layout_device(pos, vel, geom)
    k = ShortRangeKick.from_device(geom, rhoc, idx, xp, vp, n_used)
    k.set_levels(np.full(geom.ns3, 2, np.int32))
    _lib.set_pp_mode(_lib.PP_PRECISE)
    try:
        k.kick(1.0, 1.0, 1.0, pm3=False, pp=True)
    finally:
        _lib.set_pp_mode(_lib.PP_REF)
    L = _lib.lib()
    nj, di = C.c_int(0), C.c_int(0)
    _lib.check(L.cube2_learned_inputs_dev(None, 0, C.byref(nj), C.byref(di)))
    X = torch.empty((nj.value, di.value, di.value, di.value, 4), dtype=torch.float32,
                    device=k.dev)
    _lib.check(L.cube2_learned_inputs_dev(_p(X), nj.value, C.byref(nj), C.byref(di)))
    cap = int(n_used)
    job = torch.empty(cap, dtype=torch.int32, device=k.dev)
    base = torch.empty((cap, 3), dtype=torch.int32, device=k.dev)
    frac = torch.empty((cap, 3), dtype=torch.float32, device=k.dev)
    slot = torch.empty(cap, dtype=torch.int32, device=k.dev)
    n = C.c_int(0)
    _lib.check(L.cube2_learned_samples_dev(occupancy, cap, _p(job), _p(base), _p(frac), _p(slot),
                                           C.byref(n)))
    n = n.value
    sel = torch.arange(n, device=k.dev)
    if max_samples is not None and n > max_samples:
        g = torch.Generator(device=k.dev)
        g.manual_seed(seed)
        sel = torch.randperm(n, device=k.dev, generator=g)[:max_samples]
    # Delta-v = a a_mid dt / (6 pi) with a_mid = dt = 1 (kick.f90's PP scaling)
    acc = k.d_vp[slot[sel].long()].to(torch.float32) * (6 * math.pi)
    ts = TrainingSet(X, di.value, job[sel].contiguous(), base[sel].contiguous(),
                     frac[sel].contiguous(), acc.contiguous(), slot[sel].contiguous())
    return ts, k


def train(net: LearnedPP, ts: TrainingSet, steps: int = 1500, batch: int = 256, lr: float = 1e-3,
          seed: int = 0, log=None):
    """Adam on the MSE of CIC-combined accelerations in units of their RMS;
    returns the loss history."""
    torch = net.torch
    g = torch.Generator(device=net.dev)
    g.manual_seed(seed)
    net.y_scale = float(torch.sqrt((ts.target.double() ** 2).mean()).item())
    y = (ts.target / net.y_scale).contiguous()
    hist = []
    for s in range(steps):
        idx = torch.randint(0, ts.n, (batch,), device=net.dev, generator=g)
        out = net.forward(ts.patches(idx))
        loss, d, _ = net.cic_mse(out, ts.frac[idx], y[idx])
        net.step(net.backward(d), lr=lr * (0.1 ** (s / steps)))
        hist.append(loss)
        if log and s % 500 == 0:
            log(f"step {s}: loss {loss:.4f}")
    return hist
