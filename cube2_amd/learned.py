"""Learned stand-in for PP in sparse cells (no counterpart in the reference).

A 3-D CNN maps the 9^3 fine-cell density stencil around a fine cell to the
PP acceleration of that cell's particles.  Four 3^3 valid convolutions,
1 -> 32 -> 32 -> 32 -> 3 channels (9 -> 7 -> 5 -> 3 -> 1 cells), ReLU after
the first three; every layer is an im2col and a BF16 GEMM on the matrix
cores (FP32 accumulation, FP32 master weights, Adam), forward and backward,
through the library's cube2_gemm_bf16_dev / cube2_cnn_*_dev kernels.

Training pairs come from the plain PP: a box's particles are kicked with PP
alone (zero initial velocities), each occupied fine cell in PP cells of
fewer than 64 particles gives (its density stencil, its particles' mean
acceleration).  ``evaluate`` assigns the predicted acceleration to every
particle of those cells and compares it with the particles' own PP
accelerations (plain PP is the judge).
"""
from __future__ import annotations

import numpy as np

from . import _lib

CHANNELS = (1, 32, 32, 32, 3)
PATCH = 9


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
        t = lambda a: torch.tensor(np.asarray(a, np.float32), device=self.dev).contiguous()
        self.W = [t(W) for W, _ in params]
        self.b = [t(b) for _, b in params]
        self.Wb = [torch.empty(w.shape, dtype=torch.int16, device=self.dev) for w in self.W]
        for w, wb in zip(self.W, self.Wb):
            _lib.check(_lib.lib().cube2_to_bf16_dev(_p(w), _p(wb), w.numel()))
        z = lambda x: torch.zeros_like(x)
        self.mW, self.vW = [z(w) for w in self.W], [z(w) for w in self.W]
        self.mb, self.vb = [z(b) for b in self.b], [z(b) for b in self.b]
        self.t = 0
        self.y_scale = 1.0

    def params_numpy(self):
        return [(w.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64))
                for w, b in zip(self.W, self.b)]

    def forward(self, x):
        """x [B][9][9][9][1] float32 on the device -> out [B][3]."""
        torch, L = self.torch, _lib.lib()
        B = x.shape[0]
        h = x.contiguous()
        self.acts, self.cols = [h], []
        D = PATCH
        for l in range(4):
            cin, cout = CHANNELS[l], CHANNELS[l + 1]
            P, Kd = D - 2, 27 * cin
            rows = B * P ** 3
            col = torch.empty((rows, Kd), dtype=torch.int16, device=self.dev)
            _lib.check(L.cube2_cnn_im2col_dev(_p(h), _p(col), B, D, cin))
            y = torch.empty((rows, cout), dtype=torch.float32, device=self.dev)
            _lib.check(L.cube2_gemm_bf16_dev(_p(col), Kd, 1, _p(self.Wb[l]), cout, 1, _p(y), cout,
                                             rows, cout, Kd, 0))
            _lib.check(L.cube2_cnn_bias_act_dev(_p(y), _p(self.b[l]), rows, cout, int(l < 3)))
            h = y.view(B, P, P, P, cout)
            self.acts.append(h)
            self.cols.append(col)
            D = P
        return h.reshape(B, 3)

    def backward(self, dout):
        """Gradients [(dW, db)] of sum(dout * out) for the last forward."""
        torch, L = self.torch, _lib.lib()
        B = dout.shape[0]
        g = dout.contiguous().clone()
        grads = [None] * 4
        for l in range(3, -1, -1):
            cin, cout = CHANNELS[l], CHANNELS[l + 1]
            D = PATCH - 2 * l
            P, Kd = D - 2, 27 * cin
            rows = B * P ** 3
            y = self.acts[l + 1]
            dy = g.reshape(rows, cout).contiguous()
            dyb = torch.empty((rows, cout), dtype=torch.int16, device=self.dev)
            db = torch.empty(cout, dtype=torch.float32, device=self.dev)
            _lib.check(L.cube2_cnn_act_bwd_dev(_p(y), _p(dy), _p(dyb), _p(db), rows, cout,
                                               int(l < 3)))
            dW = torch.empty((Kd, cout), dtype=torch.float32, device=self.dev)
            # dW = col^T dY
            _lib.check(L.cube2_gemm_bf16_dev(_p(self.cols[l]), 1, Kd, _p(dyb), cout, 1, _p(dW),
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

    def mse(self, out, target):
        """(loss, dloss/dout) on the device."""
        torch = self.torch
        d = torch.empty_like(out)
        loss = torch.empty(1, dtype=torch.float32, device=self.dev)
        _lib.check(_lib.lib().cube2_cnn_mse_dev(_p(out), _p(target.contiguous()), _p(d), _p(loss),
                                                out.numel()))
        return float(loss.item()), d

    @staticmethod
    def features(patches):
        """Network input of particle-count patches: log(1 + count)."""
        return np.log1p(patches).astype(np.float32)

    def predict(self, patches, batch: int = 4096):
        """Accelerations (physical units) for count patches [N][9][9][9]."""
        torch = self.torch
        X = torch.tensor(self.features(patches), device=self.dev)
        out = []
        for s in range(0, X.shape[0], batch):
            x = X[s:s + batch].reshape(-1, PATCH, PATCH, PATCH, 1)
            out.append(self.forward(x.contiguous()) * self.y_scale)
        return torch.cat(out) if out else torch.empty((0, 3), device=self.dev)


# --------------------------------------------------------------- data
def training_box(seed: int, z: float, npd: int = 40):
    """C1-style box at a small size: a Zel'dovich lattice of npd^3 particles
    at redshift z (60%) plus Plummer spheres with the rest (40%), in the
    64-standard-cell geometry of one 4 x 4 x 4-coarse-cell-tile image."""
    from .geometry import Geometry
    from .synth import plummer_spheres, zeldovich_lattice
    g = Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, app=0.05)
    p1, _ = zeldovich_lattice(npd, g, 100.0, z, seed)
    n2 = int(round(npd ** 3 * 0.4 / 0.6))
    p2, _, _ = plummer_spheres(n2, max(8, n2 // 2000), g, seed + 1000)
    return g, np.concatenate([p1, p2])


def pp_pairs(g, pos, occupancy: int = 64):
    """Plain-PP training pairs of one box: (density patches [N][9][9][9]
    float32, mean accelerations [N][3], fine-cell index [N], per-particle
    (cell row, acceleration)) for the occupied fine cells of PP cells with
    fewer than ``occupancy`` particles, all subtiles at level 2."""
    import torch
    from .kick import ShortRangeKick
    from .layout import build_layout
    lay = build_layout(pos, None, g)
    k = ShortRangeKick(lay)
    k.set_levels(np.full(g.ns3, 2, np.int32))
    mass, amid, dt = 1.0, 1.0, 1.0
    _lib.set_pp_mode(_lib.PP_PRECISE)
    try:
        k.kick(mass, amid, dt, pm3=False, pp=True)
    finally:
        _lib.set_pp_mode(_lib.PP_REF)  # the library's default arithmetic
    phys = lay.physical_mask()
    acc = np.zeros((pos.shape[0], 3))
    # Delta-v = a * a_mid * dt / (6 pi) per particle (kick.f90's PP scaling)
    acc[lay.src[phys]] = k.vp_host()[phys].astype(np.float64) * 6 * np.pi
    nf = 2 * g.ng                                     # fine cells per side (level 2)
    fc = np.floor(np.mod(pos, g.ng) * 2).astype(np.int64) % nf
    fid = (fc[:, 2] * nf + fc[:, 1]) * nf + fc[:, 0]
    rho = np.bincount(fid, minlength=nf ** 3).reshape(nf, nf, nf).astype(np.float32)
    ppc = 4                                           # fine cells per PP cell side at level 2
    pc = fc // ppc
    npc = nf // ppc
    pid = (pc[:, 2] * npc + pc[:, 1]) * npc + pc[:, 0]
    pcount = np.bincount(pid, minlength=npc ** 3)
    sparse = pcount[pid] < occupancy
    cells, inv = np.unique(fid[sparse], return_inverse=True)
    cnt = np.bincount(inv)
    target = np.stack([np.bincount(inv, weights=acc[sparse, d]) for d in range(3)], 1) / cnt[:, None]
    cz, cy, cx = cells // (nf * nf), (cells // nf) % nf, cells % nf
    o = np.arange(-4, 5)
    iz = (cz[:, None, None, None] + o[:, None, None]) % nf
    iy = (cy[:, None, None, None] + o[None, :, None]) % nf
    ix = (cx[:, None, None, None] + o[None, None, :]) % nf
    patches = rho[iz, iy, ix]
    return patches.astype(np.float32), target.astype(np.float32), inv, acc[sparse]


def train(net: LearnedPP, patches, target, steps: int = 1500, batch: int = 256, lr: float = 1e-3,
          seed: int = 0, clip: float | None = None, log=None):
    """Adam on MSE of accelerations in units of their RMS (optionally
    clipped at ``clip``: close pairs give a heavy tail that no density
    stencil predicts; measured, clipping at 2 or 4 RMS did worse on held-out
    particles than none); returns the loss history."""
    torch = net.torch
    rng = np.random.default_rng(seed)
    net.y_scale = float(np.sqrt((target ** 2).mean()))
    X = torch.tensor(net.features(patches), device=net.dev)
    y = target / net.y_scale
    Y = torch.tensor(np.clip(y, -clip, clip) if clip else y, device=net.dev)
    hist = []
    for s in range(steps):
        idx = torch.tensor(rng.integers(0, X.shape[0], batch), device=net.dev)
        out = net.forward(X[idx].reshape(batch, PATCH, PATCH, PATCH, 1).contiguous())
        loss, d = net.mse(out, Y[idx])
        net.step(net.backward(d), lr=lr * (0.1 ** (s / steps)))
        hist.append(loss)
        if log and s % 250 == 0:
            log(f"step {s}: loss {loss:.4f}")
    return hist


def evaluate(net: LearnedPP, patches, target, inv, acc_particles):
    """Relative RMS errors against plain PP of (cells) the predicted cell
    accelerations and (particles) the accelerations they give the cells'
    particles; the method without the network (no PP force in these cells)
    scores 1 on both."""
    pred = net.predict(patches).cpu().numpy().astype(np.float64)
    rel = lambda e, t: float(np.sqrt((e ** 2).sum(1).mean() / (t ** 2).sum(1).mean()))
    return dict(cells=rel(pred - target, target), particles=rel(pred[inv] - acc_particles,
                                                                  acc_particles))
