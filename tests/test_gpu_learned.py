"""The learned PP stand-in on the GPU: the BF16 MFMA GEMM against exact
integer products (operand and result lane layouts, transposed strides, ragged
sizes), the network's forward and backward against the FP64 oracle, and a
short training on plain-PP pairs that must beat the method without the
network on a held-out box (plain PP judges)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_lib():
    import torch
    assert torch.cuda.is_available()
    from cube2_amd import _lib
    _lib.lib()
    return torch, _lib


def _bf16_round(a):
    """FP64 values rounded to BF16 (to nearest even, through FP32)."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _bf16_bits(a):
    """Exact BF16 bit patterns of values representable in BF16."""
    u = np.asarray(a, np.float32).view(np.uint32)
    assert np.all((u & 0xFFFF) == 0)
    return (u >> 16).astype(np.uint16).view(np.int16)


@pytest.mark.parametrize("M,N,K,ta,tb", [(16, 16, 16, 0, 0), (37, 50, 29, 0, 0), (64, 3, 864, 0, 0),
                                         (45, 33, 70, 1, 0), (20, 70, 40, 0, 1), (130, 96, 17, 1, 1),
                                         (864, 32, 20011, 1, 0), (27, 32, 9000, 1, 0)])
def test_gemm_bf16_exact_on_integers(torch_lib, M, N, K, ta, tb):
    """(the last two shapes take the split-K path: few output tiles, long K)"""
    torch, _lib = torch_lib
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.integers(-8, 9, (M, K)).astype(np.float32)
    B = rng.integers(-8, 9, (K, N)).astype(np.float32)
    a_store = A.T.copy() if ta else A
    b_store = B.T.copy() if tb else B
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    da = torch.tensor(_bf16_bits(a_store), device="cuda")
    db = torch.tensor(_bf16_bits(b_store), device="cuda")
    C = torch.full((M, N), 5.0, dtype=torch.float32, device="cuda")
    print(f"gemm M={M} N={N} K={K} transA={ta} transB={tb}", flush=True)
    _lib.check(_lib.lib().cube2_gemm_bf16_dev(da.data_ptr(), sam, sak, db.data_ptr(), sbk, sbn,
                                              C.data_ptr(), N, M, N, K, 0))
    assert np.array_equal(C.cpu().numpy(), A @ B)
    _lib.check(_lib.lib().cube2_gemm_bf16_dev(da.data_ptr(), sam, sak, db.data_ptr(), sbk, sbn,
                                              C.data_ptr(), N, M, N, K, 1))
    assert np.array_equal(C.cpu().numpy(), 2 * (A @ B))


def _rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def test_forward_backward_match_fp64_oracle(torch_lib):
    """The network (matrix-core convolutions forward, BF16 GEMMs backward,
    the CIC-combined MSE) against the FP64 oracle with the same BF16
    weights: outputs, loss and every layer's gradients."""
    torch, _lib = torch_lib
    from cube2_amd.learned import LearnedPP
    from oracle import cnn
    rng = np.random.default_rng(4)
    params = [(W, rng.normal(0, 0.05, b.shape)) for W, b in cnn.init_params(7)]
    net = LearnedPP(params)
    x = rng.uniform(-0.5, 2, (6, 10, 10, 10, 4)).astype(np.float32)
    t = rng.normal(size=(6, 3)).astype(np.float32)
    frac = rng.uniform(0, 1, (6, 3)).astype(np.float32)
    out = net.forward(torch.tensor(x, device="cuda"))
    loss, d, pred = net.cic_mse(out, torch.tensor(frac, device="cuda"), torch.tensor(t, device="cuda"))
    grads = net.backward(d)
    # the oracle with the network's BF16 weights and BF16-rounded input, in FP64
    pq = [(_bf16_round(w), b) for (w, b) in net.params_numpy()]
    o_out, cache = cnn.forward(pq, _bf16_round(x))
    o_pred = cnn.cic(o_out, frac.astype(np.float64))
    o_loss, o_dp = cnn.mse(o_pred, t.astype(np.float64))
    w = cnn.cic_weights(frac)
    o_grads = cnn.backward(pq, cache, (w[:, :, None] * o_dp[:, None, :]).reshape(6, 24))
    e_out = _rel(out.cpu().numpy().reshape(6, 24), o_out)
    e_pred = _rel(pred.cpu().numpy(), o_pred)
    print(f"forward rel RMS {e_out:.2e}, CIC prediction {e_pred:.2e}; loss {loss:.5f} vs "
          f"{o_loss:.5f}", flush=True)
    assert e_out < 2e-2 and e_pred < 2e-2
    assert abs(loss - o_loss) <= 4e-2 * o_loss
    for l in range(4):
        eW = _rel(grads[l][0].cpu().numpy(), o_grads[l][0])
        eb = _rel(grads[l][1].cpu().numpy(), o_grads[l][1])
        print(f"layer {l}: dW rel RMS {eW:.2e}, db rel RMS {eb:.2e}", flush=True)
        assert eW < 5e-2 and eb < 5e-2


SMALL = dict(npd=54, n_total=64 ** 3, box_mpc=75.0)   # C1's particle density, 1/64 of its volume


def _small_geom():
    from cube2_amd.geometry import Geometry
    return Geometry(ng=64, nnt=2, nns=2, ratio_cs=4, ngb=8, app=0.05)


def test_learned_kick_matches_training_path_and_beats_no_network(torch_lib):
    """Four C1-style boxes (C1's particle density at 1/64 of its volume, z =
    10, 3, 1, 0), samples of the particles the network stands in for, 1500
    Adam steps.  On a held-out box (seed 99, z = 0), with PP alone:
    * the learned kick (CUBE2_KICK_LEARNED: the network on each level-2
      subtile's grid, the pair kernel skipping those particles) gives the
      substituted particles the training path's prediction (gathered
      patches) to FP32 rounding, and every other particle the plain kick's
      Delta-v bit for bit;
    * its error against plain PP on the substituted particles beats the
      method without the network (no PP force for them: relative RMS 1)."""
    torch, _lib = torch_lib
    from cube2_amd.kick import ShortRangeKick
    from cube2_amd.learned import LearnedPP, TrainingSet, box_samples, c1_style_box, train
    g = _small_geom()
    sets = []
    for seed, z in ((10, 10.0), (11, 3.0), (12, 1.0), (13, 0.0)):
        ts, _ = box_samples(g, c1_style_box(seed, z, g, **SMALL), max_samples=60000, seed=seed)
        sets.append(ts)
    ts = TrainingSet.concat(sets)
    print(f"training samples: {ts.n}", flush=True)
    net = LearnedPP(seed=0)
    hist = train(net, ts, steps=1500, log=lambda m: print(m, flush=True))
    hold, kp = box_samples(g, c1_style_box(99, 0.0, g, **SMALL))
    pred = net.predict(hold).cpu().numpy().astype(np.float64)
    plain_vp = kp.d_vp.clone()
    net.load()
    kl = ShortRangeKick.from_device(g, kp.d_rhoc, kp.d_idx, kp.d_xp, torch.zeros_like(kp.d_vp),
                                    kp.n_used)
    kl.set_levels(np.full(g.ns3, 2, np.int32))
    _lib.set_pp_mode(_lib.PP_PRECISE)
    try:
        rl = kl.kick(1.0, 1.0, 1.0, pm3=False, pp=True, learned=True)
    finally:
        _lib.set_pp_mode(_lib.PP_REF)
        _lib.lib().cube2_learned_clear()
    v_l = kl.d_vp.cpu().numpy().astype(np.float64)
    v_p = plain_vp.cpu().numpy().astype(np.float64)
    # the substituted particles: the held-out box's samples (all of them)
    slots = hold.slot.cpu().numpy().astype(np.int64)
    sub = np.zeros(v_l.shape[0], bool)
    sub[slots] = True
    a_kick = v_l[slots] * 6 * np.pi
    a_plain = hold.target.cpu().numpy().astype(np.float64)
    e_paths = _rel(a_kick, pred)
    e_net = _rel(a_kick, a_plain)
    same = np.array_equal(v_l[~sub].view(np.uint64), v_p[~sub].view(np.uint64))
    print(f"held-out: {hold.n} substituted particles of {int((v_p != 0).any(1).sum())} kicked; "
          f"kick vs training path {e_paths:.2e}; error vs plain PP {e_net:.3f} (no network: 1); "
          f"others bit-identical: {same}; P_acc {rl.p_acc}; loss {np.mean(hist[:50]):.3f} -> "
          f"{np.mean(hist[-100:]):.3f}", flush=True)
    assert e_paths < 1e-5
    assert same
    assert e_net < 0.95


def _pack_and_conv(torch, _lib, X, W, b, cin, cout, relu, x_bf16=False, y_bf16=False):
    """cube2_cnn_conv_pack_dev + cube2_cnn_conv3_dev on device copies."""
    B, D = X.shape[0], X.shape[1]
    P = D - 2
    L = _lib.lib()
    dW = torch.tensor(W.astype(np.float32), device="cuda")
    taps = 27 if cin == 32 else 32
    coutp = 16 if cout == 3 else 32
    Wp = torch.empty(taps * coutp * cin, dtype=torch.int16, device="cuda")
    _lib.check(L.cube2_cnn_conv_pack_dev(dW.data_ptr(), Wp.data_ptr(), cin, cout))
    if x_bf16:
        dX = torch.tensor(_bf16_bits(X.astype(np.float32)), device="cuda")
    else:
        dX = torch.tensor(X.astype(np.float32), device="cuda")
    db = torch.tensor(b.astype(np.float32), device="cuda")
    Y = torch.empty((B, P, P, P, cout), dtype=torch.int16 if y_bf16 else torch.float32, device="cuda")
    _lib.check(L.cube2_cnn_conv3_dev(dX.data_ptr(), int(x_bf16), Wp.data_ptr(), db.data_ptr(),
                                     Y.data_ptr(), int(y_bf16), B, D, cin, cout, int(relu)))
    y = Y.cpu().numpy()
    if y_bf16:
        y = (y.view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    return y


@pytest.mark.parametrize("B,D,cin,cout,relu", [(3, 10, 4, 32, 1), (2, 9, 32, 32, 1), (2, 5, 32, 3, 0),
                                               (1, 3, 32, 32, 0), (1, 19, 4, 32, 0), (2, 12, 32, 3, 0)])
def test_conv3_exact_on_integers(torch_lib, B, D, cin, cout, relu):
    """The MFMA convolution with integer operands (exact in BF16, sums exact
    in FP32) equals the oracle's im2col product exactly: the lane layout of
    v_mfma_f32_16x16x32_bf16, the tap/channel order of K, the block tails."""
    torch, _lib = torch_lib
    from oracle import cnn
    rng = np.random.default_rng(B * 100 + D * 10 + cin + cout)
    X = rng.integers(-4, 5, (B, D, D, D, cin)).astype(np.float64)
    W = rng.integers(-4, 5, (27 * cin, cout)).astype(np.float64)
    b = rng.integers(-8, 9, cout).astype(np.float64)
    print(f"conv3 B={B} D={D} {cin}->{cout} relu={relu}", flush=True)
    got = _pack_and_conv(torch, _lib, X, W, b, cin, cout, relu)
    want = cnn.im2col(X) @ W + b
    if relu:
        want = np.maximum(want, 0.0)
    assert np.array_equal(got.reshape(-1, cout), want)


def _conv_ref(X, W, b, relu):
    """Valid 3^3 convolution in FP64 as 27 tap products (no im2col)."""
    B, D, cin = X.shape[0], X.shape[1], X.shape[4]
    P = D - 2
    cout = W.shape[1]
    y = np.zeros((B, P, P, P, cout)) + b
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                t = (dz * 3 + dy) * 3 + dx
                y += X[:, dz:dz + P, dy:dy + P, dx:dx + P, :] @ W[t * cin:(t + 1) * cin]
    return np.maximum(y, 0.0) if relu else y


@pytest.mark.parametrize("B,D,cin,cout,xb,yb", [(150, 10, 32, 32, 0, 0), (8, 26, 4, 32, 0, 1),
                                               (6, 26, 32, 32, 1, 1), (9, 26, 32, 3, 1, 0),
                                               (300, 10, 4, 32, 0, 0), (4, 42, 32, 32, 1, 1)])
def test_conv3_many_blocks_per_workgroup(torch_lib, B, D, cin, cout, xb, yb):
    """Grids of more output blocks than resident workgroups (the persistent
    loop runs several blocks per workgroup), every input/output type the
    learned kick and training use; integer operands, so the FP32 results are
    exact and the BF16 ones the BF16 rounding of the exact sums."""
    torch, _lib = torch_lib
    rng = np.random.default_rng(B + D + cin + cout)
    X = rng.integers(-3, 4, (B, D, D, D, cin)).astype(np.float64)
    W = rng.integers(-2, 3, (27 * cin, cout)).astype(np.float64)
    b = rng.integers(-4, 5, cout).astype(np.float64)
    P = D - 2
    print(f"conv3 B={B} D={D} {cin}->{cout} x_bf16={xb} y_bf16={yb}: "
          f"{B * (-(-P // 4)) ** 2 * (-(-P // 8))} blocks", flush=True)
    got = _pack_and_conv(torch, _lib, X, W, b, cin, cout, 1, x_bf16=bool(xb), y_bf16=bool(yb))
    want = _conv_ref(X, W, b, 1)
    if yb:
        want = _bf16_round(want)
    bad = np.argwhere(got != want)
    assert bad.size == 0, (bad[:5], got[tuple(bad[0])], want[tuple(bad[0])])


def test_conv3_bf16_activations_match_fp64(torch_lib):
    """BF16 in and out (the in-kick inference layers) against FP64 with the
    same BF16-rounded operands: within BF16 output rounding."""
    torch, _lib = torch_lib
    from oracle import cnn
    rng = np.random.default_rng(3)
    X = _bf16_round(rng.standard_normal((2, 14, 14, 14, 32)))
    W = _bf16_round(rng.standard_normal((27 * 32, 32)) * 0.1)
    b = rng.standard_normal(32) * 0.1
    got = _pack_and_conv(torch, _lib, X, W, b, 32, 32, 1, x_bf16=True, y_bf16=True)
    want = np.maximum(cnn.im2col(X) @ W + b, 0.0)
    err = _rel(got.reshape(-1, 32), want)
    print(f"conv3 bf16 in/out vs FP64: relative RMS {err:.2e}", flush=True)
    assert err < 4e-3
