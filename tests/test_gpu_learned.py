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
    torch, _lib = torch_lib
    from cube2_amd.learned import LearnedPP
    from oracle import cnn
    rng = np.random.default_rng(4)
    params = [(W, rng.normal(0, 0.05, b.shape)) for W, b in cnn.init_params(7)]
    net = LearnedPP(params)
    x = rng.uniform(0, 2, (6, 9, 9, 9, 1)).astype(np.float32)
    t = rng.normal(size=(6, 3)).astype(np.float32)
    out = net.forward(torch.tensor(x, device="cuda"))
    loss, d = net.mse(out, torch.tensor(t, device="cuda"))
    grads = net.backward(d)
    # the oracle with the network's BF16 weights, in FP64
    pq = [(w.astype(np.float64), b) for (w, b) in net.params_numpy()]
    pq = [(np.asarray(torch.tensor(w, dtype=torch.float32).to(torch.bfloat16).to(torch.float64)), b)
          for w, b in pq]
    o_out, cache = cnn.forward(pq, x.astype(np.float64))
    o_loss, o_d = cnn.mse(o_out, t.astype(np.float64))
    o_grads = cnn.backward(pq, cache, o_d)
    e_out = _rel(out.cpu().numpy(), o_out)
    print(f"forward rel RMS {e_out:.2e}; loss {loss:.5f} vs {o_loss:.5f}", flush=True)
    assert e_out < 2e-2
    assert abs(loss - o_loss) <= 4e-2 * o_loss
    for l in range(4):
        eW = _rel(grads[l][0].cpu().numpy(), o_grads[l][0])
        eb = _rel(grads[l][1].cpu().numpy(), o_grads[l][1])
        print(f"layer {l}: dW rel RMS {eW:.2e}, db rel RMS {eb:.2e}", flush=True)
        assert eW < 5e-2 and eb < 5e-2


def test_training_beats_the_method_without_it(torch_lib):
    """Four C1-style training boxes (z = 3, 1, 0, 0), plain-PP pairs of
    sparse cells, 1500 Adam steps; on the held-out box (seed 99) the
    network's accelerations must beat the method without it (no PP force in
    sparse cells: relative RMS error 1).  Measured on MI355X: cells 0.72,
    particles 0.85 (the close-pair part of PP is not in a count stencil)."""
    torch, _lib = torch_lib
    from cube2_amd.learned import LearnedPP, evaluate, pp_pairs, train, training_box
    xs, ys = [], []
    for seed, z in ((10, 3.0), (11, 1.0), (12, 0.0), (13, 0.0)):
        g, pos = training_box(seed, z, npd=32)
        p, y, _, _ = pp_pairs(g, pos)
        xs.append(p)
        ys.append(y)
    X, Y = np.concatenate(xs), np.concatenate(ys)
    print(f"training pairs: {X.shape[0]}", flush=True)
    net = LearnedPP(seed=0)
    hist = train(net, X, Y, steps=1500, log=lambda m: print(m, flush=True))
    g, pos = training_box(99, 0.0, npd=32)
    p, y, inv, acc = pp_pairs(g, pos)
    r = evaluate(net, p, y, inv, acc)
    print(f"held-out box: {p.shape[0]} sparse cells, {acc.shape[0]} particles; relative RMS "
          f"error with the network {r}, without it 1.0; loss {np.mean(hist[:50]):.3f} -> "
          f"{np.mean(hist[-100:]):.3f}", flush=True)
    assert np.mean(hist[-100:]) < 0.7 * np.mean(hist[:50])
    assert r["cells"] < 0.8 and r["particles"] < 0.92
