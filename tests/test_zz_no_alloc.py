"""The allocation-free product (cube2_gemm_bf16_noalloc_dev) through the C
ABI against cube2_gemm_bf16_dev: bit-identical on two calls in a row, the
second smaller (both split over K, so both use the static scratch), on a
shape that is not split, and the refusal of a split shape whose partials
exceed the static scratch."""
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


def _bf16(torch, rng, shape):
    a = rng.standard_normal(shape).astype(np.float32)
    bits = (a.view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    return torch.tensor(bits, device="cuda"), (bits.view(np.uint16).astype(np.uint32) << 16).view(np.float32)


def _run(torch, _lib, fn, A, B, M, N, K, acc, c0):
    C = c0.clone()
    rc = getattr(_lib.lib(), fn)(A.data_ptr(), K, 1, B.data_ptr(), N, 1, C.data_ptr(), N, M, N, K, acc)
    return rc, C.cpu().numpy()


# (M, N, K): split over K (few tiles, long K), then a smaller split shape,
# then an unsplit one
SHAPES = [(864, 32, 20011), (100, 32, 9000), (130, 96, 300)]


def test_noalloc_gemm_equals_allocating_gemm(torch_lib):
    torch, _lib = torch_lib
    rng = np.random.default_rng(11)
    for M, N, K in SHAPES:
        A, Ah = _bf16(torch, rng, (M, K))
        B, Bh = _bf16(torch, rng, (K, N))
        c0 = torch.tensor(rng.standard_normal((M, N)).astype(np.float32), device="cuda")
        for acc in (0, 1):
            print(f"gemm_noalloc M={M} N={N} K={K} acc={acc}", flush=True)
            rc1, want = _run(torch, _lib, "cube2_gemm_bf16_dev", A, B, M, N, K, acc, c0)
            rc2, got = _run(torch, _lib, "cube2_gemm_bf16_noalloc_dev", A, B, M, N, K, acc, c0)
            _lib.check(rc1)
            _lib.check(rc2)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
            ref = Ah.astype(np.float64) @ Bh.astype(np.float64) + (c0.cpu().numpy() if acc else 0)
            assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()


def test_noalloc_gemm_refuses_partials_beyond_the_scratch(torch_lib):
    torch, _lib = torch_lib
    M, N, K = 1000, 64, 100000  # 16 tiles, split 32 ways: 8.2 MB of partials
    A = torch.zeros((M, K), dtype=torch.int16, device="cuda")
    B = torch.zeros((K, N), dtype=torch.int16, device="cuda")
    C = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    rc = _lib.lib().cube2_gemm_bf16_noalloc_dev(A.data_ptr(), K, 1, B.data_ptr(), N, 1,
                                                C.data_ptr(), N, M, N, K, 0)
    assert rc == _lib.ERR_CAPACITY
    assert b"scratch" in _lib.lib().cube2_last_error()
