"""geo_gemm_bf16_noalloc: the split-K GEMM with its partial sums in the
library's compiled-in scratch array instead of stream-ordered allocations.

It must give the bits of geo_gemm_bf16 (same kernels, split rule and
summation order) on two calls in a row, the second smaller, through the C
ABI; both are checked against an FP64 product of the same BF16 operands
(rtol 2e-5 of the operand scale, the GEMM's FP32-accumulation bar).  The
shapes exercise the largest split the rule picks (512 split x tile blocks:
the full 8 MiB scratch), a smaller split, and an unsplit product.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev16(torch, bits):
    return torch.as_tensor(np.ascontiguousarray(bits).view(np.int16), device="cuda")


def _run(lib, fn, torch, A, B, M, N, K, a_rs, a_cs, b_rs, b_cs, bias, accumulate, start):
    from ionosphere_fdtd_amd import _native

    out32 = torch.as_tensor(start, device="cuda").clone()
    out16 = torch.zeros((M, N), dtype=torch.int16, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _native.check(getattr(lib, fn)(M, N, K, A.data_ptr(), a_rs, a_cs, B.data_ptr(), b_rs, b_cs,
                                   bias.data_ptr(), out32.data_ptr(), out16.data_ptr(), N,
                                   accumulate, stream), lib)
    torch.cuda.synchronize()
    return out32.cpu().numpy(), out16.cpu().numpy()


@pytest.mark.parametrize("shapes", [
    # (M, N, K, transposed A, accumulate): 256 x 256 over K = 65536 splits into
    # 32 pieces of 2048 (16 tiles x 32 = 512 blocks, the whole scratch); the
    # second call is smaller (8 splits of 1024 over 2 tiles), the third unsplit
    ((256, 256, 65536, True, 0), (100, 64, 8192, False, 1), (70, 45, 37, False, 0)),
])
def test_noalloc_gemm_equals_allocating_gemm(shapes):
    import torch

    from ionosphere_fdtd_amd import _native
    from ionosphere_fdtd_amd.surrogate import bf16_bits, bf16_to_f32

    lib = _native.load_library()
    for M, N, K, ta, accumulate in shapes:
        print("noalloc gemm case", M, N, K, "transposed A" if ta else "", accumulate, flush=True)
        rng = np.random.default_rng(M * 7 + N * 3 + K)
        a16 = bf16_bits(rng.standard_normal((M, K)).astype(np.float32))
        b16 = bf16_bits(rng.standard_normal((K, N)).astype(np.float32))
        bias_np = rng.standard_normal(N).astype(np.float32)
        start = rng.standard_normal((M, N)).astype(np.float32)
        A = _dev16(torch, a16.T.copy() if ta else a16)
        B = _dev16(torch, b16)
        bias = torch.as_tensor(bias_np, device="cuda")
        a_rs, a_cs = (1, M) if ta else (K, 1)
        args = (torch, A, B, M, N, K, a_rs, a_cs, N, 1, bias, accumulate, start)
        want32, want16 = _run(lib, "geo_gemm_bf16", *args)
        got32, got16 = _run(lib, "geo_gemm_bf16_noalloc", *args)
        assert np.array_equal(got32, want32)
        assert np.array_equal(got16, want16)
        exact = (bf16_to_f32(a16).astype(np.float64) @ bf16_to_f32(b16).astype(np.float64)
                 + bias_np + (start if accumulate else 0.0))
        np.testing.assert_allclose(got32, exact, rtol=0, atol=2e-5 * np.sqrt(K) + 1e-5)


def test_noalloc_gemm_rejects_bad_arguments():
    import torch

    from ionosphere_fdtd_amd import _native

    lib = _native.load_library()
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.geo_gemm_bf16_noalloc(0, 4, 4, None, 4, 1, None, 4, 1, None, None, None, 4, 0,
                                   stream)
    assert rc == -1
    assert b"bad sizes" in lib.geo_last_error()
