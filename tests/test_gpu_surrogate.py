"""Trace-surrogate kernels on the MI355X against NumPy and the CPU oracle.

Bars: the BF16 MFMA GEMM equals an FP64 product of the same BF16 operands
within FP32 accumulation error (rtol 2e-5 of the operand scale); layout and
reduction kernels are exact or FP32-rounded; the network's gradients match
the oracle (same BF16 rounding points, FP64 accumulation) within 2e-3
relative L2 per tensor -- BF16 rounding of values that FP32 and FP64
accumulation place on different sides of a rounding boundary.
"""

from __future__ import annotations

import importlib.util

import numpy as np
import pytest

import scenarios as S

pytestmark = pytest.mark.gpu


def _oracle():
    spec = importlib.util.spec_from_file_location("surrogate_oracle",
                                                  S.ORACLE_DIR / "surrogate_oracle.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _dev16(torch, bits):
    return torch.as_tensor(np.ascontiguousarray(bits).view(np.int16), device="cuda")


@pytest.mark.parametrize("M,N,K,ta,tb", [(70, 45, 37, False, False), (128, 64, 64, True, False),
                                          (1, 1, 1, False, True), (300, 17, 96, True, True),
                                          (64, 256, 512, False, False),
                                          (80, 16, 40000, True, False)])  # split K
def test_gemm_bf16_matches_fp64_product_of_bf16_operands(M, N, K, ta, tb):
    import torch

    from ionosphere_fdtd_amd import _native
    from ionosphere_fdtd_amd.surrogate import bf16_bits, bf16_to_f32

    lib = _native.load_library()
    rng = np.random.default_rng(M * 1000 + N + K)
    a = rng.standard_normal((M, K)).astype(np.float32)
    b = rng.standard_normal((K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    a16, b16 = bf16_bits(a), bf16_bits(b)
    want = bf16_to_f32(a16).astype(np.float64) @ bf16_to_f32(b16).astype(np.float64) + bias
    # operands stored transposed are read through strides
    A = _dev16(torch, a16.T.copy() if ta else a16)
    B = _dev16(torch, b16.T.copy() if tb else b16)
    a_rs, a_cs = (1, M) if ta else (K, 1)
    b_rs, b_cs = (1, K) if tb else (N, 1)
    out32 = torch.full((M, N), 3.0, dtype=torch.float32, device="cuda")
    out16 = torch.zeros((M, N), dtype=torch.int16, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _native.check(lib.geo_gemm_bf16(M, N, K, A.data_ptr(), a_rs, a_cs, B.data_ptr(), b_rs, b_cs,
                                    torch.as_tensor(bias, device="cuda").data_ptr(),
                                    out32.data_ptr(), out16.data_ptr(), N, 1, stream), lib)
    got = out32.cpu().numpy().astype(np.float64) - 3.0
    scale = np.sqrt(K) * 1.0
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * scale + 1e-5)
    # the BF16 output is the FP32 result rounded once (relative 2^-9)
    got16 = bf16_to_f32(out16.cpu().numpy().view(np.uint16)).astype(np.float64)
    assert np.all(np.abs(got16 - want) <= 2.0 ** -8 * np.abs(want) + 2e-5 * scale + 1e-5)


def test_gemm_taller_than_one_launch_runs_in_row_chunks():
    """A product with more than 65535 row tiles of 64 (grid.y) runs as row
    chunks of 32767 x 128 rows: rows on both sides of a chunk boundary and in
    the short last chunk equal the same rows computed alone, bit for bit, and
    the whole product matches the FP64 product of the BF16 operands."""
    import torch

    from ionosphere_fdtd_amd import _native
    from ionosphere_fdtd_amd.surrogate import bf16_bits, bf16_to_f32

    lib = _native.load_library()
    stream = torch.cuda.current_stream().cuda_stream
    chunk = 32767 * 128
    M, N, K = chunk + 200, 24, 40
    rng = np.random.default_rng(3)
    a16 = bf16_bits(rng.standard_normal((M, K), dtype=np.float32))
    b16 = bf16_bits(rng.standard_normal((K, N), dtype=np.float32))
    A, B = _dev16(torch, a16), _dev16(torch, b16)
    bias = torch.as_tensor(rng.standard_normal(N, dtype=np.float32), device="cuda")
    full = torch.empty((M, N), dtype=torch.float32, device="cuda")
    _native.check(lib.geo_gemm_bf16(M, N, K, A.data_ptr(), K, 1, B.data_ptr(), N, 1,
                                    bias.data_ptr(), full.data_ptr(), None, N, 0, stream), lib)
    got = full.cpu().numpy()
    for m0, rows in ((chunk - 100, 200), (chunk, 200), (0, 64)):
        part = torch.empty((rows, N), dtype=torch.float32, device="cuda")
        _native.check(lib.geo_gemm_bf16(rows, N, K, A.data_ptr() + 2 * m0 * K, K, 1, B.data_ptr(),
                                        N, 1, bias.data_ptr(), part.data_ptr(), None, N, 0,
                                        stream), lib)
        assert part.cpu().numpy().tobytes() == got[m0:m0 + rows].tobytes(), m0
    want = (bf16_to_f32(a16).astype(np.float64) @ bf16_to_f32(b16).astype(np.float64)
            + bias.cpu().numpy())
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * np.sqrt(K) + 1e-5)


@pytest.mark.parametrize("M,N,K", [(2056, 2056, 77), (200, 136, 77), (129, 300, 96),
                                   (1000, 130, 40)])
def test_gemm_tile_paths_and_operand_layouts_give_the_same_bits(M, N, K):
    """Every operand layout (K contiguous, row contiguous, a stride of 2 that
    takes the element-wise loads) and both block tiles: a product with M, N
    >= 128, K > 64 and at least 256 tiles of 128 x 128 runs on them (2056 x
    2056: 289 tiles, the last row and column ragged), the same rows in chunks
    of 64 on 64 x 64 tiles, as every smaller product; each output tile
    accumulates the same K slices in the same order, so all must agree bit
    for bit, and with the FP64 product of the BF16 operands to FP32
    accumulation error.  Outputs with a row stride that is a multiple of 4
    take the staged 16-byte row stores, the others (N = 130) the per-element
    ones."""
    import torch

    from ionosphere_fdtd_amd import _native
    from ionosphere_fdtd_amd.surrogate import bf16_bits, bf16_to_f32

    lib = _native.load_library()
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(M + 7 * N + 13 * K)
    a16 = bf16_bits(rng.standard_normal((M, K)).astype(np.float32))
    b16 = bf16_bits(rng.standard_normal((K, N)).astype(np.float32))
    want = bf16_to_f32(a16).astype(np.float64) @ bf16_to_f32(b16).astype(np.float64)

    def operand(x16, layout):
        # (device buffer, row stride, column stride) of the (rows, cols) matrix x
        rows, cols = x16.shape
        if layout == "row-major":
            return _dev16(torch, x16), cols, 1
        if layout == "col-major":
            return _dev16(torch, x16.T.copy()), 1, rows
        wide = np.zeros((rows, 2 * cols), dtype=np.uint16)  # every other column
        wide[:, ::2] = x16
        return _dev16(torch, wide), 2 * cols, 2

    results = []
    for la in ("row-major", "col-major", "strided"):
        for lb in ("row-major", "col-major", "strided"):
            A, a_rs, a_cs = operand(a16, la)
            B, b_rs, b_cs = operand(b16, lb)
            full = torch.empty((M, N), dtype=torch.float32, device="cuda")
            _native.check(lib.geo_gemm_bf16(M, N, K, A.data_ptr(), a_rs, a_cs, B.data_ptr(), b_rs,
                                            b_cs, None, full.data_ptr(), None, N, 0, stream), lib)
            chunks = torch.empty((M, N), dtype=torch.float32, device="cuda")
            for m0 in range(0, M, 64):
                rows = min(64, M - m0)
                off = (m0 * a_rs) * 2  # bytes
                _native.check(lib.geo_gemm_bf16(rows, N, K, A.data_ptr() + off, a_rs, a_cs,
                                                B.data_ptr(), b_rs, b_cs, None,
                                                chunks[m0:].data_ptr(), None, N, 0, stream), lib)
            got = full.cpu().numpy()
            assert got.tobytes() == chunks.cpu().numpy().tobytes(), (la, lb)
            np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * np.sqrt(K) + 1e-5)
            results.append(got)
    for r in results[1:]:
        assert r.tobytes() == results[0].tobytes()


def test_layout_reduction_and_optimizer_kernels():
    import torch

    from ionosphere_fdtd_amd import _native
    from ionosphere_fdtd_amd.surrogate import bf16_bits, bf16_to_f32

    so = _oracle()
    lib = _native.load_library()
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(3)
    # im2col with upsampling, and its adjoint
    x = rng.standard_normal((3, 10, 6)).astype(np.float32)
    x16 = bf16_bits(x)
    cols = torch.empty((3 * 20, 30), dtype=torch.int16, device="cuda")
    _native.check(lib.geo_im2col_up(3, 10, 6, 2, 5, _dev16(torch, x16).data_ptr(),
                                    cols.data_ptr(), stream), lib)
    want = so.im2col_up(bf16_to_f32(x16).astype(np.float64), 2, 5)
    np.testing.assert_array_equal(bf16_to_f32(cols.cpu().numpy().view(np.uint16)), want)
    dc = rng.standard_normal((60, 30)).astype(np.float32)
    dx = torch.empty((3, 10, 6), dtype=torch.float32, device="cuda")
    _native.check(lib.geo_col2im_up(3, 10, 6, 2, 5, torch.as_tensor(dc, device="cuda").data_ptr(),
                                    dx.data_ptr(), stream), lib)
    np.testing.assert_allclose(dx.cpu().numpy(), so.col2im_up(dc.astype(np.float64), 3, 10, 6,
                                                                 2, 5), rtol=1e-5, atol=1e-5)
    # SiLU forward / backward
    pre = rng.standard_normal(1000).astype(np.float32) * 4
    dout = rng.standard_normal(1000).astype(np.float32)
    pre_d = torch.as_tensor(pre, device="cuda")
    out = torch.empty(1000, dtype=torch.int16, device="cuda")
    _native.check(lib.geo_silu_forward(1000, pre_d.data_ptr(), out.data_ptr(), stream), lib)
    ref = so.silu(pre.astype(np.float64))
    np.testing.assert_allclose(bf16_to_f32(out.cpu().numpy().view(np.uint16)), ref,
                               rtol=8e-3, atol=1e-6)
    dpre = torch.empty(1000, dtype=torch.float32, device="cuda")
    _native.check(lib.geo_silu_backward(1000, pre_d.data_ptr(),
                                        torch.as_tensor(dout, device="cuda").data_ptr(),
                                        dpre.data_ptr(), None, stream), lib)
    np.testing.assert_allclose(dpre.cpu().numpy(), dout * so.silu_grad(pre.astype(np.float64)),
                               rtol=1e-5, atol=1e-6)
    # column sums
    m = rng.standard_normal((777, 70)).astype(np.float32)
    cs = torch.empty(70, dtype=torch.float32, device="cuda")
    _native.check(lib.geo_column_sums(777, 70, torch.as_tensor(m, device="cuda").data_ptr(), 70,
                                      cs.data_ptr(), stream), lib)
    np.testing.assert_allclose(cs.cpu().numpy(), m.astype(np.float64).sum(axis=0), rtol=1e-5,
                               atol=1e-4)
    # the training shapes (262 144 x 16 bias gradients), one column, more than
    # 256 columns, a row pitch above N; the same bits on a second call
    for rows, cols, ld in ((262144, 16, 16), (300001, 1, 1), (1000, 300, 300), (5, 3, 7)):
        m = rng.standard_normal((rows, ld)).astype(np.float32)
        md_ = torch.as_tensor(m, device="cuda")
        got = []
        for _ in range(2):
            cs = torch.empty(cols, dtype=torch.float32, device="cuda")
            _native.check(lib.geo_column_sums(rows, cols, md_.data_ptr(), ld, cs.data_ptr(),
                                              stream), lib)
            got.append(cs.cpu().numpy())
        want = m[:, :cols].astype(np.float64).sum(axis=0)
        np.testing.assert_allclose(got[0], want, rtol=0, atol=2e-6 * np.sqrt(rows) * 4)
        assert got[0].tobytes() == got[1].tobytes()
    # Adam: one step from zero moments moves every parameter by -lr * sign(g)
    p = rng.standard_normal(500).astype(np.float32)
    g = rng.standard_normal(500).astype(np.float32)
    pd = torch.as_tensor(p.copy(), device="cuda")
    md = torch.zeros(500, device="cuda")
    vd = torch.zeros(500, device="cuda")
    _native.check(lib.geo_adam_step(500, pd.data_ptr(), torch.as_tensor(g, device="cuda")
                                    .data_ptr(), md.data_ptr(), vd.data_ptr(), None, 1e-2, 0.9,
                                    0.999, 1e-12, 1, stream), lib)
    np.testing.assert_allclose(pd.cpu().numpy(), p - 1e-2 * np.sign(g), rtol=0, atol=1e-6)


def test_network_gradients_match_cpu_oracle():
    import torch

    from ionosphere_fdtd_amd.surrogate import SurrogateConfig, TraceSurrogate, initial_parameters

    so = _oracle()
    cfg = SurrogateConfig()
    rng = np.random.default_rng(17)
    params = initial_parameters(cfg, seed=5)
    for k in params:
        if not k.startswith("W"):
            params[k] = (0.1 * rng.standard_normal(params[k].shape)).astype(np.float32)
    net = TraceSurrogate(cfg, params)
    B = 6
    x = rng.standard_normal((B, cfg.in_features)).astype(np.float32)
    st = rng.standard_normal((B, cfg.length)).astype(np.float32)
    at = rng.standard_normal(B).astype(np.float32)
    loss = net.loss_and_grad(torch.as_tensor(x, device="cuda"), torch.as_tensor(st, device="cuda"),
                             torch.as_tensor(at, device="cuda"))
    shape, amp, cache = so.forward(params, x.astype(np.float64), cfg, True)
    want_loss, ds, da = so.trace_loss(shape, amp, st, at, cfg.amp_weight)
    grads = so.backward(cache, ds, da, cfg, True)
    print(f"loss gpu {loss:.6f} oracle {want_loss:.6f}", flush=True)
    assert abs(loss - want_loss) <= 1e-3 * abs(want_loss)
    got_shape = net._cache.shape.cpu().numpy()
    assert S.rel_l2(got_shape, shape) <= 1e-3
    worst = 0.0
    for name, g in grads.items():
        got = net.grads[name].cpu().numpy().reshape(g.shape)
        err = S.rel_l2(got, g)
        worst = max(worst, err)
        assert err <= 2e-3, (name, err)
    print(f"worst gradient rel-L2 {worst:.2e}", flush=True)


def test_adam_training_reduces_the_loss_of_a_fixed_batch():
    import torch

    from ionosphere_fdtd_amd.surrogate import TraceSurrogate

    net = TraceSurrogate(seed=1)
    rng = np.random.default_rng(2)
    B, L = 32, net.cfg.length
    x = torch.as_tensor(rng.standard_normal((B, net.cfg.in_features)).astype(np.float32),
                        device="cuda")
    t = np.linspace(0, 1, L)
    st = torch.as_tensor(np.stack([np.sin(2 * np.pi * (1 + i % 3) * t) for i in range(B)])
                         .astype(np.float32), device="cuda")
    at = torch.as_tensor(rng.standard_normal(B).astype(np.float32), device="cuda")
    losses = [net.train_step(x, st, at, lr=2e-3) for _ in range(60)]
    print("loss", losses[0], "->", losses[-1], flush=True)
    assert losses[-1] < 0.2 * losses[0]
