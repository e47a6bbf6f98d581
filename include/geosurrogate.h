/*
 * geosurrogate.h -- C ABI of the trace-surrogate kernels (gfx950).
 *
 * No counterpart in the reference (ruddyscent/ionosphere-fdtd has no learned
 * model): these kernels train and run a small network -- a parameter MLP
 * feeding a 1-D convolutional trace decoder -- that predicts receiver E_r
 * traces of the FDTD solver.  Every matrix product of its forward and
 * backward passes goes through geo_gemm_bf16 (BF16 operands on the gfx950
 * MFMA units, FP32 accumulation); the remaining steps are the elementwise,
 * reduction and layout kernels below.  All pointers are device pointers;
 * every call is enqueued on `stream` and returns a GEO_* code (geofdtd.h).
 *
 * BF16 values are passed as uint16 bit patterns (round to nearest even).
 */
#ifndef GEOSURROGATE_H
#define GEOSURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C = A B (+ bias) with A (M x K), B (K x N) BF16, addressed by strides in
 * elements: A(m,k) = A[m*a_rs + k*a_cs], B(k,n) = B[k*b_rs + n*b_cs], so
 * transposed operands need no copy.  FP32 accumulation.  Outputs (row-major,
 * leading dimension ldc): c_f32 (written, or incremented when accumulate is
 * 1) and/or c_bf16; bias (N FP32 values) may be NULL.  Any M, N, K >= 1
 * (a product of more than 32767 x 128 rows runs as row chunks of that many,
 * with the same kernels and bits).
 * A long K with few output tiles is split over blocks; the split-K partials
 * (at most 8 MiB) are a stream-ordered allocation from a pool the library
 * keeps per device, which holds freed memory for the next call. */
int geo_gemm_bf16(int64_t M, int64_t N, int64_t K, const uint16_t* A, int64_t a_rs,
                  int64_t a_cs, const uint16_t* B, int64_t b_rs, int64_t b_cs, const float* bias,
                  float* c_f32, uint16_t* c_bf16, int64_t ldc, int32_t accumulate, void* stream);

/* geo_gemm_bf16 that allocates nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels, split rule and results, bit for bit, as geo_gemm_bf16; the split-K
 * partials go to an 8 MiB FP32 scratch array compiled into the library (one
 * per device, held until the process ends), which covers every split the
 * rule chooses -- a larger request is refused with GEO_ERR_INVALID.  The
 * scratch is shared by all calls on a device and keeps the last call's
 * partials (each call writes every partial it reads), so calls that overlap
 * on one device -- on different streams or host threads -- race. */
int geo_gemm_bf16_noalloc(int64_t M, int64_t N, int64_t K, const uint16_t* A, int64_t a_rs,
                          int64_t a_cs, const uint16_t* B, int64_t b_rs, int64_t b_cs,
                          const float* bias, float* c_f32, uint16_t* c_bf16, int64_t ldc,
                          int32_t accumulate, void* stream);

/* out = bf16(x) for n values. */
int geo_f32_to_bf16(int64_t n, const float* x, uint16_t* out, void* stream);

/* SiLU x * sigmoid(x): out_bf16 = bf16(silu(pre)); backward
 * dpre = dout * sigmoid(pre) * (1 + pre * (1 - sigmoid(pre))) in FP32 and
 * (when dpre_bf16 is not NULL) BF16. */
int geo_silu_forward(int64_t n, const float* pre, uint16_t* out_bf16, void* stream);
int geo_silu_backward(int64_t n, const float* pre, const float* dout, float* dpre,
                      uint16_t* dpre_bf16, void* stream);

/* out[c] = sum over rows r < M of x[r*ld + c], c < N: fixed chunks of about
 * 256 rows summed per block, then the chunk sums in order, so the result is
 * the same bits on every call.  Takes a stream-ordered allocation for the
 * chunk sums from the library's own pool, which keeps freed memory for the
 * next call (as for geo_gemm_bf16's split-K partials). */
int geo_column_sums(int64_t M, int64_t N, const float* x, int64_t ld, float* out, void* stream);

/* Nearest-neighbour upsampling by `up` followed by the im2col of a
 * `ksize`-tap convolution with zero padding ksize/2, channels last:
 * x (batch, length, channels) -> cols (batch * length * up, ksize * channels),
 * cols[(b*Lo + l)*K + j*C + c] = x[b, (l + j - ksize/2) / up, c] when
 * 0 <= l + j - ksize/2 < Lo = length * up, else 0. */
int geo_im2col_up(int64_t batch, int64_t length, int64_t channels, int32_t up, int32_t ksize,
                  const uint16_t* x, uint16_t* cols, void* stream);
/* Adjoint of geo_im2col_up in FP32: dx[b, l', c] = sum of the dcols entries
 * that geo_im2col_up filled from x[b, l', c]. */
int geo_col2im_up(int64_t batch, int64_t length, int64_t channels, int32_t up, int32_t ksize,
                  const float* dcols, float* dx, void* stream);

/* Adam with bias correction (step t >= 1) on FP32 parameters; writes the
 * BF16 copy used by the GEMMs when param_bf16 is not NULL. */
int geo_adam_step(int64_t n, float* param, const float* grad, float* m, float* v,
                  uint16_t* param_bf16, float lr, float beta1, float beta2, float eps,
                  int64_t t, void* stream);

/* Trace loss of a batch: shape (batch, length) and log-amplitude (batch,
 * amp_ld-strided) predictions against targets.
 * loss = mean_b [ mean_l (s - s_t)^2 + amp_weight (a - a_t)^2 ];
 * writes d_shape, d_amp (same layouts) and the loss to loss_out (1 FP32),
 * `partial` is scratch of 1024 FP32 values. */
int geo_trace_loss(int64_t batch, int64_t length, const float* shape, const float* shape_t,
                   const float* amp, int64_t amp_ld, const float* amp_t, float amp_weight,
                   float* d_shape, float* d_amp, float* partial, float* loss_out,
                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GEOSURROGATE_H */
