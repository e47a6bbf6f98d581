// Kernels of the trace surrogate (include/geosurrogate.h): a BF16 MFMA GEMM
// with FP32 accumulation for every matrix product of the network's forward
// and backward passes, and the elementwise / reduction / layout kernels
// around it.  gfx950 only.
#include "geofdtd.h"
#include "geosurrogate.h"
#include "geofdtd_internal.h"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <mutex>
#include <string>

// Split-K partial sums of geo_gemm_bf16_noalloc.  The split rule of
// gemm_bf16 keeps splits x output tiles <= 512, so the partials (splits x M x
// N FP32 values) never exceed 512 full 64 x 64 tiles: 8 MiB.  Every device
// the library runs on holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns on that
// device (the name itself is a host-side stand-in, not a device address).
constexpr long long kGemmScratchFloats = 512LL * 64 * 64;
static __device__ __align__(256) float scratch[kGemmScratchFloats];

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define SG_HIP(expr)                                                                        \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      return geo_internal_fail(GEO_ERR_HIP,                                                 \
                               (std::string(#expr) + " failed: " + hipGetErrorString(_e)).c_str()); \
  } while (0)

// Round to nearest even; NaN stays a quiet NaN.
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float sigmoid(float x) { return 1.0f / (1.0f + __expf(-x)); }

// ---------------------------------------------------------------------------
// GEMM.  256 threads = 4 waves in a 2 x 2 grid over a TM x TN block tile
// (64 x 64 or 128 x 128), K step 32.  A wave owns a (TM/2) x (TN/2) quadrant
// as (TM/32) x (TN/32) tiles of v_mfma_f32_16x16x32_bf16.  Operand layout of
// that instruction: lane l holds A[l % 16][8 (l / 16) + e] and
// B[8 (l / 16) + e][l % 16], e = 0..7, and returns D[4 (l / 16) + r][l % 16],
// r = 0..3.  LDS holds A as [m][k] and B transposed as [n][k], so every
// lane's operand is one 16-byte LDS read; rows are padded by 8 values against
// bank conflicts.  Every output tile accumulates the same 32-wide K slices
// in K order with the same instruction whatever the block tile, so the two
// tile sizes give the same bits.
//
// Operand tiles are fetched into registers one K step ahead and written to
// the other LDS buffer after the current step's MFMAs (double buffering, one
// barrier per step).  A tile is fetched with 16-byte loads when its K index
// is contiguous (8 consecutive k of one row per load) or its row index is
// contiguous (8 consecutive rows at one k, scattered into the [row][k] LDS
// layout); other strides, and the edges of the matrix, load element by
// element with zero fill.
// ---------------------------------------------------------------------------
constexpr int kTK = 32, kLdsRow = kTK + 8;

// LDS column of value k of tile row `row`: the four 8-value k groups of a row
// are permuted by bits 3-4 of the row (an XOR swizzle).  The 16-byte group
// reads and writes stay aligned, and the transposing stores of row-contiguous
// operands (8 rows apart per lane, one k each) spread over four banks instead
// of one (rows 8 apart are 160 words apart: the same bank).
__device__ __forceinline__ int lds_col(int row, int k) {
  return (((k >> 3) ^ ((row >> 3) & 3)) << 3) | (k & 7);
}
constexpr int kTM = 64, kTN = 64;  // tile of the split-K path (scratch sizing)
enum : int { kOpGeneric = 0, kOpKContig = 1, kOpRowContig = 2 };

// ROWS x kTK values of one operand, 8 per register vector, ROWS / 64 vectors
// per thread.  X(row, k) = X[row * rs + k * ks].
template <int ROWS, int MODE>
struct OperandTile {
  uint4 v[ROWS / 64];

  __device__ __forceinline__ void fetch(const uint16_t* __restrict__ X, int64_t rs, int64_t ks,
                                        int64_t row0, int64_t rows, int64_t k0, int64_t k_end) {
#pragma unroll
    for (int j = 0; j < ROWS / 64; ++j) {
      const int q = (int)threadIdx.x + 256 * j;
      // kOpRowContig: 8 consecutive rows at one k; otherwise 8 consecutive k of one row
      const int kk = MODE == kOpRowContig ? q / (ROWS / 8) : (q & 3) * 8;
      const int rr = MODE == kOpRowContig ? (q % (ROWS / 8)) * 8 : q >> 2;
      const int64_t row = row0 + rr, k = k0 + kk;
      if (MODE == kOpKContig && row < rows && k + 8 <= k_end) {
        v[j] = *reinterpret_cast<const uint4*>(X + row * rs + k);
      } else if (MODE == kOpRowContig && k < k_end && row + 8 <= rows) {
        v[j] = *reinterpret_cast<const uint4*>(X + row + k * ks);
      } else {
        uint16_t e8[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int64_t r = MODE == kOpRowContig ? row + e : row;
          const int64_t kx = MODE == kOpRowContig ? k : k + e;
          e8[e] = (r < rows && kx < k_end) ? X[r * rs + kx * ks] : (uint16_t)0;
        }
        v[j].x = (uint32_t)e8[0] | ((uint32_t)e8[1] << 16);
        v[j].y = (uint32_t)e8[2] | ((uint32_t)e8[3] << 16);
        v[j].z = (uint32_t)e8[4] | ((uint32_t)e8[5] << 16);
        v[j].w = (uint32_t)e8[6] | ((uint32_t)e8[7] << 16);
      }
    }
  }

  __device__ __forceinline__ void store(uint16_t* lds) const {
#pragma unroll
    for (int j = 0; j < ROWS / 64; ++j) {
      const int q = (int)threadIdx.x + 256 * j;
      if (MODE == kOpRowContig) {
        // Lanes ROWS/8 apart hold the same 8 rows at k and k + 1 (k even).
        // They swap halves: the even lane keeps rows rr..rr+3, the odd lane
        // rows rr+4..rr+7, and each writes its 4 rows as (k, k + 1) pairs,
        // 4 32-bit stores in place of 8 16-bit ones.
        const int kk = q / (ROWS / 8), rr = (q % (ROWS / 8)) * 8;
        const bool odd = (kk & 1) != 0;
        const uint32_t keep0 = odd ? v[j].z : v[j].x, keep1 = odd ? v[j].w : v[j].y;
        const uint32_t send0 = odd ? v[j].x : v[j].z, send1 = odd ? v[j].y : v[j].w;
        uint32_t got0, got1;
        if constexpr (ROWS / 8 == 8) {
          // partner lane ^ 8: a rotation by 8 within each row of 16 lanes
          // (DPP row_ror:8), with no trip through the LDS crossbar
          got0 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send0, 0x128, 0xf, 0xf, false);
          got1 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)send1, 0x128, 0xf, 0xf, false);
        } else {
          got0 = (uint32_t)__shfl_xor((int)send0, ROWS / 8);
          got1 = (uint32_t)__shfl_xor((int)send1, ROWS / 8);
        }
        const uint32_t lo[2] = {odd ? got0 : keep0, odd ? got1 : keep1};  // values at k
        const uint32_t hi[2] = {odd ? keep0 : got0, odd ? keep1 : got1};  // values at k + 1
        const int r0 = rr + (odd ? 4 : 0), k_even = kk & ~1;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t a = lo[e >> 1], b = hi[e >> 1];
          const uint32_t pair = (e & 1) ? (a >> 16) | (b & 0xffff0000u) : (a & 0xffffu) | (b << 16);
          *reinterpret_cast<uint32_t*>(&lds[(r0 + e) * kLdsRow + lds_col(r0 + e, k_even)]) = pair;
        }
      } else {
        const int row = q >> 2;
        *reinterpret_cast<uint4*>(&lds[row * kLdsRow + lds_col(row, (q & 3) * 8)]) = v[j];
      }
    }
  }
};

// Split K: block z covers k in [z * k_chunk, (z + 1) * k_chunk) and, with
// more than one split, writes its raw partial product to partial[z][m][n];
// gemm_reduce_kernel then adds the partials in split order (deterministic).
template <int TM, int TN, int AMODE, int BMODE>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(TM == 64 ? 5 : 1))) gemm_bf16_kernel(
    int64_t M, int64_t N, int64_t K, const uint16_t* __restrict__ A, int64_t a_rs, int64_t a_cs,
    const uint16_t* __restrict__ B, int64_t b_rs, int64_t b_cs, const float* __restrict__ bias,
    float* __restrict__ c32, uint16_t* __restrict__ c16, int64_t ldc, int32_t accumulate,
    int64_t k_chunk, float* __restrict__ partial) {
  constexpr int FI = TM / 32, FJ = TN / 32;  // 16 x 16 tiles per wave along m, n
  // the operand buffers, and after the K loop the output staging (64 rows of
  // TN + 4 FP32 values at a time)
  constexpr int kStageRow = TN + 4;
  constexpr int kOperandBytes = 2 * (TM + TN) * kLdsRow * 2, kStageBytes = 64 * kStageRow * 4;
  __shared__ __attribute__((aligned(16)))
      uint16_t smem[(kOperandBytes > kStageBytes ? kOperandBytes : kStageBytes) / 2];
  auto As = reinterpret_cast<uint16_t(*)[TM * kLdsRow]>(smem);
  auto Bt = reinterpret_cast<uint16_t(*)[TN * kLdsRow]>(smem + 2 * TM * kLdsRow);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t m0 = (int64_t)blockIdx.y * TM, n0 = (int64_t)blockIdx.x * TN;
  const int wm = (wave >> 1) * (TM / 2), wn = (wave & 1) * (TN / 2);
  f32x4 acc[FI][FJ];
#pragma unroll
  for (int i = 0; i < FI; ++i)
#pragma unroll
    for (int j = 0; j < FJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
  const int64_t k_begin = (int64_t)blockIdx.z * k_chunk;
  const int64_t k_end = k_begin + k_chunk < K ? k_begin + k_chunk : K;
  OperandTile<TM, AMODE> ta;
  OperandTile<TN, BMODE> tb;
  ta.fetch(A, a_rs, a_cs, m0, M, k_begin, k_end);
  tb.fetch(B, b_cs, b_rs, n0, N, k_begin, k_end);
  ta.store(As[0]);
  tb.store(Bt[0]);
  __syncthreads();
  int buf = 0;
  for (int64_t k0 = k_begin; k0 < k_end; k0 += kTK) {
    const bool more = k0 + kTK < k_end;
    if (more) {  // next step's operands in flight during this step's MFMAs
      ta.fetch(A, a_rs, a_cs, m0, M, k0 + kTK, k_end);
      tb.fetch(B, b_cs, b_rs, n0, N, k0 + kTK, k_end);
    }
    bf16x8 a[FI], b[FJ];
#pragma unroll
    for (int i = 0; i < FI; ++i) {
      const int row = wm + i * 16 + r16;
      a[i] = __builtin_bit_cast(
          bf16x8, *reinterpret_cast<const uint4*>(&As[buf][row * kLdsRow + lds_col(row, kq)]));
    }
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
      const int row = wn + j * 16 + r16;
      b[j] = __builtin_bit_cast(
          bf16x8, *reinterpret_cast<const uint4*>(&Bt[buf][row * kLdsRow + lds_col(row, kq)]));
    }
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
      for (int j = 0; j < FJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (more) {  // the other buffer: every wave finished reading it before the last barrier
      ta.store(As[buf ^ 1]);
      tb.store(Bt[buf ^ 1]);
    }
    __syncthreads();
    buf ^= 1;
  }
  // Output rows as 16-byte vectors where every output allows them: the
  // accumulators (4 rows x 16 columns per instruction, 64-byte row pieces)
  // go through LDS, 64 tile rows at a time, and each thread then writes 4
  // consecutive columns of one row.  Same values, same operations per value.
  const auto al = [](const void* x, uintptr_t a) { return (reinterpret_cast<uintptr_t>(x) & (a - 1)) == 0; };
  const bool vec_out = partial != nullptr
                           ? (N % 4 == 0 && al(partial, 16))
                           : (ldc % 4 == 0 && al(c32, 16) && al(c16, 8) && al(bias, 16));
  if (vec_out) {
    float* stage = reinterpret_cast<float*>(smem);
    constexpr int kVecPerRow = TN / 4, kRowsPerIter = 256 / kVecPerRow;
#pragma unroll
    for (int pass = 0; pass < TM / 64; ++pass) {
      if (wm / 64 == pass) {  // the waves whose quadrant holds these 64 rows
#pragma unroll
        for (int i = 0; i < FI; ++i)
#pragma unroll
          for (int j = 0; j < FJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              stage[(wm - pass * 64 + i * 16 + (lane >> 4) * 4 + r) * kStageRow + wn + j * 16 + r16] =
                  acc[i][j][r];
      }
      __syncthreads();
#pragma unroll
      for (int it = 0; it < 64 / kRowsPerIter; ++it) {
        const int lr = it * kRowsPerIter + t / kVecPerRow, c = (t % kVecPerRow) * 4;
        const int64_t m = m0 + pass * 64 + lr, n = n0 + c;
        if (m >= M || n >= N) continue;
        const float4 s4 = *reinterpret_cast<const float4*>(&stage[lr * kStageRow + c]);
        float v[4] = {s4.x, s4.y, s4.z, s4.w};
        if (n + 4 <= N) {
          if (partial != nullptr) {
            *reinterpret_cast<float4*>(partial + ((int64_t)blockIdx.z * M + m) * N + n) = s4;
            continue;
          }
          if (bias != nullptr) {
            const float4 b4 = *reinterpret_cast<const float4*>(bias + n);
            v[0] += b4.x; v[1] += b4.y; v[2] += b4.z; v[3] += b4.w;
          }
          if (c32 != nullptr) {
            float4* p = reinterpret_cast<float4*>(c32 + m * ldc + n);
            float4 o = make_float4(v[0], v[1], v[2], v[3]);
            if (accumulate) {
              const float4 q = *p;
              o = make_float4(q.x + v[0], q.y + v[1], q.z + v[2], q.w + v[3]);
            }
            *p = o;
          }
          if (c16 != nullptr)
            *reinterpret_cast<uint2*>(c16 + m * ldc + n) =
                make_uint2((uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                           (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16));
        } else {  // the last columns of the matrix
          for (int e = 0; e < 4 && n + e < N; ++e) {
            float x = v[e];
            if (partial != nullptr) {
              partial[((int64_t)blockIdx.z * M + m) * N + n + e] = x;
              continue;
            }
            if (bias != nullptr) x += bias[n + e];
            if (c32 != nullptr) {
              float* p = c32 + m * ldc + n + e;
              *p = accumulate ? *p + x : x;
            }
            if (c16 != nullptr) c16[m * ldc + n + e] = f2bf(x);
          }
        }
      }
      if (pass + 1 < TM / 64) __syncthreads();  // the staging rows are rewritten
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < FI; ++i)
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        const int64_t n = n0 + wn + j * 16 + r16;
        if (m >= M || n >= N) continue;
        float v = acc[i][j][r];
        if (partial != nullptr) {
          partial[((int64_t)blockIdx.z * M + m) * N + n] = v;
          continue;
        }
        if (bias != nullptr) v += bias[n];
        if (c32 != nullptr) {
          float* p = c32 + m * ldc + n;
          *p = accumulate ? *p + v : v;
        }
        if (c16 != nullptr) c16[m * ldc + n] = f2bf(v);
      }
}

__global__ void gemm_reduce_kernel(int64_t M, int64_t N, int32_t splits,
                                   const float* __restrict__ partial, const float* __restrict__ bias,
                                   float* __restrict__ c32, uint16_t* __restrict__ c16,
                                   int64_t ldc, int32_t accumulate) {
  const int64_t total = M * N;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = i / N, n = i - m * N;
    float v = 0.f;
    for (int32_t z = 0; z < splits; ++z) v += partial[(int64_t)z * total + i];
    if (bias != nullptr) v += bias[n];
    if (c32 != nullptr) {
      float* p = c32 + m * ldc + n;
      *p = accumulate ? *p + v : v;
    }
    if (c16 != nullptr) c16[m * ldc + n] = f2bf(v);
  }
}

int64_t grid_for(int64_t n, int threads = 256, int64_t cap = 65536) {
  const int64_t g = (n + threads - 1) / threads;
  return g < 1 ? 1 : (g > cap ? cap : g);
}

__global__ void f32_to_bf16_kernel(int64_t n, const float* __restrict__ x,
                                   uint16_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = f2bf(x[i]);
}

__global__ void silu_forward_kernel(int64_t n, const float* __restrict__ pre,
                                    uint16_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float x = pre[i];
    out[i] = f2bf(x * sigmoid(x));
  }
}

__global__ void silu_backward_kernel(int64_t n, const float* __restrict__ pre,
                                     const float* __restrict__ dout, float* __restrict__ dpre,
                                     uint16_t* __restrict__ dpre16) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float x = pre[i], s = sigmoid(x);
    const float d = dout[i] * s * (1.0f + x * (1.0f - s));
    dpre[i] = d;
    if (dpre16 != nullptr) dpre16[i] = f2bf(d);
  }
}

// Column sums in two deterministic stages.  Stage 1: block b sums rows
// [b * chunk, (b + 1) * chunk) of every column.  Its 256 threads cover
// `width` = min(N, 256) adjacent columns (coalesced row segments) times
// 256 / width row lanes; lane r takes rows r, r + lanes, ... and the lanes
// of a column are added in lane order.  Stage 2 adds the block partials of
// each column in block order.  The chunking depends on M only, so a sum is
// the same bits on every call.
constexpr int kColSumMaxBlocks = 1024;

__global__ void __launch_bounds__(256) column_partials_kernel(int64_t M, int64_t N,
                                                              const float* __restrict__ x,
                                                              int64_t ld, int64_t chunk,
                                                              float* __restrict__ partial) {
  __shared__ float red[256];
  const int t = threadIdx.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < M ? lo + chunk : M;
  const int width = N >= 256 ? 256 : (int)N;
  const int lanes = 256 / width;
  const int r = t / width, c = t - r * width;
  for (int64_t c0 = 0; c0 < N; c0 += width) {
    const int64_t col = c0 + c;
    float sum = 0.f;
    if (r < lanes && col < N)
      for (int64_t m = lo + r; m < hi; m += lanes) sum += x[m * ld + col];
    red[t] = sum;
    __syncthreads();
    if (t < width && col < N) {
      float acc = red[t];
      for (int q = 1; q < lanes; ++q) acc += red[q * width + t];
      partial[(int64_t)blockIdx.x * N + col] = acc;
    }
    __syncthreads();
  }
}

__global__ void column_reduce_kernel(int64_t N, int32_t blocks, const float* __restrict__ partial,
                                     float* __restrict__ out) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= N) return;
  float acc = partial[col];
  for (int32_t b = 1; b < blocks; ++b) acc += partial[(int64_t)b * N + col];
  out[col] = acc;
}

__global__ void im2col_up_kernel(int64_t batch, int64_t length, int64_t channels, int32_t up,
                                 int32_t ksize, const uint16_t* __restrict__ x,
                                 uint16_t* __restrict__ cols) {
  const int64_t lo = length * up, kc = (int64_t)ksize * channels;
  const int64_t total = batch * lo * kc;
  const int32_t half = ksize / 2;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = i / kc, col = i - row * kc;
    const int64_t b = row / lo, l = row - b * lo;
    const int64_t j = col / channels, c = col - j * channels;
    const int64_t u = l + j - half;
    cols[i] = (u >= 0 && u < lo) ? x[(b * length + u / up) * channels + c] : (uint16_t)0;
  }
}

__global__ void col2im_up_kernel(int64_t batch, int64_t length, int64_t channels, int32_t up,
                                 int32_t ksize, const float* __restrict__ dcols,
                                 float* __restrict__ dx) {
  const int64_t lo = length * up, kc = (int64_t)ksize * channels;
  const int64_t total = batch * length * channels;
  const int32_t half = ksize / 2;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t bl = i / channels, c = i - bl * channels;
    const int64_t b = bl / length, lp = bl - b * length;
    float s = 0.f;
    for (int32_t q = 0; q < up; ++q) {
      const int64_t u = lp * up + q;  // upsampled position filled from x[b, lp, c]
      for (int32_t j = 0; j < ksize; ++j) {
        const int64_t l = u - j + half;  // output row whose tap j reads u
        if (l >= 0 && l < lo) s += dcols[(b * lo + l) * kc + j * channels + c];
      }
    }
    dx[i] = s;
  }
}

__global__ void adam_kernel(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, uint16_t* __restrict__ p16,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (p16 != nullptr) p16[i] = f2bf(pi);
  }
}

constexpr int kLossBlocks = 1024;

__global__ void __launch_bounds__(256) trace_loss_kernel(
    int64_t batch, int64_t length, const float* __restrict__ s, const float* __restrict__ st,
    const float* __restrict__ a, int64_t a_ld, const float* __restrict__ at, float w,
    float* __restrict__ ds, float* __restrict__ da, float* __restrict__ partial) {
  __shared__ float red[256];
  const int64_t total = batch * length;
  const float inv_shape = 1.0f / (float)total;
  const float inv_batch = 1.0f / (float)batch;
  float acc = 0.f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float d = s[i] - st[i];
    acc += d * d * inv_shape;
    ds[i] = 2.f * d * inv_shape;
  }
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < batch;
       b += (int64_t)gridDim.x * blockDim.x) {
    const float d = a[b * a_ld] - at[b];
    acc += w * d * d * inv_batch;
    da[b * a_ld] = 2.f * w * d * inv_batch;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) sum_partials_kernel(const float* __restrict__ partial,
                                                           int32_t n, float* __restrict__ out) {
  __shared__ float red[256];
  float acc = 0.f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

constexpr int64_t kMaxSplitTiles = 512;
static_assert(kGemmScratchFloats == kMaxSplitTiles * kTM * kTN, "scratch holds the largest split");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call per device (no allocation: the array lives in the code object).
int gemm_scratch_address(float** out) {
  int dev = 0;
  SG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return geo_internal_fail(GEO_ERR_UNAVAILABLE, "gemm_noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    SG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return geo_internal_fail(GEO_ERR_HIP, "gemm_noalloc: scratch symbol not found");
    g_scratch_ptr[dev] = static_cast<float*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return GEO_OK;
}

// Stream-ordered pool of the split-K partials of gemm_bf16 and the chunk sums
// of geo_column_sums, one per device.
// Its release threshold keeps freed memory in the pool: the default pool
// hands it back at every synchronization, so that each weight gradient
// re-mapped its partials (~35 us of host time inside the call for 8 MiB).
hipMemPool_t g_partial_pool[kMaxDevices] = {};

int partial_pool(hipStream_t s, hipMemPool_t* out) {
  int dev = 0;
  if (s != nullptr) SG_HIP(hipStreamGetDevice(s, &dev));
  else SG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return geo_internal_fail(GEO_ERR_UNAVAILABLE, "gemm: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_partial_pool[dev]) {
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t pool = nullptr;
    SG_HIP(hipMemPoolCreate(&pool, &props));
    uint64_t keep = UINT64_MAX;
    SG_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    g_partial_pool[dev] = pool;
  }
  *out = g_partial_pool[dev];
  return GEO_OK;
}

// Load mode of an operand X(row, k) = X[row * rs + k * ks]: 16-byte vectors
// along whichever index is contiguous, when the base and the other stride
// keep every vector 16-byte aligned (rows and K steps start at multiples of 8).
int operand_mode(const uint16_t* X, int64_t rs, int64_t ks) {
  const bool aligned = (reinterpret_cast<uintptr_t>(X) & 15u) == 0;
  if (aligned && ks == 1 && rs % 8 == 0) return kOpKContig;
  if (aligned && rs == 1 && ks % 8 == 0) return kOpRowContig;
  return kOpGeneric;
}

template <int TM, int TN, int AM, int BM>
void launch_gemm_kernel(dim3 grid, hipStream_t s, int64_t M, int64_t N, int64_t K,
                        const uint16_t* A, int64_t a_rs, int64_t a_cs, const uint16_t* B,
                        int64_t b_rs, int64_t b_cs, const float* bias, float* c32, uint16_t* c16,
                        int64_t ldc, int32_t acc, int64_t k_chunk, float* partial) {
  hipLaunchKernelGGL((gemm_bf16_kernel<TM, TN, AM, BM>), grid, dim3(256), 0, s, M, N, K, A, a_rs,
                     a_cs, B, b_rs, b_cs, bias, c32, c16, ldc, acc, k_chunk, partial);
}

template <int TM, int TN>
void launch_gemm(int am, int bm, dim3 grid, hipStream_t s, int64_t M, int64_t N, int64_t K,
                 const uint16_t* A, int64_t a_rs, int64_t a_cs, const uint16_t* B, int64_t b_rs,
                 int64_t b_cs, const float* bias, float* c32, uint16_t* c16, int64_t ldc,
                 int32_t acc, int64_t k_chunk, float* partial) {
#define GEMM_CASE(a_, b_)                                                                    \
  case a_ * 3 + b_:                                                                          \
    launch_gemm_kernel<TM, TN, a_, b_>(grid, s, M, N, K, A, a_rs, a_cs, B, b_rs, b_cs, bias, \
                                       c32, c16, ldc, acc, k_chunk, partial);                \
    break;
  switch (am * 3 + bm) {
    GEMM_CASE(0, 0) GEMM_CASE(0, 1) GEMM_CASE(0, 2)
    GEMM_CASE(1, 0) GEMM_CASE(1, 1) GEMM_CASE(1, 2)
    GEMM_CASE(2, 0) GEMM_CASE(2, 1) GEMM_CASE(2, 2)
    default: break;
  }
#undef GEMM_CASE
}

// Rows per launch of a product taller than 65535 row tiles of 64 (grid.y):
// a multiple of 128 and of 8, so every chunk keeps the operand and output
// alignment of the whole product.
constexpr int64_t kGemmRowChunk = 32767LL * 128;

// C = A B (+ bias).  `scratch_partial` NULL: split-K partials come from the
// library's stream-ordered pool (partial_pool); otherwise they go to that buffer of
// kGemmScratchFloats values (device address of `scratch`).  `allow_split`
// false keeps one K pass per tile (the row chunks of a tall product).
int gemm_bf16(int64_t M, int64_t N, int64_t K, const uint16_t* A, int64_t a_rs, int64_t a_cs,
              const uint16_t* B, int64_t b_rs, int64_t b_cs, const float* bias, float* c_f32,
              uint16_t* c_bf16, int64_t ldc, int32_t accumulate, hipStream_t s,
              float* scratch_partial, bool allow_split = true) {
  if (M < 1 || N < 1 || K < 1 || !A || !B || (!c_f32 && !c_bf16) || ldc < N)
    return geo_internal_fail(GEO_ERR_INVALID, "gemm: bad sizes or null operands");
  if (accumulate && !c_f32)
    return geo_internal_fail(GEO_ERR_INVALID, "gemm: accumulation needs the FP32 output");
  const int64_t gy = (M + kTM - 1) / kTM, gx = (N + kTN - 1) / kTN;
  if (gx >= (1LL << 24))  // grid.x times 256 threads stays below 2^32
    return geo_internal_fail(GEO_ERR_INVALID, "gemm: more than 2^24 column tiles");
  if (M > kGemmRowChunk) {
    // Rows are independent: a tall product runs as row chunks with the same
    // kernels, one K pass per tile in every chunk (the last one included, so
    // that no row's sum is split where the others' are not).
    for (int64_t m = 0; m < M; m += kGemmRowChunk) {
      const int64_t rows = M - m < kGemmRowChunk ? M - m : kGemmRowChunk;
      const int rc = gemm_bf16(rows, N, K, A + m * a_rs, a_rs, a_cs, B, b_rs, b_cs, bias,
                               c_f32 ? c_f32 + m * ldc : nullptr,
                               c_bf16 ? c_bf16 + m * ldc : nullptr, ldc, accumulate, s,
                               scratch_partial, false);
      if (rc != GEO_OK) return rc;
    }
    return GEO_OK;
  }
  // Few output tiles over a long K (weight gradients: K = batch x length):
  // split K so that the launch fills the chip, at most 256 splits of >= 1024
  // and at most kMaxSplitTiles (split, tile) blocks.
  int64_t splits = 1;
  if (allow_split && gx * gy < kMaxSplitTiles && K >= 2048) {
    splits = kMaxSplitTiles / (gx * gy);
    if (splits > K / 1024) splits = K / 1024;
    if (splits > 256) splits = 256;
    if (splits < 1) splits = 1;
  }
  const int64_t k_chunk = ((K + splits - 1) / splits + kTK - 1) / kTK * kTK;
  splits = (K + k_chunk - 1) / k_chunk;
  const int am = operand_mode(A, a_rs, a_cs), bm = operand_mode(B, b_cs, b_rs);
  if (splits == 1) {
    // 128 x 128 block tiles where both sides fill them, K takes more than two
    // steps and the tiles still give every CU a block; 64 x 64 otherwise
    if (M >= 128 && N >= 128 && K > 2 * kTK && ((M + 127) / 128) * ((N + 127) / 128) >= 256) {
      const int64_t gy2 = (M + 127) / 128, gx2 = (N + 127) / 128;
      launch_gemm<128, 128>(am, bm, dim3((uint32_t)gx2, (uint32_t)gy2, 1), s, M, N, K, A, a_rs,
                            a_cs, B, b_rs, b_cs, bias, c_f32, c_bf16, ldc, accumulate, K,
                            (float*)nullptr);
    } else {
      launch_gemm<kTM, kTN>(am, bm, dim3((uint32_t)gx, (uint32_t)gy, 1), s, M, N, K, A, a_rs,
                            a_cs, B, b_rs, b_cs, bias, c_f32, c_bf16, ldc, accumulate, K,
                            (float*)nullptr);
    }
    SG_HIP(hipGetLastError());
    return GEO_OK;
  }
  float* partial = scratch_partial;
  if (scratch_partial) {
    if (splits * M * N > kGemmScratchFloats)
      return geo_internal_fail(GEO_ERR_INVALID,
                               "gemm_noalloc: split-K partials exceed the 8 MiB scratch");
  } else {
    hipMemPool_t pool = nullptr;
    const int rc = partial_pool(s, &pool);
    if (rc != GEO_OK) return rc;
    SG_HIP(hipMallocFromPoolAsync(reinterpret_cast<void**>(&partial), sizeof(float) * splits * M * N,
                                  pool, s));
  }
  launch_gemm<kTM, kTN>(am, bm, dim3((uint32_t)gx, (uint32_t)gy, (uint32_t)splits), s, M, N, K,
                        A, a_rs, a_cs, B, b_rs, b_cs, (const float*)nullptr, (float*)nullptr,
                        (uint16_t*)nullptr, ldc, 0, k_chunk, partial);
  hipLaunchKernelGGL(gemm_reduce_kernel, dim3(grid_for(M * N)), dim3(256), 0, s, M, N,
                     (int32_t)splits, (const float*)partial, bias, c_f32, c_bf16, ldc, accumulate);
  SG_HIP(hipGetLastError());
  if (!scratch_partial) SG_HIP(hipFreeAsync(partial, s));
  return GEO_OK;
}

}  // namespace

extern "C" {

int geo_gemm_bf16(int64_t M, int64_t N, int64_t K, const uint16_t* A, int64_t a_rs,
                  int64_t a_cs, const uint16_t* B, int64_t b_rs, int64_t b_cs, const float* bias,
                  float* c_f32, uint16_t* c_bf16, int64_t ldc, int32_t accumulate, void* stream) {
  return gemm_bf16(M, N, K, A, a_rs, a_cs, B, b_rs, b_cs, bias, c_f32, c_bf16, ldc, accumulate,
                   static_cast<hipStream_t>(stream), nullptr);
}

int geo_gemm_bf16_noalloc(int64_t M, int64_t N, int64_t K, const uint16_t* A, int64_t a_rs,
                          int64_t a_cs, const uint16_t* B, int64_t b_rs, int64_t b_cs,
                          const float* bias, float* c_f32, uint16_t* c_bf16, int64_t ldc,
                          int32_t accumulate, void* stream) {
  float* partial = nullptr;
  const int rc = gemm_scratch_address(&partial);
  if (rc != GEO_OK) return rc;
  return gemm_bf16(M, N, K, A, a_rs, a_cs, B, b_rs, b_cs, bias, c_f32, c_bf16, ldc, accumulate,
                   static_cast<hipStream_t>(stream), partial);
}

int geo_f32_to_bf16(int64_t n, const float* x, uint16_t* out, void* stream) {
  if (n < 0 || (n && (!x || !out))) return geo_internal_fail(GEO_ERR_INVALID, "null argument");
  if (n == 0) return GEO_OK;
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_for(n)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, x, out);
  SG_HIP(hipGetLastError());
  return GEO_OK;
}

int geo_silu_forward(int64_t n, const float* pre, uint16_t* out_bf16, void* stream) {
  if (n < 0 || (n && (!pre || !out_bf16))) return geo_internal_fail(GEO_ERR_INVALID, "null argument");
  if (n == 0) return GEO_OK;
  hipLaunchKernelGGL(silu_forward_kernel, dim3(grid_for(n)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, pre, out_bf16);
  SG_HIP(hipGetLastError());
  return GEO_OK;
}

int geo_silu_backward(int64_t n, const float* pre, const float* dout, float* dpre,
                      uint16_t* dpre_bf16, void* stream) {
  if (n < 0 || (n && (!pre || !dout || !dpre)))
    return geo_internal_fail(GEO_ERR_INVALID, "null argument");
  if (n == 0) return GEO_OK;
  hipLaunchKernelGGL(silu_backward_kernel, dim3(grid_for(n)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, pre, dout, dpre, dpre_bf16);
  SG_HIP(hipGetLastError());
  return GEO_OK;
}

int geo_column_sums(int64_t M, int64_t N, const float* x, int64_t ld, float* out, void* stream) {
  if (M < 1 || N < 1 || ld < N || !x || !out)
    return geo_internal_fail(GEO_ERR_INVALID, "column sums: bad sizes");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // about 256 rows per block, at most kColSumMaxBlocks blocks
  int64_t blocks = (M + 255) / 256;
  if (blocks > kColSumMaxBlocks) blocks = kColSumMaxBlocks;
  const int64_t chunk = (M + blocks - 1) / blocks;
  blocks = (M + chunk - 1) / chunk;
  float* partial = nullptr;
  hipMemPool_t pool = nullptr;
  const int rc = partial_pool(s, &pool);
  if (rc != GEO_OK) return rc;
  SG_HIP(hipMallocFromPoolAsync(reinterpret_cast<void**>(&partial), sizeof(float) * blocks * N, pool, s));
  hipLaunchKernelGGL(column_partials_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, M, N, x, ld,
                     chunk, partial);
  hipLaunchKernelGGL(column_reduce_kernel, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, s, N,
                     (int32_t)blocks, (const float*)partial, out);
  SG_HIP(hipGetLastError());
  SG_HIP(hipFreeAsync(partial, s));
  return GEO_OK;
}

int geo_im2col_up(int64_t batch, int64_t length, int64_t channels, int32_t up, int32_t ksize,
                  const uint16_t* x, uint16_t* cols, void* stream) {
  if (batch < 1 || length < 1 || channels < 1 || up < 1 || ksize < 1 || !(ksize & 1) || !x ||
      !cols)
    return geo_internal_fail(GEO_ERR_INVALID, "im2col: bad sizes (odd ksize required)");
  const int64_t total = batch * length * up * ksize * channels;
  hipLaunchKernelGGL(im2col_up_kernel, dim3(grid_for(total)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), batch, length, channels, up, ksize, x,
                     cols);
  SG_HIP(hipGetLastError());
  return GEO_OK;
}

int geo_col2im_up(int64_t batch, int64_t length, int64_t channels, int32_t up, int32_t ksize,
                  const float* dcols, float* dx, void* stream) {
  if (batch < 1 || length < 1 || channels < 1 || up < 1 || ksize < 1 || !(ksize & 1) ||
      !dcols || !dx)
    return geo_internal_fail(GEO_ERR_INVALID, "col2im: bad sizes (odd ksize required)");
  const int64_t total = batch * length * channels;
  hipLaunchKernelGGL(col2im_up_kernel, dim3(grid_for(total)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), batch, length, channels, up, ksize, dcols,
                     dx);
  SG_HIP(hipGetLastError());
  return GEO_OK;
}

int geo_adam_step(int64_t n, float* param, const float* grad, float* m, float* v,
                  uint16_t* param_bf16, float lr, float beta1, float beta2, float eps,
                  int64_t t, void* stream) {
  if (n < 1 || !param || !grad || !m || !v || t < 1)
    return geo_internal_fail(GEO_ERR_INVALID, "adam: bad arguments");
  const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, param, grad, m, v, param_bf16, lr,
                     beta1, beta2, eps, c1, c2);
  SG_HIP(hipGetLastError());
  return GEO_OK;
}

int geo_trace_loss(int64_t batch, int64_t length, const float* shape, const float* shape_t,
                   const float* amp, int64_t amp_ld, const float* amp_t, float amp_weight,
                   float* d_shape, float* d_amp, float* partial, float* loss_out,
                   void* stream) {
  if (batch < 1 || length < 1 || amp_ld < 1 || !shape || !shape_t || !amp || !amp_t ||
      !d_shape || !d_amp || !partial || !loss_out)
    return geo_internal_fail(GEO_ERR_INVALID, "trace loss: bad arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(trace_loss_kernel, dim3(kLossBlocks), dim3(256), 0, s, batch, length, shape,
                     shape_t, amp, amp_ld, amp_t, amp_weight, d_shape, d_amp, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, (const float*)partial,
                     kLossBlocks, loss_out);
  SG_HIP(hipGetLastError());
  return GEO_OK;
}

}  // extern "C"
