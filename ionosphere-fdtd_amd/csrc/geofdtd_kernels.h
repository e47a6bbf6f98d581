// Device code of the GeodesicFDTD step: two fused kernels per leapfrog step.
//
// Layout.  Every field is stored radial-innermost: one row per mesh entity
// (vertex, edge or face) holding its radial column, rows padded to a multiple
// of one 16-byte strip.  A thread owns one strip (4 FP32 / 2 FP64 / 8 FP16
// values) of one column, so the own-column traffic is 16-byte vector loads
// and stores, and each unstructured neighbour gather (the 2 Er columns of an
// edge, 3 Et columns of a face, <= 6 Ht columns of a dual cell, 2 Hr columns
// of an edge) reads a contiguous column slab that the strips of one wave
// cover with coalesced 16-byte requests.
//
// Kernels.  h_step_kernel advances Ht and Hr (reference solver.py:877-921),
// the surface-impedance lower boundary (surface_impedance.py:225-270) and the
// plasma ADE current (plasma.py:413-442, 489-500) which reads the same Et
// columns as the Hr curl; one block of it samples receivers
// (solver.py:1122-1124).  e_step_kernel advances Er and Et
// (solver.py:923-989) including the plasma current scatter
// (plasma.py:503-527) and the source terms (solver.py:936-959, 973-987).
// Blocks of the different segments are interleaved by a host-built block map
// and remapped so that each XCD sweeps one contiguous range of the mesh.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace geo {

constexpr int kBlock = 256;
constexpr int kMaxSource = 16;
constexpr int kMaxSiTerms = 32;
constexpr int kSiFastTerms = 16;  // terms held in registers (the reference default)
constexpr int kMaxSpecies = 4;
constexpr int kNumXcd = 8;

enum Segment : uint32_t {
  kSegHEdge = 0, kSegHFace = 1, kSegSample = 3,
  kSegEVertex = 4, kSegEEdge = 5, kSegHSample = 6,
};

// Column segment geometry.  Whole-row mapping: an integer number of rows
// per 256-thread block, column index within the block by a 16-bit
// reciprocal multiply (exact for tid < 256 and strips <= 256).  Flat
// mapping (one strip per thread only): thread t of block b takes strip
// b * 256 + t of the segment's row-major strip sequence, so rows may
// straddle blocks and every lane has work.
struct SegGeom {
  int32_t n_rows;
  int32_t strips;         // thread columns per row: ceil(row_strips / S)
  int32_t rows_per_block;
  uint32_t magic;         // ceil(65536 / strips)
  int32_t row_strips;     // 16-byte strips per row
  int32_t flat;           // 1: flat mapping (strips == row_strips)
};

struct Dynamic {
  int64_t counter;        // steps completed since geo_step started
  int64_t trace_rows;     // rows in the trace buffer (0: no trace)
  void* trace;            // (trace_rows, n_receivers) compute type
  const void* currents;   // (n_currents,) compute type, or nullptr
  int64_t n_currents;
  // Narrow storage: stored = value * scale, per field in the order er, et,
  // hr, ht.  Powers of two, so scaling is exact; the rescale kernels change
  // them between steps, never while a step kernel runs.
  float scale[4];
  float inv_scale[4];
  float factor[4];        // pending exact rescale of the stored values
  int32_t overflow;       // sticky: a non-finite stored value was seen
  int32_t sr_mode;        // narrow storage: 0 nearest, 1 independent, 2 rotation thresholds
  // Step index of the stochastic rounding (narrow storage).  The H kernel
  // rounds with sr_h and its sampling block (the only H writer of this block)
  // copies it to sr_e; the E kernel rounds with sr_e and one of its threads
  // writes sr_h = sr_e + 1 for the next step.  No launch reads a value that
  // the same launch writes.
  uint32_t sr_h;
  uint32_t sr_e;
  // H traces (record_h_observations): (h_rows, n) compute-type rows of the
  // weighted Hr / Ht samples, or null.  The E launch of the step that
  // completes step c (counter == c) writes row c / h_every when h_every
  // divides c, and row h_total / h_every + 1 after the last step h_total
  // otherwise; row 0 (the state before the call) is written separately.
  void* h_trace_r;
  void* h_trace_t;
  int64_t h_rows;
  int64_t h_every;
  int64_t h_total;
};

enum : int32_t { kFieldEr = 0, kFieldEt = 1, kFieldHr = 2, kFieldHt = 3 };

template <typename T> struct Strip {
  static constexpr int N = 16 / sizeof(T);
  T v[N];
} __attribute__((aligned(16)));

template <typename T> __device__ __forceinline__ constexpr bool is_narrow() {
  return sizeof(T) == 2;
}

template <typename C> __device__ __forceinline__ C to_c(float x) { return (C)x; }
template <typename C> __device__ __forceinline__ C to_c(double x) { return (C)x; }
template <typename C> __device__ __forceinline__ C to_c(__half x) { return (C)__half2float(x); }
template <typename C> __device__ __forceinline__ C to_c(__hip_bfloat16 x) {
  return (C)__bfloat162float(x);
}
template <typename T> __device__ __forceinline__ T from_c(float x);
template <> __device__ __forceinline__ float from_c<float>(float x) { return x; }
template <> __device__ __forceinline__ __half from_c<__half>(float x) { return __float2half(x); }
template <> __device__ __forceinline__ __hip_bfloat16 from_c<__hip_bfloat16>(float x) {
  return __float2bfloat16(x);
}
template <typename T> __device__ __forceinline__ T from_c(double x);
template <> __device__ __forceinline__ double from_c<double>(double x) { return x; }

// Stochastic rounding of narrow stores.  A stored FP16/BF16 value that
// rounds to nearest loses every increment smaller than half an ulp, so a
// slowly varying field stops following its update (quasi-static regions near
// a low-frequency source).  v_cvt_sr_{f16,bf16}_f32 round up with probability
// equal to the dropped fraction of an ulp, read from the high bits of a
// 32-bit threshold word u.  Two ways of drawing u, both functions of (step,
// field, element) only, so runs are bitwise repeatable:
//  * independent (mode 1): u = hash(step, field, element).  Each element's
//    rounding error is a sum of independent draws and random-walks: after N
//    steps with a steady sub-ulp increment it is ~sqrt(N p (1 - p)) ulp.
//  * rotation (mode 2): u = hash(field, element) + step * 0x9e3779b9, a random
//    phase per element advanced by the golden ratio each step (a Kronecker
//    sequence).  The thresholds of one element are equidistributed with
//    discrepancy O(log N / N), so a steady increment is followed to O(log N)
//    ulp; the per-element phase keeps neighbouring elements uncorrelated.
__device__ __forceinline__ uint32_t mix32(uint32_t h) {  // murmur3 finaliser
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

typedef short sr_v2s __attribute__((ext_vector_type(2)));

template <typename T> __device__ __forceinline__ T round_stochastic(float x, uint32_t u);
template <> __device__ __forceinline__ __half round_stochastic<__half>(float x, uint32_t u) {
  const sr_v2s v = __builtin_amdgcn_cvt_sr_f16_f32(sr_v2s{0, 0}, x, u, false);
  return __builtin_bit_cast(__half, v[0]);
}
template <> __device__ __forceinline__ __hip_bfloat16 round_stochastic<__hip_bfloat16>(float x,
                                                                                      uint32_t u) {
  const sr_v2s v = __builtin_amdgcn_cvt_sr_bf16_f32(sr_v2s{0, 0}, x, u, false);
  return __builtin_bit_cast(__hip_bfloat16, v[0]);
}

// How a launch rounds one field's stores.  mode 0: to nearest; otherwise the
// threshold of element `elem` is mix32(elem * K ^ key) + offset.
struct StoreRounding {
  uint32_t key;
  uint32_t offset;
  int32_t mode;
};

__device__ __forceinline__ StoreRounding store_rounding(int32_t mode, uint32_t step,
                                                        int32_t field) {
  const uint32_t salt = 0x632be5abu + (uint32_t)field * 0x5bd1e995u;
  if (mode == 2) return StoreRounding{mix32(salt), step * 0x9e3779b9u, 2};
  return StoreRounding{mix32(step * 4u + salt), 0u, mode};
}

__device__ __forceinline__ uint32_t sr_threshold(const StoreRounding& rd, uint32_t elem) {
  return mix32(elem * 0x9e3779b1u ^ rd.key) + rd.offset;
}

// Load one strip of N values and convert to compute type (undoing the
// power-of-two storage scale for narrow types).
template <typename T, typename C>
__device__ __forceinline__ void load_strip(const T* __restrict__ p, C* out, C inv_scale) {
  constexpr int N = Strip<T>::N;
  Strip<T> s = *reinterpret_cast<const Strip<T>*>(p);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if constexpr (is_narrow<T>()) out[i] = to_c<C>(s.v[i]) * inv_scale;
    else out[i] = to_c<C>(s.v[i]);
  }
}

// Narrow values of one strip: scaled, then rounded to nearest or
// stochastically (element ids elem0 + i).
template <typename T, typename C>
__device__ __forceinline__ Strip<T> narrow_strip(const C* in, C scale, StoreRounding rd,
                                                 uint32_t elem0) {
  constexpr int N = Strip<T>::N;
  Strip<T> s;
  if (rd.mode != 0) {
#pragma unroll
    for (int i = 0; i < N; ++i)
      s.v[i] = round_stochastic<T>((float)(in[i] * scale), sr_threshold(rd, elem0 + (uint32_t)i));
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) s.v[i] = from_c<T>(in[i] * scale);
  }
  return s;
}

template <typename T, typename C>
__device__ __forceinline__ void store_strip(T* __restrict__ p, const C* in, C scale,
                                            StoreRounding rd, uint32_t elem0) {
  constexpr int N = Strip<T>::N;
  Strip<T> s;
  if constexpr (is_narrow<T>()) {
    s = narrow_strip<T>(in, scale, rd, elem0);
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) s.v[i] = from_c<T>(in[i]);
  }
  *reinterpret_cast<Strip<T>*>(p) = s;
}

template <typename T, typename C>
__device__ __forceinline__ C load_one(const T* __restrict__ p, C inv_scale) {
  if constexpr (is_narrow<T>()) return to_c<C>(*p) * inv_scale;
  else return to_c<C>(*p);
}

// Storage scales of the four fields as seen by one kernel launch: read once
// from the device block for narrow types, identically 1 otherwise.
template <typename C>
struct Scales {
  C er, et, hr, ht, er_inv, et_inv, hr_inv, ht_inv;
  // rounding of the fields this launch stores (narrow storage)
  StoreRounding rd_er, rd_et, rd_hr, rd_ht;
};

// Storage scales, and the store rounding of a step kernel: the H kernel
// (E_SIDE false) stores Hr / Ht with step index sr_h, the E kernel Er / Et
// with sr_e.
template <typename T, typename C, bool E_SIDE = false>
__device__ __forceinline__ Scales<C> load_scales(const Dynamic* dyn) {
  Scales<C> s;
  s.rd_er = s.rd_et = s.rd_hr = s.rd_ht = StoreRounding{0u, 0u, 0};
  if constexpr (is_narrow<T>()) {
    const int32_t mode = dyn->sr_mode;
    const uint32_t step = E_SIDE ? dyn->sr_e : dyn->sr_h;
    if (E_SIDE) {
      s.rd_er = store_rounding(mode, step, kFieldEr);
      s.rd_et = store_rounding(mode, step, kFieldEt);
    } else {
      s.rd_hr = store_rounding(mode, step, kFieldHr);
      s.rd_ht = store_rounding(mode, step, kFieldHt);
    }
    s.er = dyn->scale[kFieldEr];
    s.et = dyn->scale[kFieldEt];
    s.hr = dyn->scale[kFieldHr];
    s.ht = dyn->scale[kFieldHt];
    s.er_inv = dyn->inv_scale[kFieldEr];
    s.et_inv = dyn->inv_scale[kFieldEt];
    s.hr_inv = dyn->inv_scale[kFieldHr];
    s.ht_inv = dyn->inv_scale[kFieldHt];
  } else {
    s.er = s.et = s.hr = s.ht = (C)1;
    s.er_inv = s.et_inv = s.hr_inv = s.ht_inv = (C)1;
  }
  return s;
}

// Aligned loads / stores of M compute-type values (16-byte multiples): the
// radial tables, loss coefficients and plasma arrays share the strip
// alignment of the fields, so one vector access replaces M lane loads.
template <typename C, int M>
struct alignas(16) CVec {
  C v[M];
};

// Per-entity rows read by every strip of an entity's column, packed so that
// one row is one 16-byte load (EdgeRow) or three (VertexRow, FP32) instead of
// separate index and weight loads.
template <typename C>
struct alignas(16) EdgeRow {
  int32_t a, b;  // H: tail, head vertex; E: left, right face
  C w;           // H: dt/(mu0 theta); E: 1/phi
};
static_assert(sizeof(EdgeRow<float>) == 16 && sizeof(EdgeRow<double>) == 16, "one 16-byte row");

// The E kernel's edge row also carries the edge's slot in its left and right
// face (slot_l | slot_r << 8; 0 without plasma), read by the plasma scatter.
template <typename C>
struct alignas(16) DualEdgeRow {
  int32_t a, b;  // left, right face
  C w;           // 1/phi
  int32_t slots;
};
static_assert(sizeof(DualEdgeRow<float>) == 16, "one 16-byte row");

template <typename C>
struct alignas(16) VertexRow {
  int32_t idx[6];  // incident edges (faces); padding slots repeat slot 0 with weight 0
  C w[6];          // sign * phi / A (face area weights)
};
static_assert(sizeof(VertexRow<float>) == 48, "three 16-byte pieces");

template <int M, typename C>
__device__ __forceinline__ void load_vec(const C* __restrict__ p, C* out) {
  const CVec<C, M> x = *reinterpret_cast<const CVec<C, M>*>(p);
#pragma unroll
  for (int i = 0; i < M; ++i) out[i] = x.v[i];
}

template <int M, typename C>
__device__ __forceinline__ void store_vec(C* __restrict__ p, const C* in) {
  CVec<C, M> x;
#pragma unroll
  for (int i = 0; i < M; ++i) x.v[i] = in[i];
  *reinterpret_cast<CVec<C, M>*>(p) = x;
}

// Streaming variants (non-temporal hint) for rows a kernel reads and writes
// once with no reuse inside the launch: the H kernel's own Ht, Hr and plasma
// rows.  Keeping them out of the caches leaves L2 to the gathered neighbour
// columns (measured: H kernel 0.262 -> 0.240 ms at s7 x 64).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename T, typename C>
__device__ __forceinline__ void load_strip_stream(const T* __restrict__ p, C* out, C inv_scale) {
  constexpr int N = Strip<T>::N;
  const u32x4 raw = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
  const Strip<T> s = __builtin_bit_cast(Strip<T>, raw);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if constexpr (is_narrow<T>()) out[i] = to_c<C>(s.v[i]) * inv_scale;
    else out[i] = to_c<C>(s.v[i]);
  }
}

template <typename T, typename C>
__device__ __forceinline__ void store_strip_stream(T* __restrict__ p, const C* in, C scale,
                                                   StoreRounding rd, uint32_t elem0) {
  constexpr int N = Strip<T>::N;
  Strip<T> s;
  if constexpr (is_narrow<T>()) {
    s = narrow_strip<T>(in, scale, rd, elem0);
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) s.v[i] = from_c<T>(in[i]);
  }
  __builtin_nontemporal_store(__builtin_bit_cast(u32x4, s), reinterpret_cast<u32x4*>(p));
}

template <int M, typename C>
__device__ __forceinline__ void load_vec_stream(const C* __restrict__ p, C* out) {
  constexpr int W = sizeof(C) * M / 16;
  static_assert(sizeof(C) * M == 16 * W, "whole 16-byte vectors");
  CVec<C, M> x;
  u32x4* dst = reinterpret_cast<u32x4*>(&x);
#pragma unroll
  for (int w = 0; w < W; ++w)
    dst[w] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p) + w);
#pragma unroll
  for (int i = 0; i < M; ++i) out[i] = x.v[i];
}

template <int M, typename C>
__device__ __forceinline__ void store_vec_stream(C* __restrict__ p, const C* in) {
  constexpr int W = sizeof(C) * M / 16;
  CVec<C, M> x;
#pragma unroll
  for (int i = 0; i < M; ++i) x.v[i] = in[i];
  const u32x4* src = reinterpret_cast<const u32x4*>(&x);
#pragma unroll
  for (int w = 0; w < W; ++w)
    __builtin_nontemporal_store(src[w], reinterpret_cast<u32x4*>(p) + w);
}


// Logical block index: hardware blocks are dispatched round-robin over the 8
// XCDs, so block b runs on XCD b % 8.  Give XCD x the contiguous logical range
// [x * per_xcd, (x + 1) * per_xcd) of the block map so neighbour columns stay
// in its L2 and blocks running together on one XCD work on adjacent regions.
__device__ __forceinline__ int32_t logical_block(int32_t per_xcd) {
  const int32_t b = blockIdx.x;
  return (b % kNumXcd) * per_xcd + b / kNumXcd;
}

__device__ __forceinline__ void row_and_strip(const SegGeom& g, int32_t block, int32_t* row,
                                              int32_t* strip) {
  const uint32_t tid = threadIdx.x;
  if (g.flat) {
    // rows * row_strips < 2^31 (checked on the host); rows past the end
    // come out >= n_rows and the caller skips them
    const uint32_t s = (uint32_t)block * (uint32_t)kBlock + tid;
    const uint32_t r = s / (uint32_t)g.row_strips;
    *row = (int32_t)r;
    *strip = (int32_t)(s - r * (uint32_t)g.row_strips);
    return;
  }
  const uint32_t local_row = (tid * g.magic) >> 16;
  *strip = (int32_t)(tid - local_row * (uint32_t)g.strips);
  *row = (local_row < (uint32_t)g.rows_per_block)
             ? block * g.rows_per_block + (int32_t)local_row
             : 0x7fffffff;
}

// Which neighbour strips of the same row sit in the adjacent lanes of this
// wave: lane - 1 holds strip - 1 when this thread is not the row's first
// column, lane + 1 holds strip + 1 when that strip exists in this pass.
struct Lanes {
  bool left, right;
};

__device__ __forceinline__ Lanes lanes_of(const SegGeom& g, int32_t col, int32_t strip) {
  const int32_t lane = (int32_t)(threadIdx.x & 63);
  return {col > 0 && lane > 0, col + 1 < g.strips && strip + 1 < g.row_strips && lane < 63};
}

template <typename C>
struct PlasmaTables {
  int32_t n_species;
  int32_t band_lo, band_hi;     // band_lo is a multiple of the strip width
  int32_t pitch;                // J row pitch (compute elements)
  int64_t species_stride;       // elements between species blocks (comps * F * pitch)
  int64_t comp_stride;          // elements between component blocks (F * pitch)
  // B = 0: (species, 3, F, pitch) holding J.n, J.t0, J.t1 (the face's edge
  // tangents are coplanar, so J.t2 = w1 J.t0 + w2 J.t1 is formed by the
  // scatter); magnetised: (species, 3, F, pitch) Cartesian J.
  C* current;
  // B = 0: (F, 8) = P[0][0..2], P[1][0..2] (P[c][c'] = t_c.R_c'), w1, w2;
  // magnetised: (F, 9) R = pinv(T).T
  const C* recon;
  const C* side_w;              // (E, 4) B = 0: (w1, w2) left side, then right; (1, 0) for stored slots
  const C* center;              // (F, 4) unit normal, padded
  const C* left_tan;            // (E, 4) padded
  const C* right_tan;           // (E, 4) padded
  const VertexRow<C>* vertex_faces;  // (V) faces around each vertex and their area weights,
                                     // in the E kernel's vertex processing order
  const C* coef;                // (species, 6, rows, pitch)
  int64_t coef_row_stride;      // 0 (uniform profile) or pitch
  int64_t coef_q_stride;        // elements between coefficient kinds
  int64_t coef_species_stride;
  const C* bdir;                // (rows, pitch, 3) unit B, magnetized only
  int64_t b_row_stride;
};

// Radii r_k = r0 + k dr (the default altitude grid is a linspace): FP32
// kernels then form 1/r_k, 1/rbar_j and the radial-derivative weights from
// these scalars instead of reading the (pn,) / (pm,) tables, two to three
// 16-byte table reads fewer per strip.  r0 is a head/tail pair so that
// r_k = r0_hi + (k dr + r0_lo) rounds once.  FP64 plans keep the tables.
template <typename C>
struct RadialGrid {
  int32_t uniform;
  C r0_hi, r0_lo, dr;
  C inv_dr;                     // 1/dr
};

// 1/r at radial position r0 + x dr (x = k for node k, j + 1/2 for midpoint j)
template <typename C>
__device__ __forceinline__ C inv_radius(const RadialGrid<C>& g, C x) {
  return (C)1 / (g.r0_hi + (x * g.dr + g.r0_lo));
}

template <typename T, typename C>
struct HParams {
  const T* __restrict__ er;
  const T* __restrict__ et;
  T* __restrict__ ht;
  T* __restrict__ hr;
  int32_t pn, pm;               // node / midpoint row pitch
  int32_t R;
  int32_t si_on;
  const EdgeRow<C>* __restrict__ edge_rows;  // (E) tail, head, dt/(mu0 theta)
  const int32_t* __restrict__ face_edges; // (F,3)
  const int32_t* __restrict__ faces;      // (F,3)
  const C* __restrict__ face_curl;        // (F,3) dt/mu0 s theta / Omega
  const C* __restrict__ inv_r;            // (pn,)
  const C* __restrict__ rd_b;             // (pn,) dt/mu0 alpha_k
  C gamma_scale;                          // dt/mu0 (spherical) or 0: dt/mu0 gamma_k = gamma_scale / r_k inside
  const C* __restrict__ inv_rbar;         // (pm,)
  RadialGrid<C> rg;
  C rb_a;                                 // dt/(mu0 dr) on a uniform grid
  // surface impedance: memory edge-major, (E, si_pitch), each edge's terms one
  // contiguous row of 16-byte pieces, read and written by the threads of the
  // edge's first radial strips (one piece each) or by the first strip's alone
  C* __restrict__ si_mem;
  int32_t si_pitch;
  const C* __restrict__ si_scale;         // (rows,) sqrt(mu/sigma)
  int32_t si_scale_stride;                // 0 or 1
  int32_t si_terms;
  C si_cs;                                // 2 dt/(mu0 dr0)
  C si_cs_metric;                         // si_cs * rbar0/r0
  C si_hw_sum;                            // sum of history weights
  // si_loss = 1 - decay of each memory term (increment form, as the loss)
  C si_loss[kMaxSiTerms], si_drive[kMaxSiTerms], si_hw[kMaxSiTerms];
  PlasmaTables<C> pl;
  // receivers + step counter
  Dynamic* dyn;
  const int32_t* __restrict__ rx_vertex;  // (n,3)
  const int32_t* __restrict__ rx_layer;   // (n,)
  const C* __restrict__ rx_w;             // (n,3)
  int32_t n_rx;
  // grid
  const uint32_t* __restrict__ block_map;
  int32_t n_blocks, per_xcd;
  SegGeom g_edge, g_face;
  int32_t n_edges;
};

template <typename T, typename C>
struct EParams {
  T* __restrict__ er;
  T* __restrict__ et;
  const T* __restrict__ ht;
  const T* __restrict__ hr;
  int32_t pn, pm, R;
  int32_t ca_er_stride, ca_et_stride;     // 0 or row pitch
  const int32_t* __restrict__ vertex_order;  // (V,) processing order
  const VertexRow<C>* __restrict__ vertex_rows;  // (V) incident edges, sign*phi/A, in processing
                                                 // order (row r describes vertex_order[r])
  const DualEdgeRow<C>* __restrict__ edge_rows;  // (E) left, right face, 1/phi, slots
  const C* __restrict__ cl_er;            // 1 - Ca (loss), zero on padding
  const C* __restrict__ cb_er;
  const C* __restrict__ cl_et;
  const C* __restrict__ cb_et;
  const C* __restrict__ inv_r;            // (pn,)
  const C* __restrict__ inv_rbar;         // (pm,)
  const C* __restrict__ rh_a;             // (pm,) alpha'_j
  int32_t spherical;                      // gamma'_j = 1/rbar_j (spherical) or 0
  RadialGrid<C> rg;
  PlasmaTables<C> pl;
  // sources: vertical (Er) and tangential (Et)
  int32_t n_src, n_tsrc;
  int32_t src_row[kMaxSource], src_layer[kMaxSource];
  C src_factor[kMaxSource];
  int32_t tsrc_row[kMaxSource], tsrc_layer[kMaxSource];
  C tsrc_factor[kMaxSource];
  Dynamic* dyn;
  // H receivers: (n, width) faces with midpoint layers, edges with node
  // layers, and their weights (record_h_observations, solver.py:1138-1229)
  const int32_t* __restrict__ hrx_face;
  const int32_t* __restrict__ hrx_flayer;
  const C* __restrict__ hrx_fw;
  const int32_t* __restrict__ hrx_edge;
  const int32_t* __restrict__ hrx_elayer;
  const C* __restrict__ hrx_ew;
  int32_t n_hrx_f, hrx_fwidth, n_hrx_e, hrx_ewidth;
  const uint32_t* __restrict__ block_map;
  int32_t n_blocks, per_xcd;
  SegGeom g_vertex, g_edge;
};

// ---------------------------------------------------------------------------
// H half-step
// ---------------------------------------------------------------------------

__device__ __forceinline__ float uniform_value(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uniform_value(double x) {
  const int64_t b = __double_as_longlong(x);
  const int32_t lo = __builtin_amdgcn_readfirstlane((int32_t)(b & 0xffffffff));
  const int32_t hi = __builtin_amdgcn_readfirstlane((int32_t)(b >> 32));
  return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

// Value q of 16-byte piece `piece` of a per-term coefficient array held in
// kernel parameters (uniform): a select chain over the pieces.
template <int SV, int P, typename C>
__device__ __forceinline__ C si_pick(const C* arr, int32_t piece, int q) {
  // every piece's value is taken as a uniform value first (scalar
  // registers), then the lane's piece is selected: selecting among the
  // kernel-argument loads themselves lets the compiler turn the chain into
  // one per-lane load from the kernel arguments at the end of the strip
  C vals[P];
#pragma unroll
  for (int s = 0; s < P; ++s) vals[s] = uniform_value(arr[s * SV + q]);
  C v = vals[0];
#pragma unroll
  for (int s = 1; s < P; ++s) v = piece == s ? vals[s] : v;
  return v;
}

template <typename T, typename C, bool SPREAD>
__device__ __forceinline__ void h_edge_strip(const HParams<T, C>& p, const Scales<C>& sc, int32_t e,
                                             int32_t strip, Lanes ln) {
  constexpr int N = Strip<T>::N;
  const int32_t k0 = strip * N;
  // Surface-impedance memory of the edge (si_pitch values, si_p pieces of 16
  // bytes): with one strip per thread the lanes of strips 0..si_p-1 hold one
  // piece each -- one 16-byte load and store per lane instead of si_p of each
  // by the first strip's lane -- when those lanes share the first strip's
  // wave; otherwise the first strip's lane updates the whole row.
  constexpr int SV = 16 / (int)sizeof(C);
  constexpr int kPieces = kSiFastTerms / SV;
  const int32_t si_p = p.si_pitch / SV;
  const int32_t lane = (int32_t)(threadIdx.x & 63), lane0 = lane - strip;
  const bool spread = SPREAD && p.si_on && p.si_pitch <= kSiFastTerms &&
                      si_p <= p.g_edge.row_strips && lane0 >= 0 && lane0 + si_p <= 64;
  C mq[SV];
#pragma unroll
  for (int q = 0; q < SV; ++q) mq[q] = (C)0;
  C* si_row = p.si_mem + (size_t)e * p.si_pitch;
  if (spread && strip < si_p) load_vec_stream<SV>(si_row + SV * strip, mq);
  C si_scale = (C)0;
  if (spread && strip == 0) si_scale = p.si_scale[e * p.si_scale_stride];
  // The edge's own rows are requested before its vertex indices, so that
  // they are in flight while the Er gathers wait on the indices (H kernel
  // 0.205 -> 0.201 ms at s7 x 64 with the same order in the face strip).
  C ert[N], erh[N], ht[N], et[N];
  T* hrow = p.ht + (size_t)e * p.pn + k0;
  load_strip_stream(hrow, ht, sc.ht_inv);
  const T* erow = p.et + (size_t)e * p.pm;
  if (k0 < p.pm) {
    load_strip(erow + k0, et, sc.et_inv);
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) et[i] = (C)0;
  }
  const EdgeRow<C> row = p.edge_rows[e];
  const int32_t tail = row.a, head = row.b;
  const C a = row.w;
  load_strip(p.er + (size_t)tail * p.pn + k0, ert, sc.er_inv);
  load_strip(p.er + (size_t)head * p.pn + k0, erh, sc.er_inv);
  C rb[N], ir[N];
  if (sizeof(C) == 4 && p.rg.uniform) {
    // dt/mu0 alpha_k on r_k = r0 + k dr: dt/mu0 (1/dr - 1/(2 r_k)) inside,
    // dt/mu0 (2/dr + 1/r_0) and dt/mu0 (2/dr - 1/r_R) at the ends on a
    // spherical grid (gamma_scale = dt/mu0), 1/dr and 2/dr times dt/mu0 on a
    // flat one (gamma_scale = 0); zero beyond R like the tables
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int32_t k = k0 + i;
      const C inv = k <= p.R ? inv_radius(p.rg, (C)k) : (C)0;
      ir[i] = inv;
      rb[i] = k > p.R ? (C)0
              : k == 0 ? (C)2 * p.rb_a + p.gamma_scale * inv
              : k == p.R ? (C)2 * p.rb_a - p.gamma_scale * inv
                         : p.rb_a - (C)0.5 * p.gamma_scale * inv;
    }
  } else {
    load_vec<N>(p.rd_b + k0, rb);
    load_vec<N>(p.inv_r + k0, ir);
  }
  // every load is issued before the first wait (the lane shuffle below)
  C prev_loaded = (C)0;
  if (!ln.left && k0 > 0) prev_loaded = load_one(erow + k0 - 1, sc.et_inv);
  const C from_left = __shfl_up(et[N - 1], 1);
  C prev = ln.left ? from_left : prev_loaded;
  const C h_old0 = ht[0];
  // Rows are padded with zeros and the radial tables are zero beyond R, so
  // the strip is updated unconditionally: Et(R) reads the zero padding (odd
  // ghost, PEC cap, solver.py:448-449) and padding lanes stay unchanged.
  // dt/mu0 gamma_k = gamma_scale / r_k at interior nodes, 0 at k = 0 and R
  // (the odd ghosts fold into alpha there).
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int32_t k = k0 + i;
    const C g = (k > 0 && k < p.R) ? p.gamma_scale : (C)0;
    const C cur = et[i];
    ht[i] = ht[i] + (ir[i] * (a * (erh[i] - ert[i]) - g * cur) - rb[i] * (cur - prev));
    prev = cur;
  }
  // node 0 under a surface-impedance ground replaces the PEC value
  if (p.si_on) {
    const C forcing = a * ir[0] * (erh[0] - ert[0]);
    if (spread) {
      C part = (C)0;
#pragma unroll
      for (int q = 0; q < SV; ++q) part += mq[q] * si_pick<SV, kPieces>(p.si_hw, strip, q);
      C hist = part;  // pieces summed in order at the first strip
#pragma unroll
      for (int s = 1; s < kPieces; ++s) {
        const C other = __shfl(part, lane0 + s);
        if (s < si_p) hist += other;
      }
      hist *= si_scale;
      const C coupling = p.si_cs * (si_scale * p.si_hw_sum);
      const C h = h_old0 + (forcing - p.si_cs_metric * et[0] + p.si_cs * hist - coupling * h_old0) /
                               ((C)1 + (C)0.5 * coupling);
      const C sum = __shfl(h + h_old0, lane0);
      if (strip < si_p) {
#pragma unroll
        for (int q = 0; q < SV; ++q)
          mq[q] = mq[q] + (sum * si_pick<SV, kPieces>(p.si_drive, strip, q) -
                           si_pick<SV, kPieces>(p.si_loss, strip, q) * mq[q]);
        store_vec_stream<SV>(si_row + SV * strip, mq);
      }
      if (strip == 0) ht[0] = h;
    } else if (k0 == 0) {
      ht[0] = si_boundary(p, e, forcing, h_old0, et[0]);
    }
  }
  store_strip_stream(hrow, ht, sc.ht, sc.rd_ht, (uint32_t)e * (uint32_t)p.pn + (uint32_t)k0);
}

template <typename C>
__device__ __forceinline__ void cross_b(const C* v, const C* b, C* out) {
  out[0] = v[1] * b[2] - v[2] * b[1];
  out[1] = v[2] * b[0] - v[0] * b[2];
  out[2] = v[0] * b[1] - v[1] * b[0];
}

template <typename T, typename C, bool PLASMA, bool MAGNETIZED>
__device__ __forceinline__ void h_face_strip(const HParams<T, C>& p, const Scales<C>& sc, int32_t f,
                                             int32_t strip, Lanes ln) {
  constexpr int N = Strip<T>::N;
  const int32_t j0 = strip * N;
  // own row first, then the face's edges (and, in the plasma band, its
  // vertices), so that the gathers wait on indices only
  C hr[N];
  T* hrow = p.hr + (size_t)f * p.pm + j0;
  load_strip_stream(hrow, hr, sc.hr_inv);
  const int32_t e0 = p.face_edges[3 * f];
  const int32_t e1 = p.face_edges[3 * f + 1];
  const int32_t e2 = p.face_edges[3 * f + 2];
  const C w0 = p.face_curl[3 * f], w1 = p.face_curl[3 * f + 1], w2 = p.face_curl[3 * f + 2];
  int32_t fv0 = 0, fv1 = 0, fv2 = 0;
  if (PLASMA && j0 >= p.pl.band_lo && j0 < p.pl.band_hi) {
    fv0 = p.faces[3 * f];
    fv1 = p.faces[3 * f + 1];
    fv2 = p.faces[3 * f + 2];
  }
  C et0[N], et1[N], et2[N];
  load_strip(p.et + (size_t)e0 * p.pm + j0, et0, sc.et_inv);
  load_strip(p.et + (size_t)e1 * p.pm + j0, et1, sc.et_inv);
  load_strip(p.et + (size_t)e2 * p.pm + j0, et2, sc.et_inv);
  C irb[N];
  if (sizeof(C) == 4 && p.rg.uniform) {
#pragma unroll
    for (int i = 0; i < N; ++i)
      irb[i] = j0 + i < p.R ? inv_radius(p.rg, (C)(j0 + i) + (C)0.5) : (C)0;
  } else {
    load_vec<N>(p.inv_rbar + j0, irb);
  }
#pragma unroll
  for (int i = 0; i < N; ++i)  // inv_rbar is zero on padding lanes
    hr[i] = hr[i] - irb[i] * (w0 * et0[i] + w1 * et1[i] + w2 * et2[i]);
  store_strip_stream(hrow, hr, sc.hr, sc.rd_hr, (uint32_t)f * (uint32_t)p.pm + (uint32_t)j0);

  if constexpr (PLASMA) {
    const PlasmaTables<C>& pl = p.pl;
    // band_lo is a strip multiple, so a strip is either fully below the band
    // or starts inside it; lanes at or beyond band_hi are J padding (zero).
    if (j0 < pl.band_lo || j0 >= pl.band_hi) return;
    // E at (face, midpoint): tangential reconstruction + radial node average
    // (plasma.py:492-500).
    const int32_t v0 = fv0, v1 = fv1, v2 = fv2;
    C a0[N + 1], a1[N + 1], a2[N + 1];
    load_strip(p.er + (size_t)v0 * p.pn + j0, a0, sc.er_inv);
    load_strip(p.er + (size_t)v1 * p.pn + j0, a1, sc.er_inv);
    load_strip(p.er + (size_t)v2 * p.pn + j0, a2, sc.er_inv);
    // node j0+N is needed only while midpoint j0+N-1 < R, i.e. j0+N <= R
    const bool top = j0 + N <= p.R;
    const C r0 = __shfl_down(a0[0], 1), r1 = __shfl_down(a1[0], 1), r2 = __shfl_down(a2[0], 1);
    if (ln.right && j0 + N < pl.band_hi) {  // lane + 1 is in the band too
      a0[N] = r0;
      a1[N] = r1;
      a2[N] = r2;
    } else {
      a0[N] = top ? load_one(p.er + (size_t)v0 * p.pn + j0 + N, sc.er_inv) : (C)0;
      a1[N] = top ? load_one(p.er + (size_t)v1 * p.pn + j0 + N, sc.er_inv) : (C)0;
      a2[N] = top ? load_one(p.er + (size_t)v2 * p.pn + j0 + N, sc.er_inv) : (C)0;
    }
    C rad[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const C lower = (a0[i] + a1[i] + a2[i]) / (C)3;
      const C upper = (a0[i + 1] + a1[i + 1] + a2[i + 1]) / (C)3;
      rad[i] = (C)0.5 * (lower + upper);
    }
    const int32_t jb0 = j0 - pl.band_lo;
    if constexpr (!MAGNETIZED) {
      // drive projections: E.n = rad, E.t_c = sum_c' P[c][c'] Et_c' (c = 0, 1)
      C rc[6];
      load_vec<4>(pl.recon + 8 * f, rc);
      rc[4] = pl.recon[8 * f + 4];
      rc[5] = pl.recon[8 * f + 5];
      C drive[3][N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        drive[0][i] = rad[i];
#pragma unroll
        for (int c = 0; c < 2; ++c)
          drive[1 + c][i] = (rc[3 * c] * et0[i] + rc[3 * c + 1] * et1[i]) + rc[3 * c + 2] * et2[i];
      }
      for (int s = 0; s < pl.n_species; ++s) {
        C* jp = pl.current + s * pl.species_stride + (size_t)f * pl.pitch + jb0;
        const C* cf = pl.coef + s * pl.coef_species_stride + (size_t)f * pl.coef_row_stride + jb0;
        // slot 0 holds 1 - exp(-nu dt): J + (d E - (1 - decay) J)
        C loss[N], dre[N];
        load_vec<N>(cf, loss);
        load_vec<N>(cf + 4 * pl.coef_q_stride, dre);
        // the three components are loaded before any is stored (the stores
        // would otherwise order each load after the previous store)
        C jx[3][N];
#pragma unroll
        for (int q = 0; q < 3; ++q) load_vec_stream<N>(jp + q * pl.comp_stride, jx[q]);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          C* x = jx[q];
#pragma unroll
          for (int i = 0; i < N; ++i)
            x[i] = (j0 + i < pl.band_hi) ? x[i] + (dre[i] * drive[q][i] - loss[i] * x[i]) : x[i];
          store_vec_stream<N>(jp + q * pl.comp_stride, x);
        }
      }
    } else {
      C rc[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) rc[q] = pl.recon[9 * f + q];
      C nrm[4];
      load_vec<4>(pl.center + 4 * f, nrm);
      C E[3][N];
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          E[d][i] = (et0[i] * rc[d] + et1[i] * rc[3 + d]) + et2[i] * rc[6 + d] + rad[i] * nrm[d];
      for (int s = 0; s < pl.n_species; ++s) {
        C* jp = pl.current + s * pl.species_stride + (size_t)f * pl.pitch + jb0;
        const C* cf = pl.coef + s * pl.coef_species_stride + (size_t)f * pl.coef_row_stride + jb0;
        // slots 0 and 1 hold 1 - exp(-nu dt) and 1 - cos(omega_c dt)
        C J[3][N], loss[N], dre[N], vers[N], sn[N], dpar[N], dim[N];
#pragma unroll
        for (int d = 0; d < 3; ++d) load_vec<N>(jp + d * pl.comp_stride, J[d]);
        load_vec<N>(cf, loss);
        load_vec<N>(cf + pl.coef_q_stride, vers);
        load_vec<N>(cf + 2 * pl.coef_q_stride, sn);
        load_vec<N>(cf + 3 * pl.coef_q_stride, dpar);
        load_vec<N>(cf + 4 * pl.coef_q_stride, dre);
        load_vec<N>(cf + 5 * pl.coef_q_stride, dim);
#pragma unroll
        for (int i = 0; i < N; ++i) {
          if (j0 + i >= pl.band_hi) continue;
          const C* bp = pl.bdir + (size_t)f * pl.b_row_stride + 3 * (jb0 + i);
          const C b[3] = {bp[0], bp[1], bp[2]};
          const C Ei[3] = {E[0][i], E[1][i], E[2][i]};
          const C Ji[3] = {J[0][i], J[1][i], J[2][i]};
          const C es = (Ei[0] * b[0] + Ei[1] * b[1]) + Ei[2] * b[2];
          const C js = (Ji[0] * b[0] + Ji[1] * b[1]) + Ji[2] * b[2];
          C epar[3], eperp[3], jpar[3], jperp[3], ecr[3], jcr[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            epar[d] = es * b[d];
            eperp[d] = Ei[d] - epar[d];
            jpar[d] = js * b[d];
            jperp[d] = Ji[d] - jpar[d];
          }
          cross_b(eperp, b, ecr);
          cross_b(jperp, b, jcr);
          // decay (J_par + cos J_perp + sin J_perp x b) + drive terms, with
          // the rotation written as J + rot, rot = sin (J_perp x b) - (1 - cos) J_perp
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const C rot = sn[i] * jcr[d] - vers[i] * jperp[d];
            J[d][i] = Ji[d] + (rot - loss[i] * (Ji[d] + rot) + dpar[i] * epar[d] +
                               dre[i] * eperp[d] + dim[i] * ecr[d]);
          }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) store_vec<N>(jp + d * pl.comp_stride, J[d]);
      }
    }
  }
}

// Implicit lower-boundary Ht of edge e with the diffusive surface-impedance
// memory (surface_impedance.py:225-270), in increment form:
// ((1 - c/2) h_old + rest) / (1 + c/2) = h_old + (rest - c h_old) / (1 + c/2)
// and m' = m + (sum drive - (1 - decay) m), so the small coupling c and the
// slow decay rates of the low-frequency poles keep their relative precision.
// Returns the new Ht(e, 0); forcing = dt/(mu0 theta r0) (Er_head - Er_tail).
template <typename T, typename C>
__device__ __forceinline__ C si_boundary(const HParams<T, C>& p, int32_t e, C forcing, C h_old,
                                         C et0) {
  const C scale = p.si_scale[e * p.si_scale_stride];
  C* mem = p.si_mem + (size_t)e * p.si_pitch;
  const C coupling = p.si_cs * (scale * p.si_hw_sum);
  constexpr int V = 16 / sizeof(C);
  if (p.si_pitch <= kSiFastTerms) {
    // the row in registers: 16-byte loads, terms past si_terms are zero with
    // zero weights, drives and losses
    C m[kSiFastTerms];
#pragma unroll
    for (int i = 0; i < kSiFastTerms; i += V) {
      if (i < p.si_pitch) load_vec_stream<V>(mem + i, m + i);
      else
#pragma unroll
        for (int q = 0; q < V; ++q) m[i + q] = (C)0;
    }
    // summed per 16-byte piece, then over the pieces in order: the order of
    // additions of the spread update in h_edge_strip, so that both give the
    // same bits
    constexpr int kPieces = kSiFastTerms / V;
    const int32_t si_p = p.si_pitch / V;
    C hist = (C)0;
#pragma unroll
    for (int s = 0; s < kPieces; ++s) {
      C part = (C)0;
#pragma unroll
      for (int q = 0; q < V; ++q) part += m[s * V + q] * p.si_hw[s * V + q];
      if (s == 0) hist = part;
      else if (s < si_p) hist += part;
    }
    hist *= scale;
    const C h = h_old + (forcing - p.si_cs_metric * et0 + p.si_cs * hist - coupling * h_old) /
                            ((C)1 + (C)0.5 * coupling);
    const C sum = h + h_old;
#pragma unroll
    for (int i = 0; i < kSiFastTerms; ++i) m[i] = m[i] + (sum * p.si_drive[i] - p.si_loss[i] * m[i]);
#pragma unroll
    for (int i = 0; i < kSiFastTerms; i += V)
      if (i < p.si_pitch) store_vec_stream<V>(mem + i, m + i);
    return h;
  }
  C hist = (C)0;
  for (int i = 0; i < p.si_terms; ++i) hist += mem[i] * p.si_hw[i];
  hist *= scale;
  const C h = h_old + (forcing - p.si_cs_metric * et0 + p.si_cs * hist - coupling * h_old) /
                          ((C)1 + (C)0.5 * coupling);
  const C sum = h + h_old;
  for (int i = 0; i < p.si_terms; ++i) {
    const C mi = mem[i];
    mem[i] = mi + (sum * p.si_drive[i] - p.si_loss[i] * mi);
  }
  return h;
}

template <typename T, typename C>
__device__ __forceinline__ void sample_block(const HParams<T, C>& p, const Scales<C>& sc) {
  // Receiver row for the Er state entering this step, then advance the
  // device step counter (only this block reads or writes it in the H kernel).
  __shared__ int64_t row;
  if (threadIdx.x == 0) row = p.dyn->counter;
  __syncthreads();
  const int64_t r = row;
  C* trace = reinterpret_cast<C*>(p.dyn->trace);
  if (trace != nullptr && r < p.dyn->trace_rows) {
    for (int i = threadIdx.x; i < p.n_rx; i += blockDim.x) {
      const int32_t layer = p.rx_layer[i];
      C acc = (C)0;
#pragma unroll
      for (int q = 0; q < 3; ++q)
        acc += load_one(p.er + (size_t)p.rx_vertex[3 * i + q] * p.pn + layer, sc.er_inv) *
               p.rx_w[3 * i + q];
      trace[r * p.n_rx + i] = acc;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    p.dyn->counter = r + 1;
    if constexpr (is_narrow<T>()) p.dyn->sr_e = p.dyn->sr_h;
  }
}

// One block of the H kernel.  With S strips per thread (a column block of
// S x 16 bytes), thread column c of a row handles strips c, c + cols, ...,
// so every load instruction of a wave stays coalesced along the row.
template <typename T, typename C, bool PLASMA, bool MAGNETIZED, int S>
__device__ __forceinline__ void h_dispatch(const HParams<T, C>& p, const Scales<C>& sc,
                                           uint32_t seg, int32_t b) {
  if (seg == kSegHEdge) {
    int32_t row, col;
    row_and_strip(p.g_edge, b, &row, &col);
    if (row < p.g_edge.n_rows) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int32_t strip = col + s * p.g_edge.strips;
        if (S == 1 || strip < p.g_edge.row_strips)
          h_edge_strip<T, C, S == 1>(p, sc, row, strip, lanes_of(p.g_edge, col, strip));
      }
    }
  } else if (seg == kSegHFace) {
    int32_t row, col;
    row_and_strip(p.g_face, b, &row, &col);
    if (row < p.g_face.n_rows) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int32_t strip = col + s * p.g_face.strips;
        if (S == 1 || strip < p.g_face.row_strips)
          h_face_strip<T, C, PLASMA, MAGNETIZED>(p, sc, row, strip, lanes_of(p.g_face, col, strip));
      }
    }
  } else {
    sample_block(p, sc);
  }
}

// Narrow storage (8-value strips) leaves the compiler at 95 VGPRs and 5
// waves per SIMD for the H kernel with Drude plasma; asking for 6 waves
// fits it in 80 VGPRs without spills (s8 x 96 FP16: 1.021 -> 0.980 ms,
// s7 x 64: 0.177 -> 0.171 ms).  The same request spills the magnetised
// variant, and the E kernel (108 VGPRs), so neither gets it.
template <typename T, bool MAGNETIZED>
constexpr int h_min_waves() { return (sizeof(T) == 2 && !MAGNETIZED) ? 6 : 1; }

template <typename T, typename C, bool PLASMA, bool MAGNETIZED>
__global__ void __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(h_min_waves<T, MAGNETIZED>())))
h_step_kernel(const HParams<T, C> p) {
  const int32_t L = logical_block(p.per_xcd);
  if (L >= p.n_blocks) return;
  const Scales<C> sc = load_scales<T, C>(p.dyn);
  const uint32_t m = p.block_map[L];
  h_dispatch<T, C, PLASMA, MAGNETIZED, 1>(p, sc, m >> 24, (int32_t)(m & 0xffffffu));
}

// ---------------------------------------------------------------------------
// E half-step
// ---------------------------------------------------------------------------

// Species sum of one stored component (projected state) for the N lanes of
// an in-band strip at j0, plus the value at j0 - 1 when that layer is in band.
template <typename C, int N, bool WANT_PREV>
__device__ __forceinline__ void plasma_component(const PlasmaTables<C>& pl, int32_t f,
                                                 int32_t comp, int32_t j0, C* strip, C* prev) {
  const bool has_strip = j0 >= pl.band_lo && j0 < pl.band_hi;
  const bool has_prev = WANT_PREV && j0 - 1 >= pl.band_lo && j0 - 1 < pl.band_hi;
#pragma unroll
  for (int i = 0; i < N; ++i) strip[i] = (C)0;
  C pv = (C)0;
  const int32_t jb0 = j0 - pl.band_lo;
  for (int s = 0; s < pl.n_species; ++s) {
    const C* base = pl.current + s * pl.species_stride + comp * pl.comp_stride + (size_t)f * pl.pitch;
    if (has_strip) {
      C x[N];
      load_vec<N>(base + jb0, x);
#pragma unroll
      for (int i = 0; i < N; ++i) strip[i] += x[i];
    }
    if (has_prev) pv += base[jb0 - 1];
  }
  if constexpr (WANT_PREV) *prev = pv;
}

// Species-summed Cartesian J(face, j0 + i) . v for the N lanes of a strip
// starting at an in-band j0 (lanes past the band read the zero padding), plus
// the value at j0 - 1 when that layer is in the band.
template <typename C, int N, bool WANT_PREV>
__device__ __forceinline__ void plasma_project(const PlasmaTables<C>& pl, int32_t f, int32_t j0,
                                               const C* v, C* strip, C* prev) {
  const bool has_strip = j0 >= pl.band_lo && j0 < pl.band_hi;
  const bool has_prev = WANT_PREV && j0 - 1 >= pl.band_lo && j0 - 1 < pl.band_hi;
  C jt[3][N];
  C jp[3] = {(C)0, (C)0, (C)0};
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int i = 0; i < N; ++i) jt[d][i] = (C)0;
  const int32_t jb0 = j0 - pl.band_lo;
  for (int s = 0; s < pl.n_species; ++s) {
    const C* base = pl.current + s * pl.species_stride + (size_t)f * pl.pitch;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (has_strip) {
        C x[N];
        load_vec<N>(base + d * pl.comp_stride + jb0, x);
#pragma unroll
        for (int i = 0; i < N; ++i) jt[d][i] += x[i];
      }
      if (has_prev) jp[d] += base[d * pl.comp_stride + jb0 - 1];
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) strip[i] = (jt[0][i] * v[0] + jt[1][i] * v[1]) + jt[2][i] * v[2];
  if constexpr (WANT_PREV) *prev = (jp[0] * v[0] + jp[1] * v[1]) + jp[2] * v[2];
}

template <typename T, typename C, bool PLASMA, bool MAGNETIZED>
__device__ __forceinline__ void e_vertex_strip(const EParams<T, C>& p, const Scales<C>& sc,
                                               int32_t row, int32_t strip, Lanes ln) {
  constexpr int N = Strip<T>::N;
  // the incidence rows are stored in processing order, so they are loaded
  // together with the vertex id rather than after it
  const int32_t v = p.vertex_order[row];
  const int32_t k0 = strip * N;
  const VertexRow<C> vrow = p.vertex_rows[row];
  const int32_t* ve = vrow.idx;
  const C* vw = vrow.w;
  C curl[N];
#pragma unroll
  for (int i = 0; i < N; ++i) curl[i] = (C)0;
#pragma unroll
  for (int s = 0; s < 6; ++s) {  // padding slots carry weight 0 and a valid edge
    C h[N];
    load_strip(p.ht + (size_t)ve[s] * p.pn + k0, h, sc.ht_inv);
#pragma unroll
    for (int i = 0; i < N; ++i) curl[i] += vw[s] * h[i];
  }
  C jr[N];
#pragma unroll
  for (int i = 0; i < N; ++i) jr[i] = (C)0;
  if constexpr (PLASMA) {
    const PlasmaTables<C>& pl = p.pl;
    // nodes k0..k0+N-1 need midpoint layers k0-1..k0+N-1 (plasma.py:511-527)
    if (k0 + N > pl.band_lo && k0 - 1 < pl.band_hi) {
      C rv[N + 1];
#pragma unroll
      for (int i = 0; i <= N; ++i) rv[i] = (C)0;
      const VertexRow<C> frow = pl.vertex_faces[row];
      const int32_t* vf = frow.idx;
      const C* vfw = frow.w;
      if (!MAGNETIZED && pl.n_species == 1) {
        // One species: the six face strips are loaded three at a time
        // (padding slots name face 0 with weight 0; the scheduling barrier
        // keeps the second three from being hoisted, which holds the kernel
        // at 72 VGPRs and 7 waves per SIMD), and the weighted sum at midpoint
        // k0 - 1 is the left lane's last sum (same vertex, same weights, same
        // order of additions).
        const int32_t jb0 = k0 - pl.band_lo;
        if (k0 >= pl.band_lo && k0 < pl.band_hi) {
#pragma unroll
          for (int g = 0; g < 6; g += 3) {
            C x[3][N];
#pragma unroll
            for (int s = 0; s < 3; ++s) load_vec<N>(pl.current + (size_t)vf[g + s] * pl.pitch + jb0, x[s]);
#pragma unroll
            for (int s = 0; s < 3; ++s)
#pragma unroll
              for (int i = 0; i < N; ++i) rv[i + 1] += x[s][i] * vfw[g + s];
            __builtin_amdgcn_sched_barrier(0);
          }
        }
        const C from_left = __shfl_up(rv[N], 1);
        const bool has_prev = k0 - 1 >= pl.band_lo && k0 - 1 < pl.band_hi;
        if (ln.left && has_prev) {
          rv[0] = from_left;
        } else if (has_prev) {
          C y[6];
#pragma unroll
          for (int s = 0; s < 6; ++s) y[s] = pl.current[(size_t)vf[s] * pl.pitch + jb0 - 1];
#pragma unroll
          for (int s = 0; s < 6; ++s) rv[0] += y[s] * vfw[s];
        }
      } else for (int s = 0; s < 6; ++s) {
        if (vfw[s] == (C)0) continue;
        C strip_v[N], prev;
        if constexpr (MAGNETIZED) {
          C nrm[4];
          load_vec<4>(pl.center + 4 * vf[s], nrm);
          plasma_project<C, N, true>(pl, vf[s], k0, nrm, strip_v, &prev);
        } else {
          plasma_component<C, N, false>(pl, vf[s], 0, k0, strip_v, &prev);
          const C from_left = __shfl_up(strip_v[N - 1], 1);
          const bool has_prev = k0 - 1 >= pl.band_lo && k0 - 1 < pl.band_hi;
          if (ln.left && has_prev) {
            prev = from_left;
          } else {
            C unused[N];
            plasma_component<C, N, true>(pl, vf[s], 0, k0, unused, &prev);
          }
        }
        rv[0] += prev * vfw[s];
#pragma unroll
        for (int i = 0; i < N; ++i) rv[i + 1] += strip_v[i] * vfw[s];
      }
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int32_t k = k0 + i;
        if (k == 0) jr[i] = rv[1];
        else if (k == p.R) jr[i] = rv[i];
        else jr[i] = (C)0.5 * (rv[i] + rv[i + 1]);
      }
    }
  }
  // The own row and its coefficients are loaded only now: live across the
  // plasma sums they held the kernel at 96 VGPRs and 5 waves per SIMD
  // against 80 and 6 (before the three-at-a-time grouping of those sums).
  T* erow = p.er + (size_t)v * p.pn + k0;
  C er[N], cl[N], cb[N], ir[N];
  load_strip(erow, er, sc.er_inv);
  load_vec<N>(p.cl_er + (size_t)v * p.ca_er_stride + k0, cl);
  load_vec<N>(p.cb_er + (size_t)v * p.ca_er_stride + k0, cb);
  if (sizeof(C) == 4 && p.rg.uniform) {
#pragma unroll
    for (int i = 0; i < N; ++i) ir[i] = k0 + i <= p.R ? inv_radius(p.rg, (C)(k0 + i)) : (C)0;
  } else {
    load_vec<N>(p.inv_r + k0, ir);
  }
  C src[N];
#pragma unroll
  for (int i = 0; i < N; ++i) src[i] = (C)0;
  for (int q = 0; q < p.n_src; ++q) {
    const int32_t i = p.src_layer[q] - k0;
    if (p.src_row[q] == v && i >= 0 && i < N) {
      const C* cur = reinterpret_cast<const C*>(p.dyn->currents);
      const C value = p.src_factor[q] * (cur != nullptr ? cur[p.dyn->counter - 1] : (C)0);
#pragma unroll
      for (int t = 0; t < N; ++t)
        if (t == i) src[t] += value;
    }
  }
  // Exponential loss in increment form, Er + (Cb * drive - (1 - Ca) * Er):
  // 1 - Ca = -expm1(-sigma dt / eps) keeps its relative precision where
  // Ca rounds to within an ulp of 1 (sigma ~ 1e-12 S/m near the ground).
  // cl/cb are zero on padding lanes, which therefore stay exactly zero.
#pragma unroll
  for (int i = 0; i < N; ++i)
    er[i] = er[i] + (cb[i] * (ir[i] * curl[i] - jr[i] - src[i]) - cl[i] * er[i]);
  store_strip(erow, er, sc.er, sc.rd_er, (uint32_t)v * (uint32_t)p.pn + (uint32_t)k0);
}

template <typename T, typename C, bool PLASMA, bool MAGNETIZED>
__device__ __forceinline__ void e_edge_strip(const EParams<T, C>& p, const Scales<C>& sc, int32_t e,
                                             int32_t strip, Lanes ln) {
  constexpr int N = Strip<T>::N;
  const int32_t j0 = strip * N;
  const DualEdgeRow<C> row = p.edge_rows[e];
  const int32_t fl = row.a, fr = row.b;
  const C inv_phi = row.w;
  C hl[N], hrr[N], ht[N + 1], et[N];
  load_strip(p.hr + (size_t)fl * p.pm + j0, hl, sc.hr_inv);
  load_strip(p.hr + (size_t)fr * p.pm + j0, hrr, sc.hr_inv);
  const T* hrow = p.ht + (size_t)e * p.pn;
  load_strip(hrow + j0, ht, sc.ht_inv);
  T* erow = p.et + (size_t)e * p.pm + j0;
  load_strip(erow, et, sc.et_inv);
  // Loss coefficients and radial weights.  Narrow storage (8 values per
  // strip) requests them only after the plasma reads: issued with the
  // other loads they held the FP16 kernel at 108 VGPRs and 4 waves per SIMD
  // against 92 and 5 (s7 x 64 FP16 E 0.122 -> 0.109 ms, s8 x 96 0.725 ->
  // 0.680 ms).  They are a few radial profiles that stay in L2.
  C cl[N], cb[N], irb[N], ra[N];
  auto load_tables = [&]() {
    load_vec<N>(p.cl_et + (size_t)e * p.ca_et_stride + j0, cl);
    load_vec<N>(p.cb_et + (size_t)e * p.ca_et_stride + j0, cb);
    if (sizeof(C) == 4 && p.rg.uniform) {
      // alpha'_j = r_j / (dr rbar_j) = 1/dr - 1/(2 rbar_j) (spherical), 1/dr (flat)
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int32_t j = j0 + i;
        const C inv = j < p.R ? inv_radius(p.rg, (C)j + (C)0.5) : (C)0;
        irb[i] = inv;
        ra[i] = j < p.R ? p.rg.inv_dr - (p.spherical ? (C)0.5 * inv : (C)0) : (C)0;
      }
    } else {
      load_vec<N>(p.inv_rbar + j0, irb);
      load_vec<N>(p.rh_a + j0, ra);
    }
  };
  if constexpr (!is_narrow<T>()) load_tables();
  // every load (the plasma reads below included; for narrow storage all but
  // the tables) is issued before the first wait, the lane shuffle after them
  C ht_next = (C)0;
  if (!ln.right && j0 + N <= p.R) ht_next = load_one(hrow + j0 + N, sc.ht_inv);
  C jt[N];
#pragma unroll
  for (int i = 0; i < N; ++i) jt[i] = (C)0;
  if constexpr (PLASMA) {
    const PlasmaTables<C>& pl = p.pl;
    if (j0 >= pl.band_lo && j0 < pl.band_hi) {
      C sl[N], sr[N], unused;
      if constexpr (MAGNETIZED) {
        C tl[4], tr[4];
        load_vec<4>(pl.left_tan + 4 * e, tl);
        load_vec<4>(pl.right_tan + 4 * e, tr);
        plasma_project<C, N, false>(pl, fl, j0, tl, sl, &unused);
        plasma_project<C, N, false>(pl, fr, j0, tr, sr, &unused);
      } else {
        // J.t of the edge's slot in each face: slots 0 and 1 are stored,
        // slot 2 is w1 J.t0 + w2 J.t1 (J.t0 read in place of the stored
        // value, J.t1 read in addition); weights (1, 0) leave a stored
        // slot's value unchanged
        const int32_t slot_l = row.slots & 0xff, slot_r = row.slots >> 8;
        const int32_t ql = slot_l < 2 ? 1 + slot_l : 1, qr = slot_r < 2 ? 1 + slot_r : 1;
        C xl[N], xr[N];
        C sw[4];
        load_vec<4>(pl.side_w + 4 * e, sw);
        const C wl1 = sw[0], wl2 = sw[1], wr1 = sw[2], wr2 = sw[3];
#pragma unroll
        for (int i = 0; i < N; ++i) xl[i] = xr[i] = (C)0;
        if (pl.n_species == 1) {  // every side's loads issued together
          const int32_t jb0 = j0 - pl.band_lo;
          load_vec<N>(pl.current + ql * pl.comp_stride + (size_t)fl * pl.pitch + jb0, sl);
          load_vec<N>(pl.current + qr * pl.comp_stride + (size_t)fr * pl.pitch + jb0, sr);
          if (slot_l == 2) load_vec<N>(pl.current + 2 * pl.comp_stride + (size_t)fl * pl.pitch + jb0, xl);
          if (slot_r == 2) load_vec<N>(pl.current + 2 * pl.comp_stride + (size_t)fr * pl.pitch + jb0, xr);
        } else {
          plasma_component<C, N, false>(pl, fl, ql, j0, sl, &unused);
          plasma_component<C, N, false>(pl, fr, qr, j0, sr, &unused);
          if (slot_l == 2) plasma_component<C, N, false>(pl, fl, 2, j0, xl, &unused);
          if (slot_r == 2) plasma_component<C, N, false>(pl, fr, 2, j0, xr, &unused);
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
          sl[i] = wl1 * sl[i] + wl2 * xl[i];
          sr[i] = wr1 * sr[i] + wr2 * xr[i];
        }
      }
#pragma unroll
      for (int i = 0; i < N; ++i) jt[i] = (C)0.5 * (sl[i] + sr[i]);
    }
  }
  if constexpr (is_narrow<T>()) load_tables();
  const C from_right = __shfl_down(ht[0], 1);
  ht[N] = ln.right ? from_right : ht_next;
  C src[N];
#pragma unroll
  for (int i = 0; i < N; ++i) src[i] = (C)0;
  for (int q = 0; q < p.n_tsrc; ++q) {
    const int32_t i = p.tsrc_layer[q] - j0;
    if (p.tsrc_row[q] == e && i >= 0 && i < N) {
      const C* cur = reinterpret_cast<const C*>(p.dyn->currents);
      const C value = p.tsrc_factor[q] * (cur != nullptr ? cur[p.dyn->counter - 1] : (C)0);
#pragma unroll
      for (int t = 0; t < N; ++t)
        if (t == i) src[t] += value;
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const C grad = irb[i] * inv_phi * (hl[i] - hrr[i]);
    const C rd = ra[i] * (ht[i + 1] - ht[i]) + (p.spherical ? irb[i] : (C)0) * ht[i + 1];
    et[i] = et[i] + (cb[i] * (grad - rd - jt[i] - src[i]) - cl[i] * et[i]);
  }
  store_strip(erow, et, sc.et, sc.rd_et, (uint32_t)e * (uint32_t)p.pm + (uint32_t)j0);
}

// Weighted H samples of one trace row: sum_q w[i,q] * H[entity[i,q], layer[i,q]]
// in the order of the reference's (selected * weights).sum(axis=1).  The E
// launch does not write H, so any of its blocks sees the H state after this
// step's H half-step.
template <typename T, typename C>
__device__ __forceinline__ void h_sample_row(const EParams<T, C>& p, const Scales<C>& sc,
                                             C* rt, C* tt, int64_t row, int64_t rows) {
  if (row < 0 || row >= rows) return;
  if (rt != nullptr)
    for (int i = threadIdx.x; i < p.n_hrx_f; i += blockDim.x) {
      C acc = (C)0;
      for (int q = 0; q < p.hrx_fwidth; ++q) {
        const int32_t k = i * p.hrx_fwidth + q;
        acc += load_one(p.hr + (size_t)p.hrx_face[k] * p.pm + p.hrx_flayer[k], sc.hr_inv) *
               p.hrx_fw[k];
      }
      rt[row * p.n_hrx_f + i] = acc;
    }
  if (tt != nullptr)
    for (int i = threadIdx.x; i < p.n_hrx_e; i += blockDim.x) {
      C acc = (C)0;
      for (int q = 0; q < p.hrx_ewidth; ++q) {
        const int32_t k = i * p.hrx_ewidth + q;
        acc += load_one(p.ht + (size_t)p.hrx_edge[k] * p.pn + p.hrx_elayer[k], sc.ht_inv) *
               p.hrx_ew[k];
      }
      tt[row * p.n_hrx_e + i] = acc;
    }
}

// The E launch's H-sampling block: the counter was advanced by this step's
// H launch, so it counts the steps completed with this one.
template <typename T, typename C>
__device__ __forceinline__ void h_sample_block(const EParams<T, C>& p, const Scales<C>& sc) {
  const Dynamic* d = p.dyn;
  if (d->h_trace_r == nullptr && d->h_trace_t == nullptr) return;
  const int64_t c = d->counter;
  const int64_t row = c % d->h_every == 0 ? c / d->h_every
                      : c == d->h_total    ? d->h_total / d->h_every + 1
                                           : -1;
  h_sample_row(p, sc, reinterpret_cast<C*>(d->h_trace_r), reinterpret_cast<C*>(d->h_trace_t),
               row, d->h_rows);
}

template <typename T, typename C, bool PLASMA, bool MAGNETIZED, int S>
__device__ __forceinline__ void e_dispatch(const EParams<T, C>& p, const Scales<C>& sc,
                                           uint32_t seg, int32_t b) {
  if (seg == kSegHSample) {
    h_sample_block(p, sc);
    return;
  }
  int32_t row, col;
  if constexpr (is_narrow<T>()) {
    // the rounding step index of the next H launch: written by one thread,
    // read by no thread of this launch
    if (seg == kSegEVertex && b == 0 && threadIdx.x == 0) p.dyn->sr_h = p.dyn->sr_e + 1u;
  }
  if (seg == kSegEVertex) {
    row_and_strip(p.g_vertex, b, &row, &col);
    if (row < p.g_vertex.n_rows) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int32_t strip = col + s * p.g_vertex.strips;
        if (S == 1 || strip < p.g_vertex.row_strips)
          e_vertex_strip<T, C, PLASMA, MAGNETIZED>(p, sc, row, strip, lanes_of(p.g_vertex, col, strip));
      }
    }
  } else {
    row_and_strip(p.g_edge, b, &row, &col);
    if (row < p.g_edge.n_rows) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int32_t strip = col + s * p.g_edge.strips;
        if (S == 1 || strip < p.g_edge.row_strips)
          e_edge_strip<T, C, PLASMA, MAGNETIZED>(p, sc, row, strip, lanes_of(p.g_edge, col, strip));
      }
    }
  }
}

template <typename T, typename C, bool PLASMA, bool MAGNETIZED>
__global__ void __launch_bounds__(kBlock) e_step_kernel(const EParams<T, C> p) {
  const int32_t L = logical_block(p.per_xcd);
  if (L >= p.n_blocks) return;
  const Scales<C> sc = load_scales<T, C, true>(p.dyn);
  const uint32_t m = p.block_map[L];
  e_dispatch<T, C, PLASMA, MAGNETIZED, 1>(p, sc, m >> 24, (int32_t)(m & 0xffffffu));
}

// ---------------------------------------------------------------------------
// Ensembles: independent runs (own radial grid, coefficients, plasma, ground,
// source and receivers; shared mesh) advanced by the same two launches per
// step.  Each run's parameter block lives in a device array; a map entry
// names the segment, the run and the block within that run's segment.
// ---------------------------------------------------------------------------
constexpr uint32_t kEnsBlockBits = 20;
constexpr uint32_t kEnsRunBits = 9;
constexpr uint32_t kEnsMaxRuns = 1u << kEnsRunBits;
constexpr uint32_t kEnsMaxBlocks = 1u << kEnsBlockBits;

__device__ __forceinline__ void ens_decode(uint32_t m, uint32_t* seg, int32_t* run, int32_t* b) {
  *seg = m >> (kEnsBlockBits + kEnsRunBits);
  *run = (int32_t)((m >> kEnsBlockBits) & (kEnsMaxRuns - 1));
  *b = (int32_t)(m & (kEnsMaxBlocks - 1));
}

template <typename T, typename C, bool PLASMA, bool MAGNETIZED, int S>
__global__ void __launch_bounds__(kBlock) h_ensemble_kernel(
    const HParams<T, C>* __restrict__ runs, const uint32_t* __restrict__ map, int32_t n_blocks,
    int32_t per_xcd) {
  const int32_t L = logical_block(per_xcd);
  if (L >= n_blocks) return;
  uint32_t seg;
  int32_t run, b;
  ens_decode(map[L], &seg, &run, &b);
  const HParams<T, C>& p = runs[run];
  h_dispatch<T, C, PLASMA, MAGNETIZED, S>(p, load_scales<T, C>(p.dyn), seg, b);
}

template <typename T, typename C, bool PLASMA, bool MAGNETIZED, int S>
__global__ void __launch_bounds__(kBlock) e_ensemble_kernel(
    const EParams<T, C>* __restrict__ runs, const uint32_t* __restrict__ map, int32_t n_blocks,
    int32_t per_xcd) {
  const int32_t L = logical_block(per_xcd);
  if (L >= n_blocks) return;
  uint32_t seg;
  int32_t run, b;
  ens_decode(map[L], &seg, &run, &b);
  const EParams<T, C>& p = runs[run];
  e_dispatch<T, C, PLASMA, MAGNETIZED, S>(p, load_scales<T, C, true>(p.dyn), seg, b);
}

// Final trace row of every run (block r samples run r and advances its counter).
template <typename T, typename C>
__global__ void __launch_bounds__(kBlock) sample_runs_kernel(const HParams<T, C>* __restrict__ runs) {
  const HParams<T, C>& p = runs[blockIdx.x];
  sample_block(p, load_scales<T, C>(p.dyn));
}

__global__ void set_dynamic_runs_kernel(Dynamic* const* __restrict__ dyns, int32_t n,
                                        const char* currents, int64_t currents_stride,
                                        char* trace, int64_t trace_stride, int64_t rows,
                                        int64_t n_currents) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Dynamic* d = dyns[i];
  d->counter = 0;
  d->currents = currents != nullptr ? currents + i * currents_stride : nullptr;
  d->n_currents = currents != nullptr ? n_currents : 0;
  d->trace = trace != nullptr ? trace + i * trace_stride : nullptr;
  d->trace_rows = trace != nullptr ? rows : 0;
  d->h_trace_r = nullptr;
  d->h_trace_t = nullptr;
  d->h_rows = 0;
}

// ---------------------------------------------------------------------------
// Small helper kernels
// ---------------------------------------------------------------------------

__global__ void set_dynamic_kernel(Dynamic* dyn, int64_t counter, void* trace, int64_t rows,
                                   const void* currents, int64_t n_currents) {
  dyn->counter = counter;
  dyn->trace = trace;
  dyn->trace_rows = rows;
  dyn->currents = currents;
  dyn->n_currents = currents != nullptr ? n_currents : 0;
  dyn->h_trace_r = nullptr;
  dyn->h_trace_t = nullptr;
  dyn->h_rows = 0;
  dyn->h_every = 1;
  dyn->h_total = 0;
}

// H traces of the call that follows (after set_dynamic_kernel), and row 0:
// the H state before its first step.
template <typename T, typename C>
__global__ void __launch_bounds__(kBlock) set_h_trace_kernel(const EParams<T, C> p, void* rt,
                                                            void* tt, int64_t rows, int64_t every,
                                                            int64_t total) {
  if (threadIdx.x == 0) {
    p.dyn->h_trace_r = rt;
    p.dyn->h_trace_t = tt;
    p.dyn->h_rows = rows;
    p.dyn->h_every = every;
    p.dyn->h_total = total;
  }
  h_sample_row(p, load_scales<T, C, true>(p.dyn), static_cast<C*>(rt), static_cast<C*>(tt),
               0, rows);
}

template <typename T, typename C>
__global__ void sample_kernel(const T* __restrict__ er, int32_t pn, const Dynamic* dyn,
                              const int32_t* __restrict__ rx_vertex,
                              const int32_t* __restrict__ rx_layer, const C* __restrict__ rx_w,
                              int32_t n_rx, C* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rx) return;
  const C s_er_inv = load_scales<T, C>(dyn).er_inv;
  const int32_t layer = rx_layer[i];
  C acc = (C)0;
#pragma unroll
  for (int q = 0; q < 3; ++q)
    acc += load_one(er + (size_t)rx_vertex[3 * i + q] * pn + layer, s_er_inv) * rx_w[3 * i + q];
  out[i] = acc;
}

}  // namespace geo
