// C ABI of the GeodesicFDTD step library (declared in include/geofdtd.h).
//
// geo_plan_create turns the reference's FP64 host tables into the device
// tables the two step kernels read (int32 indices, compute-type metric and
// loss coefficients folded with dt), builds the XCD-aware block maps, and
// owns them until geo_plan_destroy.  Field state stays in caller buffers.
#include "geofdtd.h"
#include "geofdtd_kernels.h"
#include "geofdtd_internal.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

constexpr double kMu0 = 1.25663706212e-6;  // constants.py:4

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define GEO_HIP(expr)                                                                    \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return fail(GEO_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(_e)); \
  } while (0)

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// A storage scale: a power of two in [2^-60, 2^60], so multiplying by it or
// by a ratio of two of them is exact in FP32 and in the narrow types' normal
// range.
bool valid_scale(double s) {
  if (!(s > 0.0) || !std::isfinite(s)) return false;
  int e = 0;
  const double m = std::frexp(s, &e);
  return m == 0.5 && e - 1 >= -60 && e - 1 <= 60;
}

// The kernels' closed-form radial tables (FP32 compute only): used when every
// spacing of the radii is within 1e-9 (relative) of their mean, as for the
// default linspace altitudes, whose spacings differ in the last bits only.
template <typename C>
geo::RadialGrid<C> radial_grid(const std::vector<double>& rad) {
  geo::RadialGrid<C> g{};
  const int64_t R = (int64_t)rad.size() - 1;
  const double dr = (rad[R] - rad[0]) / (double)R;
  bool uniform = sizeof(C) == 4;
  for (int64_t k = 0; k < R && uniform; ++k)
    uniform = std::fabs((rad[k + 1] - rad[k]) - dr) <= 1e-9 * dr;
  g.uniform = uniform ? 1 : 0;
  g.r0_hi = (C)rad[0];
  g.r0_lo = (C)(rad[0] - (double)g.r0_hi);
  g.dr = (C)dr;
  g.inv_dr = (C)(1.0 / dr);
  return g;
}

// Owns device allocations made at plan creation.
struct DeviceArena {
  std::vector<void*> blocks;
  ~DeviceArena() {
    for (void* p : blocks) (void)hipFree(p);
  }
  template <typename X>
  int upload(const std::vector<X>& host, X** out) {
    *out = nullptr;
    if (host.empty()) return GEO_OK;
    void* p = nullptr;
    GEO_HIP(hipMalloc(&p, host.size() * sizeof(X)));
    blocks.push_back(p);
    GEO_HIP(hipMemcpy(p, host.data(), host.size() * sizeof(X), hipMemcpyHostToDevice));
    *out = static_cast<X*>(p);
    return GEO_OK;
  }
  int zeroed(size_t bytes, void** out) {
    void* p = nullptr;
    GEO_HIP(hipMalloc(&p, bytes));
    blocks.push_back(p);
    GEO_HIP(hipMemset(p, 0, bytes));
    *out = p;
    return GEO_OK;
  }
};

struct SegSpec {
  uint32_t segment;
  int32_t blocks;
};

// A block-map entry holds the block index within its segment in 24 bits.
constexpr int32_t kMaxSegmentBlocks = 1 << 24;

bool blocks_fit(const std::vector<SegSpec>& segs) {
  for (const SegSpec& sp : segs)
    if (sp.blocks < 0 || sp.blocks >= kMaxSegmentBlocks) return false;
  return true;
}

// Interleave the blocks of several segments in proportion to their sizes so
// that, e.g., edge and face blocks covering the same mesh region run close
// together; entry = segment << 24 | block index within the segment.
std::vector<uint32_t> build_block_map(const std::vector<SegSpec>& segs) {
  struct Item { double key; uint32_t seg_order; uint32_t value; };
  std::vector<Item> items;
  uint32_t order = 0;
  for (const SegSpec& s : segs) {
    for (int32_t i = 0; i < s.blocks; ++i)
      items.push_back({(i + 0.5) / s.blocks, order, (s.segment << 24) | (uint32_t)i});
    ++order;
  }
  std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
    return a.key < b.key || (a.key == b.key && a.seg_order < b.seg_order);
  });
  std::vector<uint32_t> map(items.size());
  for (size_t i = 0; i < items.size(); ++i) map[i] = items[i].value;
  return map;
}

// Rows of `pitch` elements in 16-byte strips of `strip_elems`, S strips per
// thread; an integer number of rows per 256-thread block (the flat mapping
// is switched on afterwards by set_flat).
geo::SegGeom make_geom(int64_t rows, int64_t pitch, int64_t strip_elems, int32_t S = 1) {
  geo::SegGeom g;
  g.n_rows = (int32_t)rows;
  g.row_strips = (int32_t)(pitch / strip_elems);
  g.strips = (g.row_strips + S - 1) / S;
  g.rows_per_block = geo::kBlock / g.strips;
  g.magic = (uint32_t)((65536 + g.strips - 1) / g.strips);
  g.flat = 0;
  return g;
}

// Lanes a whole-row block leaves without a strip.
int32_t idle_lanes(const geo::SegGeom& g) { return geo::kBlock - g.rows_per_block * g.strips; }

// Whether a kernel whose segments are `a` and `b` runs flat: the mapping
// asked for, or under AUTO when whole-row blocks would leave a sixteenth of
// their lanes idle in either segment.  Measured on the MI355X (ms per
// launch, whole-row -> flat): at 96 radial cells (rows of 25 and 24
// strips, 6 and 16 idle lanes) E 0.311 -> 0.235 and H 0.299 -> 0.288 at
// s7, E 1.235 -> 0.933 at s8; at 80 cells (4 and 16 idle) E 0.169 ->
// 0.139; at 64 cells (1 and 0 idle) flat is slower: E 0.148 -> 0.150,
// H 0.184 -> 0.189.  Flat indexing needs rows * strips < 2^31 (flat_fits).
bool flat_fits(const geo::SegGeom& a, const geo::SegGeom& b) {
  const auto fits = [](const geo::SegGeom& g) {
    return (int64_t)g.n_rows * g.row_strips + geo::kBlock < (int64_t)INT32_MAX;
  };
  return fits(a) && fits(b);
}

bool use_flat(int32_t mapping, const geo::SegGeom& a, const geo::SegGeom& b) {
  if (!flat_fits(a, b) || mapping == GEO_MAPPING_ROWS) return false;
  if (mapping == GEO_MAPPING_FLAT) return true;
  return idle_lanes(a) >= geo::kBlock / 16 || idle_lanes(b) >= geo::kBlock / 16;
}

void set_flat(geo::SegGeom* g) {
  g->flat = 1;
  g->strips = g->row_strips;
}

int32_t geom_blocks(const geo::SegGeom& g) {
  if (g.flat)
    return (int32_t)(((int64_t)g.n_rows * g.row_strips + geo::kBlock - 1) / geo::kBlock);
  return (int32_t)((g.n_rows + g.rows_per_block - 1) / g.rows_per_block);
}

// Step index of the stochastic rounding of narrow stores (sr_h: the next H
// launch rounds with it), so that a run resumed at step n rounds as the
// uninterrupted run does.
__global__ void set_step_index_kernel(geo::Dynamic* dyn, uint32_t index) {
  dyn->sr_h = index;
  dyn->sr_e = index;
}

struct PlanBase {
  virtual ~PlanBase() = default;
  virtual int bind(const GeoState* s) = 0;
  virtual int step(const void* currents, int64_t n, int32_t chunk, void* trace,
                   hipStream_t stream) = 0;
  virtual int prepare(int32_t chunk) = 0;
  virtual int set_step_index(int64_t index, hipStream_t stream) = 0;
  virtual int set_receivers(const int64_t* v, const int64_t* l, const double* w, int32_t n) = 0;
  virtual int set_h_receivers(const int64_t* f, const int64_t* fl, const double* fw, int32_t nf,
                              int32_t wf, const int64_t* e, const int64_t* el, const double* ew,
                              int32_t ne, int32_t we) = 0;
  virtual int step_h(const void* currents, int64_t n, int32_t chunk, int32_t every, void* rt,
                     void* tt, hipStream_t stream) = 0;
  virtual int sample(void* row, hipStream_t stream) = 0;
  virtual int diagnostics(GeoDiagOut* out, hipStream_t stream) = 0;
  virtual int radial_profiles(const GeoProfiles* out, hipStream_t stream) = 0;
  virtual int time_kernels(int64_t n, hipStream_t stream, double* out_ms) = 0;
  virtual int get_scales(double* out, hipStream_t stream) = 0;
  virtual int set_scales(const double* in, hipStream_t stream) = 0;
  GeoLayout layout{};
  uint64_t generation = 0;  // advanced whenever kernel parameters change (bind, receivers)
};

// FP64 reduction kernels for diagnostics: per-block partials, then one block.
struct DiagAcc {
  double max_abs;
  double energy;
  long long nonfinite;
  double loss;
  double sumsq;
};

template <typename T, typename C>
__global__ void diag_field_kernel(const T* __restrict__ x, int64_t rows, int32_t cols,
                                  int32_t pitch, const geo::Dynamic* __restrict__ dyn,
                                  int32_t field, const double* __restrict__ row_w,
                                  const double* __restrict__ col_w,
                                  const double* __restrict__ eps, int64_t eps_stride,
                                  const double* __restrict__ sigma, int64_t sigma_stride,
                                  double coef, DiagAcc* __restrict__ partial) {
  __shared__ double s_max[geo::kBlock];
  __shared__ double s_en[geo::kBlock];
  __shared__ long long s_nf[geo::kBlock];
  __shared__ double s_ls[geo::kBlock];
  __shared__ double s_sq[geo::kBlock];
  double mx = 0.0, en = 0.0, ls = 0.0, sq = 0.0;
  long long nf = 0;
  const C inv_scale = geo::is_narrow<T>() ? (C)dyn->inv_scale[field] : (C)1;
  const int64_t total = rows * cols;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols;
    const int32_t c = (int32_t)(i - r * cols);
    const double v = (double)geo::load_one<T, C>(x + r * pitch + c, inv_scale);
    if (!isfinite(v)) {
      ++nf;
      continue;
    }
    mx = fmax(mx, fabs(v));
    const double e = eps != nullptr ? eps[r * eps_stride + c] : 1.0;
    const double w = v * v * row_w[r] * col_w[c];
    en += coef * e * w;
    if (sigma != nullptr) ls += sigma[r * sigma_stride + c] * w;
    sq += v * v;
  }
  s_max[threadIdx.x] = mx;
  s_en[threadIdx.x] = en;
  s_nf[threadIdx.x] = nf;
  s_ls[threadIdx.x] = ls;
  s_sq[threadIdx.x] = sq;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + s]);
      s_en[threadIdx.x] += s_en[threadIdx.x + s];
      s_nf[threadIdx.x] += s_nf[threadIdx.x + s];
      s_ls[threadIdx.x] += s_ls[threadIdx.x + s];
      s_sq[threadIdx.x] += s_sq[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = {s_max[0], s_en[0], s_nf[0], s_ls[0], s_sq[0]};
}

// Per-layer horizontal sums of one field for the radial profiles: column c of
// partial block b holds sum_r w_r eps(r, c) x(r, c)^2 (eps = 1 when absent) and
// sum_r w_r sigma(r, c) x(r, c)^2 over the block's rows; non-finite values are
// skipped as in diag_field_kernel.  One wave per block, lanes across layers.
template <typename T, typename C>
__global__ void diag_profile_kernel(const T* __restrict__ x, int64_t rows, int32_t cols,
                                    int32_t pitch, const geo::Dynamic* __restrict__ dyn,
                                    int32_t field, const double* __restrict__ row_w,
                                    const double* __restrict__ eps, int64_t eps_stride,
                                    const double* __restrict__ sigma, int64_t sigma_stride,
                                    int64_t rows_per_block, double* __restrict__ partial) {
  const C inv_scale = geo::is_narrow<T>() ? (C)dyn->inv_scale[field] : (C)1;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  for (int32_t c = threadIdx.x; c < cols; c += blockDim.x) {
    double en = 0.0, ls = 0.0;
    for (int64_t r = r0; r < r1; ++r) {
      const double v = (double)geo::load_one<T, C>(x + r * pitch + c, inv_scale);
      if (!isfinite(v)) continue;
      const double q = v * v * row_w[r];
      en += (eps != nullptr ? eps[r * eps_stride + c] : 1.0) * q;
      if (sigma != nullptr) ls += sigma[r * sigma_stride + c] * q;
    }
    partial[((int64_t)blockIdx.x * 2) * cols + c] = en;
    partial[((int64_t)blockIdx.x * 2 + 1) * cols + c] = ls;
  }
}

// Sums the per-block profile partials in block order: out = (2, cols).
__global__ void diag_profile_reduce_kernel(const double* __restrict__ partial, int32_t nb,
                                           int32_t cols, double* __restrict__ out) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * cols;
       i += gridDim.x * blockDim.x) {
    const int32_t q = i / cols, c = i - q * cols;
    double s = 0.0;
    for (int32_t b = 0; b < nb; ++b) s += partial[((int64_t)b * 2 + q) * cols + c];
    out[i] = s;
  }
}

// Discrete div-B per prism cell (f, k) in FP64; partial maxima of |D| and of
// the flux scale per block.
template <typename T, typename C>
__global__ void diag_divb_kernel(const T* __restrict__ hr, const T* __restrict__ ht, int32_t pn,
                                 int32_t pm, int32_t R, int64_t F, int32_t k_start,
                                 const geo::Dynamic* __restrict__ dyn,
                                 const int32_t* __restrict__ face_edges,
                                 const double* __restrict__ face_w,
                                 const double* __restrict__ face_omega,
                                 const double* __restrict__ rb2, const double* __restrict__ rl,
                                 DiagAcc* __restrict__ partial) {
  __shared__ double s_d[geo::kBlock];
  __shared__ double s_s[geo::kBlock];
  double dmax = 0.0, smax = 0.0;
  const geo::Scales<C> sc = geo::load_scales<T, C>(dyn);
  const C s_hr_inv = sc.hr_inv, s_ht_inv = sc.ht_inv;
  const int64_t total = F * (R + 1);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = i / (R + 1);
    const int32_t k = (int32_t)(i - f * (R + 1));
    if (k < k_start) continue;
    const double om = face_omega[f];
    const double up = k < R ? (double)geo::load_one<T, C>(hr + f * pm + k, s_hr_inv) * om * rb2[k] : 0.0;
    const double lo = k > 0 ? (double)geo::load_one<T, C>(hr + f * pm + k - 1, s_hr_inv) * om * rb2[k - 1] : 0.0;
    double tang = 0.0, atang = 0.0;
    for (int c = 0; c < 3; ++c) {
      const int32_t e = face_edges[3 * f + c];
      const double flux = (double)geo::load_one<T, C>(ht + (int64_t)e * pn + k, s_ht_inv) * rl[k];
      tang += face_w[3 * f + c] * flux;
      atang += fabs(face_w[3 * f + c] * flux);
    }
    dmax = fmax(dmax, fabs(up - lo - tang));
    smax = fmax(smax, fabs(up) + fabs(lo) + atang);
  }
  s_d[threadIdx.x] = dmax;
  s_s[threadIdx.x] = smax;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s_d[threadIdx.x] = fmax(s_d[threadIdx.x], s_d[threadIdx.x + s]);
      s_s[threadIdx.x] = fmax(s_s[threadIdx.x], s_s[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = {s_d[0], s_s[0], 0};
}

__global__ void diag_reduce_kernel(const DiagAcc* __restrict__ partial, int32_t n,
                                   DiagAcc* __restrict__ out) {
  __shared__ double s_max[geo::kBlock];
  __shared__ double s_en[geo::kBlock];
  __shared__ long long s_nf[geo::kBlock];
  __shared__ double s_ls[geo::kBlock];
  __shared__ double s_sq[geo::kBlock];
  double mx = 0.0, en = 0.0, ls = 0.0, sq = 0.0;
  long long nf = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    mx = fmax(mx, partial[i].max_abs);
    en += partial[i].energy;
    nf += partial[i].nonfinite;
    ls += partial[i].loss;
    sq += partial[i].sumsq;
  }
  s_max[threadIdx.x] = mx;
  s_en[threadIdx.x] = en;
  s_nf[threadIdx.x] = nf;
  s_ls[threadIdx.x] = ls;
  s_sq[threadIdx.x] = sq;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + s]);
      s_en[threadIdx.x] += s_en[threadIdx.x + s];
      s_nf[threadIdx.x] += s_nf[threadIdx.x + s];
      s_ls[threadIdx.x] += s_ls[threadIdx.x + s];
      s_sq[threadIdx.x] += s_sq[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = {s_max[0], s_en[0], s_nf[0], s_ls[0], s_sq[0]};
}

// Final reduction for the div-B partials: both fields are maxima.
__global__ void diag_reduce_max_kernel(const DiagAcc* __restrict__ partial, int32_t n,
                                       DiagAcc* __restrict__ out) {
  __shared__ double s_a[geo::kBlock];
  __shared__ double s_b[geo::kBlock];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    a = fmax(a, partial[i].max_abs);
    b = fmax(b, partial[i].energy);
  }
  s_a[threadIdx.x] = a;
  s_b[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s_a[threadIdx.x] = fmax(s_a[threadIdx.x], s_a[threadIdx.x + s]);
      s_b[threadIdx.x] = fmax(s_b[threadIdx.x], s_b[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = {s_a[0], s_b[0], 0, 0.0, 0.0};
}

// ---------------------------------------------------------------------------
// Dynamic power-of-two rescaling of narrow (FP16/BF16) field storage.
//
// Before every block of steps (one graph chunk), three small launches:
//   scale_max_kernel     per-block maxima of |stored value| of each field;
//   scale_decide_kernel  one block: physical maxima, the source drive over
//                        the coming steps, and a new power-of-two scale for
//                        every field whose stored bound would leave
//                        [2^7, 2^12] (re-placed at 2^10: 64x headroom below
//                        the FP16 limit 65504; precision does not depend on
//                        the placement while values stay above 2^-14);
//   scale_apply_kernel   multiplies the stored values of the rescaled fields
//                        by the exact ratio of the two scales.
// All state lives in the device block, so captured graphs stay valid.
// ---------------------------------------------------------------------------
constexpr int kScaleBlocks = 512;
constexpr int kRescaleInterval = 32;  // steps between rescale points without graphs
constexpr int kMaxScaleExp = 60;
static_assert(offsetof(geo::Dynamic, inv_scale) == offsetof(geo::Dynamic, scale) + 4 * sizeof(float),
              "set_scales writes scale[] and inv_scale[] with one copy");

template <typename T>
__global__ void __launch_bounds__(geo::kBlock) scale_max_kernel(
    const T* __restrict__ er, const T* __restrict__ et, const T* __restrict__ hr,
    const T* __restrict__ ht, int64_t n_er, int64_t n_et, int64_t n_hr, int64_t n_ht,
    float* __restrict__ partial) {
  __shared__ float red[geo::kBlock];
  const T* base[4] = {er, et, hr, ht};
  const int64_t strips[4] = {n_er, n_et, n_hr, n_ht};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int f = 0; f < 4; ++f) {
    const geo::Strip<T>* src = reinterpret_cast<const geo::Strip<T>*>(base[f]);
    float m = 0.f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < strips[f];
         i += stride) {
      const geo::Strip<T> v = src[i];
#pragma unroll
      for (int l = 0; l < geo::Strip<T>::N; ++l) {
        const float a = fabsf(geo::to_c<float>(v.v[l]));
        m = (a <= FLT_MAX) ? fmaxf(m, a) : INFINITY;  // NaN and inf count as inf
      }
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[4 * blockIdx.x + f] = red[0];
    __syncthreads();
  }
}

// gain_er / gain_et bound the change of Er / Et per step and ampere of source
// current (sum over support points of |Cb * factor|).  The bound of a field
// is that of its family (Er with Et, Hr with Ht) and of the other family
// across the impedance of free space (377 ohm ~ 2^8): a field at roundoff
// level (Hr under a vertical source) or one about to be driven (H fields
// before the first wave, E fields seeded far below eta0 * H) is scaled for
// the values the coupled update can bring within the next chunk.
__global__ void __launch_bounds__(geo::kBlock) scale_decide_kernel(
    geo::Dynamic* __restrict__ dyn, const float* __restrict__ partial, int32_t n_partial,
    int32_t interval, float gain_er, float gain_et) {
  __shared__ float red[5][geo::kBlock];
  const int t = threadIdx.x;
  float m[4] = {0.f, 0.f, 0.f, 0.f};
  for (int b = t; b < n_partial; b += blockDim.x)
    for (int f = 0; f < 4; ++f) m[f] = fmaxf(m[f], partial[4 * b + f]);
  // the step that starts with the device counter at c injects currents[c]
  float drive = 0.f;
  const float* cur = static_cast<const float*>(dyn->currents);
  if (cur != nullptr) {
    const int64_t c0 = dyn->counter;
    const int64_t c1 = c0 + interval < dyn->n_currents ? c0 + interval : dyn->n_currents;
    for (int64_t i = c0 + t; i < c1; i += blockDim.x) drive += fabsf(cur[i]);
  }
  for (int f = 0; f < 4; ++f) red[f][t] = m[f];
  red[4][t] = drive;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) {
      for (int f = 0; f < 4; ++f) red[f][t] = fmaxf(red[f][t], red[f][t + s]);
      red[4][t] += red[4][t + s];
    }
    __syncthreads();
  }
  if (t != 0) return;
  constexpr float kHi = 4096.f, kLo = 128.f, kTarget = 1024.f, kEta = 256.f;
  float phys[4];
  bool bad = false;
  for (int f = 0; f < 4; ++f) {
    if (!(red[f][0] <= FLT_MAX)) bad = true;
    phys[f] = red[f][0] * dyn->inv_scale[f];
  }
  if (bad) {
    dyn->overflow = 1;
    for (int f = 0; f < 4; ++f) dyn->factor[f] = 1.f;
    return;
  }
  const float d = red[4][0];
  const float b[4] = {phys[0] + gain_er * d, phys[1] + gain_et * d, phys[2], phys[3]};
  const float eb = fmaxf(b[0], b[1]), hb = fmaxf(b[2], b[3]);
  const float be = fmaxf(eb, kEta * hb), bh = fmaxf(hb, eb / kEta);
  const float bounds[4] = {be, be, bh, bh};
  for (int f = 0; f < 4; ++f) {
    const float bound = bounds[f];
    float fac = 1.f;
    if (bound > 0.f && bound <= FLT_MAX) {
      const float stored = bound * dyn->scale[f];
      if (stored > kHi || stored < kLo) {
        int k = (int)floorf(log2f(kTarget / bound));
        k = k > kMaxScaleExp ? kMaxScaleExp : (k < -kMaxScaleExp ? -kMaxScaleExp : k);
        const float ns = ldexpf(1.f, k);
        if (phys[f] > 0.f) fac = ns * dyn->inv_scale[f];  // ratio of powers of two: exact
        dyn->scale[f] = ns;
        dyn->inv_scale[f] = ldexpf(1.f, -k);
      }
    }
    dyn->factor[f] = fac;
  }
}

template <typename T>
__global__ void __launch_bounds__(geo::kBlock) scale_apply_kernel(
    T* __restrict__ er, T* __restrict__ et, T* __restrict__ hr, T* __restrict__ ht,
    int64_t n_er, int64_t n_et, int64_t n_hr, int64_t n_ht,
    const geo::Dynamic* __restrict__ dyn) {
  T* base[4] = {er, et, hr, ht};
  const int64_t strips[4] = {n_er, n_et, n_hr, n_ht};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int f = 0; f < 4; ++f) {
    const float fac = dyn->factor[f];
    if (fac == 1.f) continue;
    geo::Strip<T>* dst = reinterpret_cast<geo::Strip<T>*>(base[f]);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < strips[f];
         i += stride) {
      geo::Strip<T> v = dst[i];
#pragma unroll
      for (int l = 0; l < geo::Strip<T>::N; ++l)
        v.v[l] = geo::from_c<T>(geo::to_c<float>(v.v[l]) * fac);
      dst[i] = v;
    }
  }
}

template <typename T, typename C>
struct Plan final : PlanBase {
  static constexpr int N = geo::Strip<T>::N;
  DeviceArena arena;
  int64_t V = 0, E = 0, F = 0, R = 0, pn = 0, pm = 0;
  bool plasma = false, magnetized = false, si_on = false;
  int32_t si_terms = 0;
  geo::HParams<T, C> hp{};
  geo::EParams<T, C> ep{};
  int32_t h_grid = 0, e_grid = 0;
  bool h_flat = false, e_flat = false;  // thread mapping of each kernel
  geo::Dynamic* dyn = nullptr;
  bool bound = false;
  int32_t rounding_mode = 0;  // narrow stores: 0 nearest, 1 / 2 stochastic thresholds
  // narrow storage with dynamic power-of-two scales
  bool dyn_scaling = false;
  float src_gain_er = 0.f, src_gain_et = 0.f;
  float* scale_partial = nullptr;
  // receivers (owned separately so repeated record calls do not grow the arena)
  int32_t n_rx = 0;
  void* rx_mem = nullptr;
  void* hrx_mem = nullptr;  // H receiver tables
  // permittivity for energy diagnostics (absolute, F/m), rows 1 or entity count
  double* eps_er_dev = nullptr;
  double* eps_et_dev = nullptr;
  int64_t eps_er_stride = 0, eps_et_stride = 0;
  // conductivity (S/m) for the conductive-loss diagnostic; null: no loss
  double* sig_er_dev = nullptr;
  double* sig_et_dev = nullptr;
  int64_t sig_er_stride = 0, sig_et_stride = 0;
  // radial profiles: per-block partials (kProfileBlocks, 2, R+1), per-field
  // sums (4, 2, R+1) and the host copies of the radial weights
  double* prof_partial = nullptr;
  double* prof_out = nullptr;
  std::vector<double> col_w_host[4];
  static constexpr int kProfileBlocks = 512;
  // graph cache
  hipStream_t capture_stream = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  int32_t graph_chunk = 0;
  // diagnostics weights (double, device)
  double *w_er_row = nullptr, *w_er_col = nullptr, *w_et_row = nullptr, *w_et_col = nullptr;
  double *w_hr_row = nullptr, *w_hr_col = nullptr, *w_ht_row = nullptr, *w_ht_col = nullptr;
  DiagAcc* diag_partial = nullptr;
  DiagAcc* diag_out = nullptr;
  double *dv_face_w = nullptr, *dv_face_omega = nullptr, *dv_rb2 = nullptr, *dv_rl = nullptr;
  static constexpr int kDiagBlocks = 1024;

  ~Plan() override {
    if (rx_mem) (void)hipFree(rx_mem);
    if (hrx_mem) (void)hipFree(hrx_mem);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }

  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_chunk = 0;
  }

  int create(const GeoMeshDesc* m, const GeoRadialDesc* r, const GeoPhysicsDesc* ph) {
    V = m->n_vertices;
    E = m->n_edges;
    F = m->n_faces;
    R = r->radial_cells;
    if (V < 3 || E < 3 || F < 2 || R < 2)
      return fail(GEO_ERR_INVALID, "mesh and radial sizes must be positive");
    if (E >= (1LL << 30) || V * (R + N) >= (1LL << 40))
      return fail(GEO_ERR_INVALID, "mesh is too large for 32-bit entity indices");
    if (m->max_degree < 1 || m->max_degree > 6)
      return fail(GEO_ERR_INVALID, "vertex degree must be at most 6");
    pn = round_up(R + 1, N);
    pm = round_up(R, N);
    if (pn / N > geo::kBlock)
      return fail(GEO_ERR_INVALID, "radial columns longer than 256 strips are not supported");
    const double dt = ph->time_step_s;
    if (!(dt > 0.0) || !std::isfinite(dt))
      return fail(GEO_ERR_INVALID, "time step must be finite and positive");
    const double dtm = dt / kMu0;
    // ---- radial tables (solver.py:260-283, 443-465) ----
    std::vector<double> rad(r->radii_m, r->radii_m + R + 1);
    for (int64_t k = 0; k < R; ++k)
      if (!(rad[k + 1] > rad[k])) return fail(GEO_ERR_INVALID, "radii must increase");
    std::vector<double> mid(R), dr(R), ctrl(R + 1);
    for (int64_t j = 0; j < R; ++j) {
      mid[j] = 0.5 * (rad[j] + rad[j + 1]);
      dr[j] = rad[j + 1] - rad[j];
    }
    ctrl[0] = 0.5 * dr[0];
    ctrl[R] = 0.5 * dr[R - 1];
    for (int64_t k = 1; k < R; ++k) ctrl[k] = mid[k] - mid[k - 1];
    const bool full = r->full_spherical != 0;
    // gamma_k = 1/r_k (interior nodes) and gamma'_j = 1/rbar_j of the
    // spherical radial derivative are the inverse radii themselves; the
    // kernels take them from inv_r / inv_rbar.
    std::vector<C> inv_r(pn, 0), rd_b(pn, 0), inv_rbar(pm, 0), rh_a(pm, 0);
    for (int64_t k = 0; k <= R; ++k) {
      inv_r[k] = (C)(1.0 / rad[k]);
      double alpha;
      if (k == 0) alpha = full ? 2.0 * mid[0] / (dr[0] * rad[0]) : 2.0 / dr[0];
      else if (k == R) alpha = full ? 2.0 * mid[R - 1] / (dr[R - 1] * rad[R]) : 2.0 / dr[R - 1];
      else alpha = full ? mid[k - 1] / ((mid[k] - mid[k - 1]) * rad[k]) : 1.0 / (mid[k] - mid[k - 1]);
      rd_b[k] = (C)(dtm * alpha);
    }
    for (int64_t j = 0; j < R; ++j) {
      inv_rbar[j] = (C)(1.0 / mid[j]);
      rh_a[j] = (C)(full ? rad[j] / (dr[j] * mid[j]) : 1.0 / dr[j]);
    }
    // ---- topology and metric (solver.py:408-442, numpy_backend.py:46-114) ----
    std::vector<geo::EdgeRow<C>> hedges(E);
    std::vector<geo::DualEdgeRow<C>> eedges(E);
    std::vector<int32_t> face_edges(3 * F), faces(3 * F);
    std::vector<C> face_curl(3 * F);
    for (int64_t e = 0; e < E; ++e) {
      const int64_t t = m->edges[2 * e], h = m->edges[2 * e + 1];
      const int64_t l = m->edge_left_faces[e], rr = m->edge_right_faces[e];
      if (t < 0 || t >= V || h < 0 || h >= V || l < 0 || l >= F || rr < 0 || rr >= F)
        return fail(GEO_ERR_INVALID, "edge incidence index out of range");
      const double theta = m->primal_edge_angles[e], phi = m->dual_edge_angles[e];
      if (!(theta > 0.0) || !(phi > 0.0)) return fail(GEO_ERR_INVALID, "degenerate edge");
      hedges[e] = {(int32_t)t, (int32_t)h, (C)(dtm / theta)};
      eedges[e] = {(int32_t)l, (int32_t)rr, (C)(1.0 / phi), 0};
    }
    for (int64_t f = 0; f < F; ++f) {
      const double omega = m->face_solid_angles[f];
      if (!(omega > 0.0)) return fail(GEO_ERR_INVALID, "degenerate face");
      for (int c = 0; c < 3; ++c) {
        const int64_t e = m->face_edges[3 * f + c];
        const int64_t v = m->faces[3 * f + c];
        if (e < 0 || e >= E || v < 0 || v >= V)
          return fail(GEO_ERR_INVALID, "face incidence index out of range");
        face_edges[3 * f + c] = (int32_t)e;
        faces[3 * f + c] = (int32_t)v;
        face_curl[3 * f + c] =
            (C)(dtm * (double)m->face_edge_signs[3 * f + c] * m->primal_edge_angles[e] / omega);
      }
    }
    // Dual-cell incidence, 6 slots per row; padding slots repeat slot 0's
    // edge with weight 0.
    std::vector<geo::VertexRow<C>> vrows(V);
    std::vector<int32_t> min_edge(V, (int32_t)E);
    for (int64_t v = 0; v < V; ++v) {
      const double area = m->dual_cell_solid_angles[v];
      if (!(area > 0.0)) return fail(GEO_ERR_INVALID, "degenerate dual cell");
      for (int64_t s = 0; s < m->max_degree; ++s) {
        const int64_t e = m->vertex_edges[v * m->max_degree + s];
        const int64_t sg = m->vertex_edge_signs[v * m->max_degree + s];
        if (e < 0 || e >= E) return fail(GEO_ERR_INVALID, "vertex incidence out of range");
        vrows[v].idx[s] = (int32_t)e;
        vrows[v].w[s] = (C)((double)sg * m->dual_edge_angles[e] / area);
        if (sg != 0) min_edge[v] = std::min(min_edge[v], (int32_t)e);
      }
      for (int64_t s = m->max_degree; s < 6; ++s) {
        vrows[v].idx[s] = vrows[v].idx[0];
        vrows[v].w[s] = (C)0;
      }
    }
    // E-kernel processing order of the dual cells: by first incident edge, so
    // the Ht columns a vertex gathers are the ones the edge segment streams at
    // the same time (storage order stays the reference's).
    std::vector<int32_t> vorder(V);
    for (int64_t v = 0; v < V; ++v) vorder[v] = (int32_t)v;
    std::stable_sort(vorder.begin(), vorder.end(),
                     [&](int32_t a, int32_t b) { return min_edge[a] < min_edge[b]; });
    {  // incidence rows in processing order (one dependent load fewer per strip)
      std::vector<geo::VertexRow<C>> by_row(V);
      for (int64_t r = 0; r < V; ++r) by_row[r] = vrows[vorder[r]];
      vrows.swap(by_row);
    }
    // ---- loss coefficients ----
    const int64_t er_rows = ph->er_coef_rows, et_rows = ph->et_coef_rows;
    if (!(er_rows == 1 || er_rows == V) || !(et_rows == 1 || et_rows == E))
      return fail(GEO_ERR_INVALID, "coefficient rows must be 1 or the entity count");
    // loss tables carry 1 - Ca (the kernels use the increment form)
    std::vector<C> cl_er(er_rows * pn, 0), cb_er(er_rows * pn, 0), cl_et(et_rows * pm, 0),
        cb_et(et_rows * pm, 0);
    for (int64_t i = 0; i < er_rows; ++i)
      for (int64_t k = 0; k <= R; ++k) {
        cl_er[i * pn + k] = (C)(1.0 - ph->ca_er[i * (R + 1) + k]);  // exact for Ca >= 1/2
        cb_er[i * pn + k] = (C)ph->cb_er[i * (R + 1) + k];
      }
    for (int64_t i = 0; i < et_rows; ++i)
      for (int64_t j = 0; j < R; ++j) {
        cl_et[i * pm + j] = (C)(1.0 - ph->ca_et[i * R + j]);
        cb_et[i * pm + j] = (C)ph->cb_et[i * R + j];
      }
    const bool narrow = sizeof(T) == 2;
    const double s_er = narrow ? ph->scale_er : 1.0, s_et = narrow ? ph->scale_et : 1.0;
    const double s_hr = narrow ? ph->scale_hr : 1.0, s_ht = narrow ? ph->scale_ht : 1.0;
    for (const double sv : {s_er, s_et, s_hr, s_ht})
      if (!valid_scale(sv))
        return fail(GEO_ERR_INVALID, "storage scales must be powers of two in [2^-60, 2^60]");
    if (ph->scaling != 0 && ph->scaling != 1)
      return fail(GEO_ERR_INVALID, "storage scaling must be 0 (static) or 1 (dynamic)");
    dyn_scaling = narrow && ph->scaling == 1;
    if (ph->thread_mapping < GEO_MAPPING_AUTO || ph->thread_mapping > GEO_MAPPING_FLAT)
      return fail(GEO_ERR_INVALID, "thread mapping must be 0 (auto), 1 (rows) or 2 (flat)");
    if (ph->storage_rounding < 0 || ph->storage_rounding > 2)
      return fail(GEO_ERR_INVALID,
                  "storage rounding must be 0 (nearest), 1 (independent) or 2 (rotation)");
    rounding_mode = narrow ? ph->storage_rounding : 0;
    // ---- surface impedance ----
    si_terms = ph->si_terms;
    si_on = si_terms > 0;
    if (si_terms < 0 || si_terms > geo::kMaxSiTerms)
      return fail(GEO_ERR_INVALID, "surface impedance terms out of range");
    std::vector<C> si_scale;
    if (si_on) {
      if (!(ph->si_scale_rows == 1 || ph->si_scale_rows == E))
        return fail(GEO_ERR_INVALID, "surface impedance scale rows must be 1 or E");
      si_scale.assign(ph->si_scale, ph->si_scale + ph->si_scale_rows);
    }
    // ---- sources ----
    if (ph->n_source < 0 || ph->n_source > geo::kMaxSource || ph->n_tsource < 0 ||
        ph->n_tsource > geo::kMaxSource)
      return fail(GEO_ERR_INVALID, "too many source support points");
    // ---- plasma ----
    plasma = ph->n_species > 0;
    magnetized = plasma && ph->magnetized != 0;
    int64_t band_lo = 0, band_hi = 0, ppitch = 0;
    std::vector<C> recon, center, ltan, rtan, coef, bdir, sidew;
    std::vector<int32_t> eslots;
    std::vector<geo::VertexRow<C>> vfaces;
    int64_t ncomp = 0;
    if (plasma) {
      if (ph->n_species > geo::kMaxSpecies) return fail(GEO_ERR_INVALID, "too many species");
      if (ph->band_lo < 0 || ph->band_hi > R || ph->band_lo >= ph->band_hi)
        return fail(GEO_ERR_INVALID, "plasma band is invalid");
      band_lo = ph->band_lo / N * N;
      band_hi = ph->band_hi;
      ppitch = round_up(band_hi - band_lo, N);
      ncomp = 3;
      if (magnetized) {
        recon.assign(ph->plasma_reconstruction, ph->plasma_reconstruction + 9 * F);
      } else {
        // Projected state for B = 0: the ADE is linear and isotropic, so the
        // projections the scatter needs (J.n, J.t_c) obey the same update as
        // J with E.n = radial average and E.t_c = sum_c' P[c][c'] Et_c',
        // P[c][c'] = t_c . R_c' (R = pinv(T).T, plasma.py:548-559).  The
        // three edge tangents of a face are orthogonal to its centre, hence
        // coplanar: t_2 = w1 t_0 + w2 t_1, so the state is J.n, J.t_0, J.t_1
        // and the scatter forms J.t_2 = w1 J.t_0 + w2 J.t_1.
        if (!ph->face_tangents) return fail(GEO_ERR_INVALID, "face tangents are required");
        recon.assign(8 * F, 0);
        for (int64_t f = 0; f < F; ++f) {
          const double* t = ph->face_tangents + 9 * f;
          for (int c = 0; c < 2; ++c)
            for (int cp = 0; cp < 3; ++cp) {
              double dot = 0.0;
              for (int d = 0; d < 3; ++d)
                dot += t[3 * c + d] * ph->plasma_reconstruction[(f * 3 + cp) * 3 + d];
              recon[8 * f + 3 * c + cp] = (C)dot;
            }
          // least squares t_2 ~ w1 t_0 + w2 t_1 (exact for coplanar tangents)
          double g00 = 0, g01 = 0, g11 = 0, b0 = 0, b1 = 0;
          for (int d = 0; d < 3; ++d) {
            g00 += t[d] * t[d];
            g01 += t[d] * t[3 + d];
            g11 += t[3 + d] * t[3 + d];
            b0 += t[d] * t[6 + d];
            b1 += t[3 + d] * t[6 + d];
          }
          const double det = g00 * g11 - g01 * g01;
          if (!(std::fabs(det) > 1e-12 * g00 * g11))
            return fail(GEO_ERR_INVALID, "face tangents are degenerate");
          recon[8 * f + 6] = (C)((b0 * g11 - b1 * g01) / det);
          recon[8 * f + 7] = (C)((b1 * g00 - b0 * g01) / det);
        }
        // local slot of every edge in its left and right face
        eslots.assign(2 * E, 0);
        for (int64_t e = 0; e < E; ++e)
          for (int side = 0; side < 2; ++side) {
            const int64_t f = side == 0 ? m->edge_left_faces[e] : m->edge_right_faces[e];
            int slot = -1;
            for (int c = 0; c < 3; ++c)
              if (m->face_edges[3 * f + c] == e) slot = c;
            if (slot < 0) return fail(GEO_ERR_INVALID, "edge is absent from an adjacent face");
            eslots[2 * e + side] = slot;
            eedges[e].slots |= slot << (8 * side);
          }
        sidew.assign(4 * E, 0);
        for (int64_t e = 0; e < E; ++e)
          for (int side = 0; side < 2; ++side) {
            const int64_t f = side == 0 ? m->edge_left_faces[e] : m->edge_right_faces[e];
            const bool dep = eslots[2 * e + side] == 2;
            sidew[4 * e + 2 * side] = dep ? recon[8 * f + 6] : (C)1;
            sidew[4 * e + 2 * side + 1] = dep ? recon[8 * f + 7] : (C)0;
          }
      }
      if (magnetized) {  // Cartesian state: normals and edge tangents on the device
        if (!ph->face_centers || !ph->left_tangents || !ph->right_tangents)
          return fail(GEO_ERR_INVALID, "face centres and edge tangents are required");
        center.assign(4 * F, 0);
        for (int64_t f = 0; f < F; ++f)
          for (int d = 0; d < 3; ++d) center[4 * f + d] = (C)ph->face_centers[3 * f + d];
        ltan.assign(4 * E, 0);
        rtan.assign(4 * E, 0);
        for (int64_t e = 0; e < E; ++e)
          for (int d = 0; d < 3; ++d) {
            ltan[4 * e + d] = (C)ph->left_tangents[3 * e + d];
            rtan[4 * e + d] = (C)ph->right_tangents[3 * e + d];
          }
      }
      if (ph->vf_width < 1 || ph->vf_width > 6)
        return fail(GEO_ERR_INVALID, "vertex-face width must be at most 6");
      // padding slots name face 0 with weight 0
      vfaces.assign(V, geo::VertexRow<C>{});
      for (int64_t v = 0; v < V; ++v)
        for (int64_t s = 0; s < ph->vf_width; ++s) {
          const int64_t f = ph->vertex_faces[v * ph->vf_width + s];
          if (f < 0 || f >= F) return fail(GEO_ERR_INVALID, "vertex face out of range");
          vfaces[v].idx[s] = (int32_t)f;
          vfaces[v].w[s] = (C)ph->vertex_face_weights[v * ph->vf_width + s];
        }
      {
        std::vector<geo::VertexRow<C>> by_row(V);
        for (int64_t r = 0; r < V; ++r) by_row[r] = vfaces[vorder[r]];
        vfaces.swap(by_row);
      }
      const int64_t rows = ph->plasma_coef_rows;
      if (!(rows == 1 || rows == F)) return fail(GEO_ERR_INVALID, "plasma rows must be 1 or F");
      coef.assign((size_t)ph->n_species * 6 * rows * ppitch, 0);
      for (int64_t s = 0; s < ph->n_species; ++s)
        for (int q = 0; q < 6; ++q)
          for (int64_t i = 0; i < rows; ++i)
            for (int64_t j = band_lo; j < band_hi; ++j)
              // slots 0 (exp(-nu dt)) and 1 (cos(omega_c dt)) are carried as
              // 1 - value, exact in FP64 for values >= 1/2
              coef[((s * 6 + q) * rows + i) * ppitch + (j - band_lo)] =
                  (C)(q < 2 ? 1.0 - ph->plasma_coef[((s * 6 + q) * rows + i) * R + j]
                            : ph->plasma_coef[((s * 6 + q) * rows + i) * R + j]);
      if (magnetized) {
        const int64_t brows = ph->plasma_b_rows;
        if (!(brows == 1 || brows == F)) return fail(GEO_ERR_INVALID, "B rows must be 1 or F");
        bdir.assign((size_t)brows * ppitch * 3, 0);
        for (int64_t i = 0; i < brows; ++i)
          for (int64_t j = band_lo; j < band_hi; ++j)
            for (int d = 0; d < 3; ++d)
              bdir[(i * ppitch + (j - band_lo)) * 3 + d] = (C)ph->plasma_b[(i * R + j) * 3 + d];
      }
    }
    // ---- upload ----
    int rc;
#define UP(vec, ptr) \
  if ((rc = arena.upload(vec, &ptr)) != GEO_OK) return rc
    geo::EdgeRow<C>* d_hedges;
    geo::DualEdgeRow<C>* d_eedges;
    geo::VertexRow<C> *d_vrows, *d_vf = nullptr;
    int32_t *d_fe, *d_faces, *d_vo;
    C *d_fc, *d_ir, *d_rdb, *d_irb, *d_rha;
    C *d_caer, *d_cber, *d_caet, *d_cbet, *d_sis = nullptr;
    C *d_recon = nullptr, *d_center = nullptr, *d_lt = nullptr, *d_rt = nullptr;
    C *d_coef = nullptr, *d_b = nullptr;
    C* d_sw = nullptr;
    UP(hedges, d_hedges); UP(eedges, d_eedges); UP(face_edges, d_fe); UP(faces, d_faces);
    UP(vrows, d_vrows); UP(vorder, d_vo); UP(face_curl, d_fc);
    UP(inv_r, d_ir); UP(rd_b, d_rdb); UP(inv_rbar, d_irb);
    UP(rh_a, d_rha); UP(cl_er, d_caer); UP(cb_er, d_cber);
    UP(cl_et, d_caet); UP(cb_et, d_cbet);
    if (si_on) UP(si_scale, d_sis);
    if (plasma) {
      UP(recon, d_recon);
      UP(vfaces, d_vf); UP(coef, d_coef);
      if (magnetized) {
        UP(center, d_center); UP(ltan, d_lt); UP(rtan, d_rt);
        UP(bdir, d_b);
      } else {
        UP(sidew, d_sw);
      }
    }
#undef UP
    void* dyn_mem = nullptr;
    if ((rc = arena.zeroed(sizeof(geo::Dynamic), &dyn_mem)) != GEO_OK) return rc;
    dyn = static_cast<geo::Dynamic*>(dyn_mem);
    {
      geo::Dynamic init{};
      init.sr_mode = rounding_mode;
      const double s0[4] = {s_er, s_et, s_hr, s_ht};
      for (int f = 0; f < 4; ++f) {
        init.scale[f] = (float)s0[f];
        init.inv_scale[f] = (float)(1.0 / s0[f]);
        init.factor[f] = 1.f;
      }
      GEO_HIP(hipMemcpy(dyn, &init, sizeof(init), hipMemcpyHostToDevice));
    }
    if (dyn_scaling) {
      void* sp = nullptr;
      if ((rc = arena.zeroed(sizeof(float) * 4 * kScaleBlocks, &sp)) != GEO_OK) return rc;
      scale_partial = static_cast<float*>(sp);
    }

    geo::PlasmaTables<C> pl{};
    if (plasma) {
      pl.n_species = ph->n_species;
      pl.band_lo = (int32_t)band_lo;
      pl.band_hi = (int32_t)band_hi;
      pl.pitch = (int32_t)ppitch;
      pl.comp_stride = F * ppitch;
      pl.species_stride = ncomp * F * ppitch;
      pl.side_w = d_sw;
      pl.recon = d_recon;
      pl.center = d_center;
      pl.left_tan = d_lt;
      pl.right_tan = d_rt;
      pl.vertex_faces = d_vf;
      pl.coef = d_coef;
      pl.coef_row_stride = ph->plasma_coef_rows == 1 ? 0 : ppitch;
      pl.coef_q_stride = ph->plasma_coef_rows * ppitch;
      pl.coef_species_stride = 6 * ph->plasma_coef_rows * ppitch;
      pl.bdir = d_b;
      pl.b_row_stride = (magnetized && ph->plasma_b_rows == F) ? 3 * ppitch : 0;
    }
    // ---- H kernel parameters and block map ----
    hp.pn = (int32_t)pn;
    hp.pm = (int32_t)pm;
    hp.R = (int32_t)R;
    hp.si_on = si_on ? 1 : 0;
    hp.edge_rows = d_hedges;
    hp.face_edges = d_fe;
    hp.faces = d_faces;
    hp.face_curl = d_fc;
    hp.inv_r = d_ir;
    hp.rd_b = d_rdb;
    hp.gamma_scale = full ? (C)dtm : (C)0;
    hp.inv_rbar = d_irb;
    hp.rg = radial_grid<C>(rad);
    hp.rb_a = (C)(dtm / ((rad[R] - rad[0]) / (double)R));
    hp.si_scale = d_sis;
    hp.si_scale_stride = (si_on && ph->si_scale_rows == E) ? 1 : 0;
    hp.si_terms = si_terms;
    hp.si_pitch = (int32_t)round_up(si_terms, 16 / (int64_t)sizeof(C));
    if (si_on) {
      const double cs = 2.0 * dt / (kMu0 * dr[0]);
      const double metric = full ? mid[0] / rad[0] : 1.0;
      hp.si_cs = (C)cs;
      hp.si_cs_metric = (C)(cs * metric);
      double hsum = 0.0;
      for (int i = 0; i < si_terms; ++i) {
        hp.si_loss[i] = (C)(1.0 - ph->si_decay[i]);
        hp.si_drive[i] = (C)ph->si_drive[i];
        hp.si_hw[i] = (C)ph->si_history_weights[i];
        hsum += (double)(C)ph->si_history_weights[i];
      }
      hp.si_hw_sum = (C)hsum;
    }
    hp.pl = pl;
    hp.dyn = dyn;
    hp.n_edges = (int32_t)E;
    hp.g_edge = make_geom(E, pn, N);
    hp.g_face = make_geom(F, pm, N);
    h_flat = use_flat(ph->thread_mapping, hp.g_edge, hp.g_face);
    if (h_flat) {
      set_flat(&hp.g_edge);
      set_flat(&hp.g_face);
    }
    {
      std::vector<SegSpec> segs = {{geo::kSegHEdge, geom_blocks(hp.g_edge)},
                                   {geo::kSegHFace, geom_blocks(hp.g_face)}};
      segs.push_back({geo::kSegSample, 1});
      if (!blocks_fit(segs))
        return fail(GEO_ERR_INVALID, "mesh too large: a kernel segment needs 2^24 blocks or more");
      std::vector<uint32_t> map = build_block_map(segs);
      uint32_t* d_map;
      if ((rc = arena.upload(map, &d_map)) != GEO_OK) return rc;
      hp.block_map = d_map;
      hp.n_blocks = (int32_t)map.size();
      hp.per_xcd = (hp.n_blocks + geo::kNumXcd - 1) / geo::kNumXcd;
      h_grid = hp.per_xcd * geo::kNumXcd;
    }
    // ---- E kernel parameters and block map ----
    ep.pn = (int32_t)pn;
    ep.pm = (int32_t)pm;
    ep.R = (int32_t)R;
    ep.ca_er_stride = er_rows == 1 ? 0 : (int32_t)pn;
    ep.ca_et_stride = et_rows == 1 ? 0 : (int32_t)pm;
    ep.vertex_order = d_vo;
    ep.vertex_rows = d_vrows;
    ep.edge_rows = d_eedges;
    ep.cl_er = d_caer;
    ep.cb_er = d_cber;
    ep.cl_et = d_caet;
    ep.cb_et = d_cbet;
    ep.inv_r = d_ir;
    ep.inv_rbar = d_irb;
    ep.rh_a = d_rha;
    ep.spherical = full ? 1 : 0;
    ep.rg = hp.rg;
    ep.pl = pl;
    ep.n_src = ph->n_source;
    for (int q = 0; q < ph->n_source; ++q) {
      const int64_t v = ph->source_vertices[q], l = ph->source_layers[q];
      if (v < 0 || v >= V || l < 0 || l > R)
        return fail(GEO_ERR_INVALID, "source support index out of range");
      ep.src_row[q] = (int32_t)v;
      ep.src_layer[q] = (int32_t)l;
      // solver.py:936-945: w * l / (A_v r_k^2 L_k)
      const double factor = ph->source_weights[q] * ph->source_length_m /
                            m->dual_cell_solid_angles[v] / (rad[l] * rad[l]) / ctrl[l];
      ep.src_factor[q] = (C)factor;
      src_gain_er += (float)std::fabs(factor * ph->cb_er[(er_rows == 1 ? 0 : v) * (R + 1) + l]);
    }
    ep.n_tsrc = ph->n_tsource;
    for (int q = 0; q < ph->n_tsource; ++q) {
      const int64_t e = ph->tsource_edges[q], l = ph->tsource_layers[q];
      if (e < 0 || e >= E || l < 0 || l >= R)
        return fail(GEO_ERR_INVALID, "tangential source index out of range");
      ep.tsrc_row[q] = (int32_t)e;
      ep.tsrc_layer[q] = (int32_t)l;
      // solver.py:973-981: w / (phi_e rbar_j dr_j)
      const double factor = ph->tsource_weights[q] / m->dual_edge_angles[e] / mid[l] / dr[l];
      ep.tsrc_factor[q] = (C)factor;
      src_gain_et += (float)std::fabs(factor * ph->cb_et[(et_rows == 1 ? 0 : e) * R + l]);
    }
    ep.dyn = dyn;
    ep.g_vertex = make_geom(V, pn, N);
    ep.g_edge = make_geom(E, pm, N);
    e_flat = use_flat(ph->thread_mapping, ep.g_vertex, ep.g_edge);
    if (e_flat) {
      set_flat(&ep.g_vertex);
      set_flat(&ep.g_edge);
    }
    {
      std::vector<SegSpec> segs = {{geo::kSegEVertex, geom_blocks(ep.g_vertex)},
                                   {geo::kSegEEdge, geom_blocks(ep.g_edge)}};
      segs.push_back({geo::kSegHSample, 1});  // H traces; returns at once without them
      if (!blocks_fit(segs))
        return fail(GEO_ERR_INVALID, "mesh too large: a kernel segment needs 2^24 blocks or more");
      std::vector<uint32_t> map = build_block_map(segs);
      uint32_t* d_map;
      if ((rc = arena.upload(map, &d_map)) != GEO_OK) return rc;
      ep.block_map = d_map;
      ep.n_blocks = (int32_t)map.size();
      ep.per_xcd = (ep.n_blocks + geo::kNumXcd - 1) / geo::kNumXcd;
      e_grid = ep.per_xcd * geo::kNumXcd;
    }
    // ---- diagnostics weights (physics_diagnostics/model.py:78-92) ----
    {
      std::vector<double> eer, eet;
      if (ph->eps_er) eer.assign(ph->eps_er, ph->eps_er + er_rows * (R + 1));
      else eer.assign(R + 1, 8.8541878128e-12);
      if (ph->eps_et) eet.assign(ph->eps_et, ph->eps_et + et_rows * R);
      else eet.assign(R, 8.8541878128e-12);
      if ((rc = arena.upload(eer, &eps_er_dev)) != GEO_OK) return rc;
      if ((rc = arena.upload(eet, &eps_et_dev)) != GEO_OK) return rc;
      eps_er_stride = (ph->eps_er && er_rows > 1) ? R + 1 : 0;
      eps_et_stride = (ph->eps_et && et_rows > 1) ? R : 0;
      if (ph->sigma_er) {
        std::vector<double> s(ph->sigma_er, ph->sigma_er + er_rows * (R + 1));
        if ((rc = arena.upload(s, &sig_er_dev)) != GEO_OK) return rc;
        sig_er_stride = er_rows > 1 ? R + 1 : 0;
      }
      if (ph->sigma_et) {
        std::vector<double> s(ph->sigma_et, ph->sigma_et + et_rows * R);
        if ((rc = arena.upload(s, &sig_et_dev)) != GEO_OK) return rc;
        sig_et_stride = et_rows > 1 ? R : 0;
      }
      std::vector<double> er_row(V), er_col(R + 1), et_row(E), et_col(R), hr_row(F), hr_col(R),
          ht_row(E), ht_col(R + 1);
      for (int64_t v = 0; v < V; ++v) er_row[v] = m->dual_cell_solid_angles[v];
      for (int64_t e = 0; e < E; ++e)
        et_row[e] = ht_row[e] = m->primal_edge_angles[e] * m->dual_edge_angles[e];
      for (int64_t f = 0; f < F; ++f) hr_row[f] = m->face_solid_angles[f];
      for (int64_t k = 0; k <= R; ++k) er_col[k] = ht_col[k] = rad[k] * rad[k] * ctrl[k];
      for (int64_t j = 0; j < R; ++j) et_col[j] = hr_col[j] = mid[j] * mid[j] * dr[j];
      if ((rc = arena.upload(er_row, &w_er_row)) != GEO_OK) return rc;
      if ((rc = arena.upload(er_col, &w_er_col)) != GEO_OK) return rc;
      if ((rc = arena.upload(et_row, &w_et_row)) != GEO_OK) return rc;
      if ((rc = arena.upload(et_col, &w_et_col)) != GEO_OK) return rc;
      if ((rc = arena.upload(hr_row, &w_hr_row)) != GEO_OK) return rc;
      if ((rc = arena.upload(hr_col, &w_hr_col)) != GEO_OK) return rc;
      if ((rc = arena.upload(ht_row, &w_ht_row)) != GEO_OK) return rc;
      if ((rc = arena.upload(ht_col, &w_ht_col)) != GEO_OK) return rc;
      col_w_host[0] = er_col;
      col_w_host[1] = et_col;
      col_w_host[2] = hr_col;
      col_w_host[3] = ht_col;
      void* pp = nullptr;
      if ((rc = arena.zeroed(sizeof(double) * (size_t)(kProfileBlocks + 4) * 2 * (R + 1), &pp)) !=
          GEO_OK)
        return rc;
      prof_partial = static_cast<double*>(pp);
      prof_out = prof_partial + (size_t)kProfileBlocks * 2 * (R + 1);
      void* p = nullptr;
      if ((rc = arena.zeroed(sizeof(DiagAcc) * (kDiagBlocks * 5 + 5), &p)) != GEO_OK) return rc;
      diag_partial = static_cast<DiagAcc*>(p);
      diag_out = diag_partial + kDiagBlocks * 5;
      // div-B tables: s_fc theta_c, Omega_f, rbar_j^2, r_k L_k
      std::vector<double> fw(3 * F), fo(F), rb2(R), rl(R + 1);
      for (int64_t f = 0; f < F; ++f) {
        fo[f] = m->face_solid_angles[f];
        for (int c = 0; c < 3; ++c)
          fw[3 * f + c] = (double)m->face_edge_signs[3 * f + c] *
                          m->primal_edge_angles[m->face_edges[3 * f + c]];
      }
      for (int64_t j = 0; j < R; ++j) rb2[j] = mid[j] * mid[j];
      for (int64_t k = 0; k <= R; ++k) rl[k] = rad[k] * ctrl[k];
      if ((rc = arena.upload(fw, &dv_face_w)) != GEO_OK) return rc;
      if ((rc = arena.upload(fo, &dv_face_omega)) != GEO_OK) return rc;
      if ((rc = arena.upload(rb2, &dv_rb2)) != GEO_OK) return rc;
      if ((rc = arena.upload(rl, &dv_rl)) != GEO_OK) return rc;
    }
    layout.node_pitch = pn;
    layout.mid_pitch = pm;
    layout.band_lo = band_lo;
    layout.band_hi = band_hi;
    layout.plasma_pitch = ppitch;
    layout.plasma_components = ncomp;
    layout.plasma_basis = plasma && !magnetized ? 1 : 0;
    layout.radial_closed_form = hp.rg.uniform;
    layout.h_thread_mapping = h_flat ? GEO_MAPPING_FLAT : GEO_MAPPING_ROWS;
    layout.e_thread_mapping = e_flat ? GEO_MAPPING_FLAT : GEO_MAPPING_ROWS;
    layout.vector_elems = N;
    layout.compute_bytes = sizeof(C);
    layout.si_pitch = si_on ? hp.si_pitch : 0;
    GEO_HIP(hipStreamCreateWithFlags(&capture_stream, hipStreamNonBlocking));
    return GEO_OK;
  }

  static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

  int bind(const GeoState* s) override {
    if (s->er_pitch != pn || s->ht_pitch != pn || s->et_pitch != pm || s->hr_pitch != pm)
      return fail(GEO_ERR_INVALID, "field row pitches do not match the plan layout");
    if (!s->er || !s->ht || !s->et || !s->hr || !aligned16(s->er) || !aligned16(s->ht) ||
        !aligned16(s->et) || !aligned16(s->hr))
      return fail(GEO_ERR_INVALID, "field buffers must be non-null and 16-byte aligned");
    if (si_on && !s->si_memory)
      return fail(GEO_ERR_INVALID, "surface impedance memory buffer is required");
    if (plasma && (!s->plasma_current || s->plasma_pitch != layout.plasma_pitch))
      return fail(GEO_ERR_INVALID, "plasma current buffer or pitch does not match the plan");
    drop_graph();
    hp.er = static_cast<const T*>(s->er);
    hp.et = static_cast<const T*>(s->et);
    hp.ht = static_cast<T*>(s->ht);
    hp.hr = static_cast<T*>(s->hr);
    hp.si_mem = static_cast<C*>(s->si_memory);
    hp.pl.current = static_cast<C*>(s->plasma_current);
    ep.er = static_cast<T*>(s->er);
    ep.et = static_cast<T*>(s->et);
    ep.ht = static_cast<const T*>(s->ht);
    ep.hr = static_cast<const T*>(s->hr);
    ep.pl.current = static_cast<C*>(s->plasma_current);
    bound = true;
    ++generation;
    return GEO_OK;
  }

  int set_receivers(const int64_t* v, const int64_t* l, const double* w, int32_t n) override {
    if (n < 0) return fail(GEO_ERR_INVALID, "receiver count must be non-negative");
    std::vector<int32_t> vv(3 * std::max(n, 1), 0), ll(std::max(n, 1), 0);
    std::vector<C> ww(3 * std::max(n, 1), 0);
    for (int32_t i = 0; i < n; ++i) {
      if (l[i] < 0 || l[i] > R) return fail(GEO_ERR_INVALID, "receiver layer out of range");
      ll[i] = (int32_t)l[i];
      for (int q = 0; q < 3; ++q) {
        if (v[3 * i + q] < 0 || v[3 * i + q] >= V)
          return fail(GEO_ERR_INVALID, "receiver vertex out of range");
        vv[3 * i + q] = (int32_t)v[3 * i + q];
        ww[3 * i + q] = (C)w[3 * i + q];
      }
    }
    drop_graph();
    if (rx_mem) (void)hipFree(rx_mem);
    rx_mem = nullptr;
    const size_t nb = std::max(n, 1);
    const size_t bytes_v = 3 * nb * sizeof(int32_t), bytes_l = nb * sizeof(int32_t);
    const size_t off_w = round_up((int64_t)(bytes_v + bytes_l), 16);
    GEO_HIP(hipMalloc(&rx_mem, off_w + 3 * nb * sizeof(C)));
    char* base = static_cast<char*>(rx_mem);
    GEO_HIP(hipMemcpy(base, vv.data(), bytes_v, hipMemcpyHostToDevice));
    GEO_HIP(hipMemcpy(base + bytes_v, ll.data(), bytes_l, hipMemcpyHostToDevice));
    GEO_HIP(hipMemcpy(base + off_w, ww.data(), 3 * nb * sizeof(C), hipMemcpyHostToDevice));
    int32_t* dv = reinterpret_cast<int32_t*>(base);
    int32_t* dl = reinterpret_cast<int32_t*>(base + bytes_v);
    C* dw = reinterpret_cast<C*>(base + off_w);
    hp.rx_vertex = dv;
    hp.rx_layer = dl;
    hp.rx_w = dw;
    hp.n_rx = n;
    n_rx = n;
    ++generation;
    return GEO_OK;
  }

  // H receivers (record_h_observations, solver.py:1138-1229): rows of
  // `width` (entity, layer, weight) triples, faces on midpoint layers and
  // edges on node layers.
  int set_h_receivers(const int64_t* f, const int64_t* fl, const double* fw, int32_t nf,
                      int32_t wf, const int64_t* e, const int64_t* el, const double* ew,
                      int32_t ne, int32_t we) override {
    if (nf < 0 || ne < 0 || wf < 0 || we < 0)
      return fail(GEO_ERR_INVALID, "H receiver counts and widths must be non-negative");
    const int64_t kf = (int64_t)nf * std::max(wf, 0), ke = (int64_t)ne * std::max(we, 0);
    std::vector<int32_t> idx(std::max<int64_t>(2 * (kf + ke), 1), 0);
    std::vector<C> w(std::max<int64_t>(kf + ke, 1), 0);
    for (int64_t k = 0; k < kf; ++k) {
      if (f[k] < 0 || f[k] >= F) return fail(GEO_ERR_INVALID, "observation face index is out of range");
      if (fl[k] < 0 || fl[k] >= R)
        return fail(GEO_ERR_INVALID, "radial H observation layer is out of range");
      idx[k] = (int32_t)f[k];
      idx[kf + k] = (int32_t)fl[k];
      w[k] = (C)fw[k];
    }
    for (int64_t k = 0; k < ke; ++k) {
      if (e[k] < 0 || e[k] >= E) return fail(GEO_ERR_INVALID, "observation edge index is out of range");
      if (el[k] < 0 || el[k] > R)
        return fail(GEO_ERR_INVALID, "tangential H observation layer is out of range");
      idx[2 * kf + k] = (int32_t)e[k];
      idx[2 * kf + ke + k] = (int32_t)el[k];
      w[kf + k] = (C)ew[k];
    }
    drop_graph();
    if (hrx_mem) (void)hipFree(hrx_mem);
    hrx_mem = nullptr;
    const size_t bi = idx.size() * sizeof(int32_t);
    const size_t off_w = round_up((int64_t)bi, 16);
    GEO_HIP(hipMalloc(&hrx_mem, off_w + w.size() * sizeof(C)));
    char* base = static_cast<char*>(hrx_mem);
    GEO_HIP(hipMemcpy(base, idx.data(), bi, hipMemcpyHostToDevice));
    GEO_HIP(hipMemcpy(base + off_w, w.data(), w.size() * sizeof(C), hipMemcpyHostToDevice));
    const int32_t* di = reinterpret_cast<const int32_t*>(base);
    const C* dw = reinterpret_cast<const C*>(base + off_w);
    ep.hrx_face = di;
    ep.hrx_flayer = di + kf;
    ep.hrx_edge = di + 2 * kf;
    ep.hrx_elayer = di + 2 * kf + ke;
    ep.hrx_fw = dw;
    ep.hrx_ew = dw + kf;
    ep.n_hrx_f = nf;
    ep.hrx_fwidth = wf;
    ep.n_hrx_e = ne;
    ep.hrx_ewidth = we;
    ++generation;
    return GEO_OK;
  }

  void launch_h(hipStream_t s) {
    if (plasma) {
      if (magnetized) hipLaunchKernelGGL((geo::h_step_kernel<T, C, true, true>), dim3(h_grid), dim3(geo::kBlock), 0, s, hp);
      else hipLaunchKernelGGL((geo::h_step_kernel<T, C, true, false>), dim3(h_grid), dim3(geo::kBlock), 0, s, hp);
    } else {
      hipLaunchKernelGGL((geo::h_step_kernel<T, C, false, false>), dim3(h_grid), dim3(geo::kBlock), 0, s, hp);
    }
  }
  void launch_e(hipStream_t s) {
    if (plasma) {
      if (magnetized) hipLaunchKernelGGL((geo::e_step_kernel<T, C, true, true>), dim3(e_grid), dim3(geo::kBlock), 0, s, ep);
      else hipLaunchKernelGGL((geo::e_step_kernel<T, C, true, false>), dim3(e_grid), dim3(geo::kBlock), 0, s, ep);
    } else {
      hipLaunchKernelGGL((geo::e_step_kernel<T, C, false, false>), dim3(e_grid), dim3(geo::kBlock), 0, s, ep);
    }
  }

  void launch_rescale(hipStream_t s, int32_t interval) {
    if constexpr (sizeof(T) == 2) {
      if (!dyn_scaling) return;
      const int64_t n_er = V * pn / N, n_et = E * pm / N, n_hr = F * pm / N, n_ht = E * pn / N;
      hipLaunchKernelGGL((scale_max_kernel<T>), dim3(kScaleBlocks), dim3(geo::kBlock), 0, s,
                         hp.er, hp.et, hp.hr, hp.ht, n_er, n_et, n_hr, n_ht, scale_partial);
      hipLaunchKernelGGL(scale_decide_kernel, dim3(1), dim3(geo::kBlock), 0, s, dyn,
                         scale_partial, kScaleBlocks, interval, src_gain_er, src_gain_et);
      hipLaunchKernelGGL((scale_apply_kernel<T>), dim3(kScaleBlocks), dim3(geo::kBlock), 0, s,
                         ep.er, ep.et, hp.hr, hp.ht, n_er, n_et, n_hr, n_ht,
                         (const geo::Dynamic*)dyn);
    }
  }

  int ensure_graph(int32_t chunk) {
    if (graph_exec && graph_chunk == chunk) return GEO_OK;
    drop_graph();
    GEO_HIP(hipStreamBeginCapture(capture_stream, hipStreamCaptureModeThreadLocal));
    launch_rescale(capture_stream, chunk);
    for (int32_t i = 0; i < chunk; ++i) {
      launch_h(capture_stream);
      launch_e(capture_stream);
    }
    hipGraph_t g = nullptr;
    GEO_HIP(hipStreamEndCapture(capture_stream, &g));
    graph = g;
    GEO_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    // upload the executable graph now, so that its first replay does not
    // carry the one-off upload (the replays themselves are unchanged)
    GEO_HIP(hipGraphUpload(graph_exec, capture_stream));
    GEO_HIP(hipStreamSynchronize(capture_stream));
    graph_chunk = chunk;
    return GEO_OK;
  }

  int set_step_index(int64_t index, hipStream_t stream) override {
    if (index < 0) return fail(GEO_ERR_INVALID, "step index must be non-negative");
    hipLaunchKernelGGL(set_step_index_kernel, dim3(1), dim3(1), 0, stream, dyn,
                       (uint32_t)(index & 0xffffffff));
    GEO_HIP(hipGetLastError());
    return GEO_OK;
  }

  int prepare(int32_t chunk) override {
    if (!bound) return fail(GEO_ERR_INVALID, "plan has no bound state");
    if (chunk < 2) return fail(GEO_ERR_INVALID, "a step graph needs a chunk of at least 2 steps");
    return ensure_graph(chunk);
  }

  int step_h(const void* currents, int64_t n, int32_t chunk, int32_t every, void* rt, void* tt,
             hipStream_t stream) override {
    if (!bound) return fail(GEO_ERR_INVALID, "plan has no bound state");
    if (n < 0) return fail(GEO_ERR_INVALID, "step count must be non-negative");
    if (every < 1) return fail(GEO_ERR_INVALID, "sample_every must be positive");
    if ((rt != nullptr && ep.n_hrx_f <= 0) || (tt != nullptr && ep.n_hrx_e <= 0))
      return fail(GEO_ERR_INVALID, "H trace recording requires H receivers");
    hipLaunchKernelGGL(geo::set_dynamic_kernel, dim3(1), dim3(1), 0, stream, dyn, (int64_t)0,
                       (void*)nullptr, (int64_t)0, currents, n);
    const int64_t rows = n / every + 1 + (n % every != 0 ? 1 : 0);
    hipLaunchKernelGGL((geo::set_h_trace_kernel<T, C>), dim3(1), dim3(geo::kBlock), 0, stream, ep,
                       rt, tt, rows, (int64_t)every, n);
    GEO_HIP(hipGetLastError());
    return run_steps(n, chunk, stream);
  }

  int step(const void* currents, int64_t n, int32_t chunk, void* trace,
           hipStream_t stream) override {
    if (!bound) return fail(GEO_ERR_INVALID, "plan has no bound state");
    if (n < 0) return fail(GEO_ERR_INVALID, "step count must be non-negative");
    if (trace != nullptr && n_rx <= 0)
      return fail(GEO_ERR_INVALID, "trace recording requires receivers");
    hipLaunchKernelGGL(geo::set_dynamic_kernel, dim3(1), dim3(1), 0, stream, dyn, (int64_t)0,
                       trace, trace ? n : (int64_t)0, currents, n);
    GEO_HIP(hipGetLastError());
    const int rc = run_steps(n, chunk, stream);
    if (rc != GEO_OK) return rc;
    if (trace != nullptr) {
      C* last = static_cast<C*>(trace) + n * n_rx;
      return sample(last, stream);
    }
    return GEO_OK;
  }

  // n steps after the device block was set for the call: full chunks as
  // graph replays, the rest as single launches.
  int run_steps(int64_t n, int32_t chunk, hipStream_t stream) {
    // rescale points (narrow storage): every `interval` steps from the start
    // of the call, identical with and without graphs for the same chunk
    const int32_t interval = chunk > 1 ? chunk : kRescaleInterval;
    int64_t done = 0;
    if (chunk > 1 && n >= chunk) {
      int rc = ensure_graph(chunk);
      if (rc != GEO_OK) return rc;
      for (; done + chunk <= n; done += chunk) GEO_HIP(hipGraphLaunch(graph_exec, stream));
    }
    for (; done < n; ++done) {
      if (done % interval == 0) launch_rescale(stream, interval);
      launch_h(stream);
      launch_e(stream);
    }
    GEO_HIP(hipGetLastError());
    return GEO_OK;
  }

  int sample(void* row, hipStream_t stream) override {
    if (!bound) return fail(GEO_ERR_INVALID, "plan has no bound state");
    if (n_rx <= 0) return fail(GEO_ERR_INVALID, "no receivers are set");
    const int threads = 64;
    hipLaunchKernelGGL((geo::sample_kernel<T, C>), dim3((n_rx + threads - 1) / threads),
                       dim3(threads), 0, stream, hp.er, hp.pn, (const geo::Dynamic*)dyn,
                       hp.rx_vertex,
                       hp.rx_layer, hp.rx_w, n_rx, static_cast<C*>(row));
    GEO_HIP(hipGetLastError());
    return GEO_OK;
  }

  int diag_one(const T* x, int64_t rows, int32_t cols, int32_t pitch, int32_t field,
               const double* row_w, const double* col_w, const double* eps, int64_t eps_stride,
               const double* sigma, int64_t sigma_stride, double coef, int slot,
               hipStream_t stream) {
    DiagAcc* part = diag_partial + slot * kDiagBlocks;
    hipLaunchKernelGGL((diag_field_kernel<T, C>), dim3(kDiagBlocks), dim3(geo::kBlock), 0,
                       stream, x, rows, cols, pitch, (const geo::Dynamic*)dyn, field, row_w,
                       col_w, eps, eps_stride, sigma, sigma_stride, coef, part);
    hipLaunchKernelGGL(diag_reduce_kernel, dim3(1), dim3(geo::kBlock), 0, stream, part,
                       kDiagBlocks, diag_out + slot);
    GEO_HIP(hipGetLastError());
    return GEO_OK;
  }

  int diagnostics(GeoDiagOut* out, hipStream_t stream) override {
    if (!bound) return fail(GEO_ERR_INVALID, "plan has no bound state");
    int rc;
    if ((rc = diag_one(hp.er, V, (int32_t)(R + 1), (int32_t)pn, geo::kFieldEr, w_er_row,
                       w_er_col, eps_er_dev, eps_er_stride, sig_er_dev, sig_er_stride, 0.5, 0,
                       stream)) != GEO_OK)
      return rc;
    if ((rc = diag_one(hp.et, E, (int32_t)R, (int32_t)pm, geo::kFieldEt, w_et_row, w_et_col,
                       eps_et_dev, eps_et_stride, sig_et_dev, sig_et_stride, 0.5, 1,
                       stream)) != GEO_OK)
      return rc;
    if ((rc = diag_one(hp.hr, F, (int32_t)R, (int32_t)pm, geo::kFieldHr, w_hr_row, w_hr_col,
                       nullptr, 0, nullptr, 0, 0.5 * kMu0, 2, stream)) != GEO_OK)
      return rc;
    if ((rc = diag_one(hp.ht, E, (int32_t)(R + 1), (int32_t)pn, geo::kFieldHt, w_ht_row,
                       w_ht_col, nullptr, 0, nullptr, 0, 0.5 * kMu0, 3, stream)) != GEO_OK)
      return rc;
    {
      DiagAcc* part = diag_partial + 4 * kDiagBlocks;
      hipLaunchKernelGGL((diag_divb_kernel<T, C>), dim3(kDiagBlocks), dim3(geo::kBlock), 0,
                         stream, hp.hr, hp.ht, (int32_t)pn, (int32_t)pm, (int32_t)R, F,
                         si_on ? 1 : 0, (const geo::Dynamic*)dyn, hp.face_edges, dv_face_w,
                         dv_face_omega, dv_rb2, dv_rl, part);
      hipLaunchKernelGGL(diag_reduce_max_kernel, dim3(1), dim3(geo::kBlock), 0, stream, part,
                         kDiagBlocks, diag_out + 4);
      GEO_HIP(hipGetLastError());
    }
    DiagAcc host[5];
    int32_t overflow = 0;
    GEO_HIP(hipMemcpyAsync(host, diag_out, sizeof(host), hipMemcpyDeviceToHost, stream));
    GEO_HIP(hipMemcpyAsync(&overflow, &dyn->overflow, sizeof(overflow), hipMemcpyDeviceToHost,
                           stream));
    GEO_HIP(hipStreamSynchronize(stream));
    out->storage_overflow = overflow;
    out->max_abs_er = host[0].max_abs;
    out->max_abs_et = host[1].max_abs;
    out->max_abs_hr = host[2].max_abs;
    out->max_abs_ht = host[3].max_abs;
    out->energy_er = host[0].energy;
    out->energy_et = host[1].energy;
    out->energy_hr = host[2].energy;
    out->energy_ht = host[3].energy;
    out->nonfinite = host[0].nonfinite + host[1].nonfinite + host[2].nonfinite +
                     host[3].nonfinite;
    out->divb_max_abs = host[4].max_abs;
    out->divb_flux_scale = host[4].energy;
    out->loss_er = host[0].loss;
    out->loss_et = host[1].loss;
    out->sumsq_er = host[0].sumsq;
    out->sumsq_et = host[1].sumsq;
    out->sumsq_hr = host[2].sumsq;
    out->sumsq_ht = host[3].sumsq;
    out->nonfinite_er = host[0].nonfinite;
    out->nonfinite_et = host[1].nonfinite;
    out->nonfinite_hr = host[2].nonfinite;
    out->nonfinite_ht = host[3].nonfinite;
    return GEO_OK;
  }

  int radial_profiles(const GeoProfiles* out, hipStream_t stream) override {
    if (!bound) return fail(GEO_ERR_INVALID, "plan has no bound state");
    struct Field {
      const T* x;
      int64_t rows;
      int32_t cols, pitch, id;
      const double *row_w, *eps, *sigma;
      int64_t eps_stride, sigma_stride;
      double coef;
      double *energy, *loss;
    };
    const Field fields[4] = {
        {hp.er, V, (int32_t)(R + 1), (int32_t)pn, geo::kFieldEr, w_er_row, eps_er_dev,
         sig_er_dev, eps_er_stride, sig_er_stride, 0.5, out->energy_er, out->loss_er},
        {hp.et, E, (int32_t)R, (int32_t)pm, geo::kFieldEt, w_et_row, eps_et_dev, sig_et_dev,
         eps_et_stride, sig_et_stride, 0.5, out->energy_et, out->loss_et},
        {hp.hr, F, (int32_t)R, (int32_t)pm, geo::kFieldHr, w_hr_row, nullptr, nullptr, 0, 0,
         0.5 * kMu0, out->energy_hr, nullptr},
        {hp.ht, E, (int32_t)(R + 1), (int32_t)pn, geo::kFieldHt, w_ht_row, nullptr, nullptr, 0,
         0, 0.5 * kMu0, out->energy_ht, nullptr}};
    const int64_t stride = 2 * (R + 1);
    for (int f = 0; f < 4; ++f) {
      const Field& d = fields[f];
      if (!d.energy && !d.loss) continue;
      const int32_t nb = (int32_t)std::min<int64_t>(kProfileBlocks, d.rows);
      const int64_t per = (d.rows + nb - 1) / nb;
      hipLaunchKernelGGL((diag_profile_kernel<T, C>), dim3(nb), dim3(64), 0, stream, d.x,
                         d.rows, d.cols, d.pitch, (const geo::Dynamic*)dyn, d.id, d.row_w,
                         d.eps, d.eps_stride, d.sigma, d.sigma_stride, per, prof_partial);
      hipLaunchKernelGGL(diag_profile_reduce_kernel, dim3(1), dim3(geo::kBlock), 0, stream,
                         prof_partial, nb, d.cols, prof_out + f * stride);
      GEO_HIP(hipGetLastError());
    }
    std::vector<double> host(4 * stride);
    GEO_HIP(hipMemcpyAsync(host.data(), prof_out, sizeof(double) * host.size(),
                           hipMemcpyDeviceToHost, stream));
    GEO_HIP(hipStreamSynchronize(stream));
    for (int f = 0; f < 4; ++f) {
      const Field& d = fields[f];
      const double* sums = host.data() + f * stride;
      for (int32_t c = 0; c < d.cols; ++c) {
        if (d.energy) d.energy[c] = d.coef * sums[c] * col_w_host[f][c];
        if (d.loss) d.loss[c] = sums[d.cols + c] * col_w_host[f][c];
      }
    }
    return GEO_OK;
  }

  int time_kernels(int64_t n, hipStream_t stream, double* out_ms) override {
    if (!bound) return fail(GEO_ERR_INVALID, "plan has no bound state");
    if (n <= 0) return fail(GEO_ERR_INVALID, "step count must be positive");
    hipLaunchKernelGGL(geo::set_dynamic_kernel, dim3(1), dim3(1), 0, stream, dyn, (int64_t)0,
                       (void*)nullptr, (int64_t)0, (const void*)nullptr, (int64_t)0);
    launch_rescale(stream, (int32_t)std::min<int64_t>(n, 1 << 30));
    std::vector<hipEvent_t> ev(3 * n);
    for (auto& e : ev) GEO_HIP(hipEventCreate(&e));
    for (int64_t i = 0; i < n; ++i) {
      GEO_HIP(hipEventRecord(ev[3 * i], stream));
      launch_h(stream);
      GEO_HIP(hipEventRecord(ev[3 * i + 1], stream));
      launch_e(stream);
      GEO_HIP(hipEventRecord(ev[3 * i + 2], stream));
    }
    GEO_HIP(hipGetLastError());
    GEO_HIP(hipEventSynchronize(ev[3 * n - 1]));
    double h = 0.0, e = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      float a = 0.f, b = 0.f;
      GEO_HIP(hipEventElapsedTime(&a, ev[3 * i], ev[3 * i + 1]));
      GEO_HIP(hipEventElapsedTime(&b, ev[3 * i + 1], ev[3 * i + 2]));
      h += a;
      e += b;
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    out_ms[0] = h / n;
    out_ms[1] = e / n;
    return GEO_OK;
  }

  int get_scales(double* out, hipStream_t stream) override {
    float host[4];
    GEO_HIP(hipMemcpyAsync(host, dyn->scale, sizeof(host), hipMemcpyDeviceToHost, stream));
    GEO_HIP(hipStreamSynchronize(stream));
    for (int f = 0; f < 4; ++f) out[f] = host[f];
    return GEO_OK;
  }

  int set_scales(const double* in, hipStream_t stream) override {
    float host[8];
    for (int f = 0; f < 4; ++f) {
      if (!valid_scale(in[f]))
        return fail(GEO_ERR_INVALID, "storage scales must be powers of two in [2^-60, 2^60]");
      if (sizeof(T) != 2 && in[f] != 1.0)
        return fail(GEO_ERR_INVALID, "only FP16/BF16 storage takes scales other than 1");
      host[f] = (float)in[f];
      host[4 + f] = (float)(1.0 / in[f]);
    }
    // scale[] and inv_scale[] are adjacent in the device block
    GEO_HIP(hipMemcpyAsync(dyn->scale, host, sizeof(host), hipMemcpyHostToDevice, stream));
    GEO_HIP(hipStreamSynchronize(stream));
    return GEO_OK;
  }
};

// Bandwidth probe: 16-byte strips, one per thread.  mode 0 copy, 1 read (sum
// into a per-thread register written only when the impossible sentinel
// matches, so the loads are not removed), 2 write; +4 = non-temporal hints.
typedef float f4v __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(geo::kBlock) stream_probe_kernel(f4v* __restrict__ dst,
                                                                   const f4v* __restrict__ src,
                                                                   int64_t n, int32_t mode) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  float acc = 0.f;
  const bool nt = (mode & 4) != 0;
  const int32_t kind = mode & 3;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    if (kind == 2) {
      const f4v v = {1.f, 2.f, 3.f, (float)i};
      if (nt) __builtin_nontemporal_store(v, dst + i);
      else dst[i] = v;
    } else {
      const f4v v = nt ? __builtin_nontemporal_load(src + i) : src[i];
      if (kind == 0) {
        if (nt) __builtin_nontemporal_store(v, dst + i);
        else dst[i] = v;
      } else {
        acc += v.x + v.y + v.z + v.w;
      }
    }
  }
  if (kind == 1 && acc == -1.2345e-38f) {
    const f4v z = {acc, acc, acc, acc};
    dst[0] = z;
  }
}

// Access-pattern probe (mode 8): the H kernel's edge strips without their
// arithmetic.  dst holds rows of 68 floats (17 strips, the s7 x 64 Ht pitch),
// each updated in place from two gathered rows of src (a third as many rows,
// like Er for Ht): src row = row / 3 + a hashed offset in [-64, 64], clamped.
// Threads map to (row, strip) and blocks to XCD ranges as in the step
// kernels, and the own row carries non-temporal hints.  Compulsory bytes: dst
// read + written, src read once.  Mode 9 (INDEXED) reads the two src rows
// from an int32 (rows, 2) table, as the step kernels read an edge's vertices,
// so that every gather waits on an index load first (+8 bytes per row).
__device__ __forceinline__ void pattern_rows(int64_t row, int64_t rows, int64_t* a, int64_t* b) {
  const int64_t vrows = rows / 3 > 0 ? rows / 3 : 1;
  uint32_t h = (uint32_t)row * 2654435761u;
  *a = std::min<int64_t>(vrows - 1, std::max<int64_t>(0, row / 3 + (int64_t)(h >> 25) - 64));
  h = h * 1597334677u + 12345u;
  *b = std::min<int64_t>(vrows - 1, std::max<int64_t>(0, row / 3 + (int64_t)(h >> 25) - 64));
}

__global__ void pattern_index_kernel(int32_t* __restrict__ idx, int64_t rows) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  int64_t a, b;
  pattern_rows(row, rows, &a, &b);
  idx[2 * row] = (int32_t)a;
  idx[2 * row + 1] = (int32_t)b;
}

template <bool INDEXED>
__global__ void __launch_bounds__(geo::kBlock) pattern_probe_kernel(f4v* __restrict__ dst,
                                                                    const f4v* __restrict__ src,
                                                                    const int32_t* __restrict__ idx,
                                                                    int64_t rows, int32_t per_xcd) {
  constexpr int kStrips = 17, kRowsPerBlock = geo::kBlock / kStrips;
  const int local_row = threadIdx.x / kStrips;
  const int strip = threadIdx.x - local_row * kStrips;
  const int64_t row = (int64_t)geo::logical_block(per_xcd) * kRowsPerBlock + local_row;
  if (local_row >= kRowsPerBlock || row >= rows) return;
  int64_t a, b;
  if constexpr (INDEXED) {
    a = idx[2 * row];
    b = idx[2 * row + 1];
  } else {
    pattern_rows(row, rows, &a, &b);
  }
  const f4v x = src[a * kStrips + strip], y = src[b * kStrips + strip];
  f4v* own = dst + row * kStrips + strip;
  f4v v = __builtin_nontemporal_load(own);
  v += 0.5f * (x - y);
  __builtin_nontemporal_store(v, own);
}

// ---------------------------------------------------------------------------
// Ensembles of independent runs on one mesh (uneven radial grids, own
// coefficients, plasma, ground, sources and receivers), advanced by one H and
// one E launch per step for all runs, full chunks replayed as one HIP graph.
//
// Column-block shape (the tuned parameters):
//   S       strips per thread (1, 2 or 4): a thread updates S 16-byte strips
//           of one row, amortising the row's index and metric loads;
//   policy  bit 0, block order: 0 run-major (each XCD sweeps whole runs, so
//           a run's neighbour columns stay in one L2), 1 runs interleaved
//           block by block (every XCD works on every run); bit 1 (one strip
//           per thread only): every run takes the flat thread mapping, where
//           otherwise each run keeps the mapping its plan took.
// ---------------------------------------------------------------------------
struct EnsembleBase {
  virtual ~EnsembleBase() = default;
  virtual int configure(int32_t S, int32_t policy) = 0;
  virtual int step(const void* currents, int64_t n, int32_t chunk, void* traces,
                   hipStream_t stream) = 0;
  virtual int tune(int32_t steps, hipStream_t stream, int32_t* best_s, int32_t* best_policy,
                   double* best_ms) = 0;
  int32_t n_runs = 0;
  int32_t S = 1, policy = 0;
};

template <typename T, typename C>
struct Ensemble final : EnsembleBase {
  static constexpr int N = geo::Strip<T>::N;
  std::vector<Plan<T, C>*> plans;
  std::vector<uint64_t> seen;
  bool plasma = false, magnetized = false;
  geo::HParams<T, C>* d_hp = nullptr;
  geo::EParams<T, C>* d_ep = nullptr;
  geo::Dynamic** d_dyns = nullptr;
  uint32_t* d_hmap = nullptr;
  uint32_t* d_emap = nullptr;
  int32_t h_blocks = 0, e_blocks = 0, h_per_xcd = 0, e_per_xcd = 0;
  std::vector<geo::HParams<T, C>> host_hp;
  std::vector<geo::EParams<T, C>> host_ep;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int32_t graph_chunk = 0;

  ~Ensemble() override {
    drop_graph();
    for (void* p : {(void*)d_hp, (void*)d_ep, (void*)d_dyns, (void*)d_hmap, (void*)d_emap})
      if (p) (void)hipFree(p);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }

  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_chunk = 0;
  }

  int create(const std::vector<Plan<T, C>*>& ps) {
    plans = ps;
    n_runs = (int32_t)ps.size();
    if (n_runs < 1 || (uint32_t)n_runs > geo::kEnsMaxRuns)
      return fail(GEO_ERR_INVALID, "an ensemble holds 1 to 512 runs");
    plasma = ps[0]->plasma;
    magnetized = ps[0]->magnetized;
    for (Plan<T, C>* p : ps) {
      if (!p->bound) return fail(GEO_ERR_INVALID, "every run must have bound state");
      if (p->plasma != plasma || p->magnetized != magnetized)
        return fail(GEO_ERR_INVALID, "runs must all have Drude, magnetised or no plasma alike");
      if (p->dyn_scaling)
        return fail(GEO_ERR_INVALID, "ensembles take FP32/FP64 or statically scaled storage");
    }
    seen.assign(n_runs, ~0ull);
    std::vector<geo::Dynamic*> dyns(n_runs);
    for (int32_t r = 0; r < n_runs; ++r) dyns[r] = ps[r]->dyn;
    GEO_HIP(hipMalloc(&d_hp, sizeof(geo::HParams<T, C>) * n_runs));
    GEO_HIP(hipMalloc(&d_ep, sizeof(geo::EParams<T, C>) * n_runs));
    GEO_HIP(hipMalloc(&d_dyns, sizeof(geo::Dynamic*) * n_runs));
    GEO_HIP(hipMemcpy(d_dyns, dyns.data(), sizeof(geo::Dynamic*) * n_runs,
                      hipMemcpyHostToDevice));
    GEO_HIP(hipStreamCreateWithFlags(&capture_stream, hipStreamNonBlocking));
    return configure(1, 0);
  }

  // Per-run parameter blocks with the segment geometry of the current S.
  void fill_params() {
    host_hp.resize(n_runs);
    host_ep.resize(n_runs);
    for (int32_t r = 0; r < n_runs; ++r) {
      Plan<T, C>& p = *plans[r];
      host_hp[r] = p.hp;
      host_ep[r] = p.ep;
      host_hp[r].g_edge = make_geom(p.E, p.pn, N, S);
      host_hp[r].g_face = make_geom(p.F, p.pm, N, S);
      host_ep[r].g_vertex = make_geom(p.V, p.pn, N, S);
      host_ep[r].g_edge = make_geom(p.E, p.pm, N, S);
      // one strip per thread: each run keeps the thread mapping its plan took,
      // or every run whose flat indices fit in 32 bits goes flat (policy bit 1)
      const bool all_flat = (policy & 2) != 0;
      if (S == 1 && (p.h_flat || (all_flat && flat_fits(host_hp[r].g_edge, host_hp[r].g_face)))) {
        set_flat(&host_hp[r].g_edge);
        set_flat(&host_hp[r].g_face);
      }
      if (S == 1 && (p.e_flat || (all_flat && flat_fits(host_ep[r].g_vertex, host_ep[r].g_edge)))) {
        set_flat(&host_ep[r].g_vertex);
        set_flat(&host_ep[r].g_edge);
      }
      seen[r] = p.generation;
    }
  }

  int upload_params(hipStream_t stream) {
    fill_params();
    GEO_HIP(hipMemcpyAsync(d_hp, host_hp.data(), sizeof(geo::HParams<T, C>) * n_runs,
                           hipMemcpyHostToDevice, stream));
    GEO_HIP(hipMemcpyAsync(d_ep, host_ep.data(), sizeof(geo::EParams<T, C>) * n_runs,
                           hipMemcpyHostToDevice, stream));
    GEO_HIP(hipStreamSynchronize(stream));  // the host copies are reused
    return GEO_OK;
  }

  static std::vector<uint32_t> combine(const std::vector<std::vector<uint32_t>>& per_run,
                                       int32_t policy) {
    std::vector<uint32_t> out;
    if ((policy & 1) == 0) {
      for (const auto& v : per_run) out.insert(out.end(), v.begin(), v.end());
      return out;
    }
    struct Item { double key; uint32_t run; uint32_t value; };
    std::vector<Item> items;
    for (size_t r = 0; r < per_run.size(); ++r)
      for (size_t i = 0; i < per_run[r].size(); ++i)
        items.push_back({(i + 0.5) / per_run[r].size(), (uint32_t)r, per_run[r][i]});
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
      return a.key < b.key || (a.key == b.key && a.run < b.run);
    });
    out.reserve(items.size());
    for (const Item& it : items) out.push_back(it.value);
    return out;
  }

  static int encode(std::vector<uint32_t>& map, uint32_t run) {
    for (uint32_t& m : map) {
      const uint32_t seg = m >> 24, b = m & 0xffffffu;
      if (b >= geo::kEnsMaxBlocks) return fail(GEO_ERR_INVALID, "run too large for an ensemble");
      m = (seg << (geo::kEnsBlockBits + geo::kEnsRunBits)) | (run << geo::kEnsBlockBits) | b;
    }
    return GEO_OK;
  }

  int configure(int32_t s, int32_t pol) override {
    if (!(s == 1 || s == 2 || s == 4)) return fail(GEO_ERR_INVALID, "strips per thread: 1, 2 or 4");
    if (pol < 0 || pol > 3) return fail(GEO_ERR_INVALID, "block order policy: 0 to 3");
    if ((pol & 2) && s != 1)
      return fail(GEO_ERR_INVALID, "the flat-mapping policy takes one strip per thread");
    S = s;
    policy = pol;
    drop_graph();
    fill_params();
    std::vector<std::vector<uint32_t>> hm(n_runs), em(n_runs);
    int rc;
    for (int32_t r = 0; r < n_runs; ++r) {
      const Plan<T, C>& p = *plans[r];
      std::vector<SegSpec> hs = {{geo::kSegHEdge, geom_blocks(host_hp[r].g_edge)},
                                 {geo::kSegHFace, geom_blocks(host_hp[r].g_face)}};
      hs.push_back({geo::kSegSample, 1});
      const std::vector<SegSpec> es = {{geo::kSegEVertex, geom_blocks(host_ep[r].g_vertex)},
                                       {geo::kSegEEdge, geom_blocks(host_ep[r].g_edge)}};
      if (!blocks_fit(hs) || !blocks_fit(es)) return fail(GEO_ERR_INVALID, "run too large for an ensemble");
      hm[r] = build_block_map(hs);
      em[r] = build_block_map(es);
      if ((rc = encode(hm[r], (uint32_t)r)) != GEO_OK) return rc;
      if ((rc = encode(em[r], (uint32_t)r)) != GEO_OK) return rc;
    }
    const std::vector<uint32_t> hmap = combine(hm, policy), emap = combine(em, policy);
    for (void* p : {(void*)d_hmap, (void*)d_emap})
      if (p) (void)hipFree(p);
    GEO_HIP(hipMalloc(&d_hmap, sizeof(uint32_t) * hmap.size()));
    GEO_HIP(hipMalloc(&d_emap, sizeof(uint32_t) * emap.size()));
    GEO_HIP(hipMemcpy(d_hmap, hmap.data(), sizeof(uint32_t) * hmap.size(), hipMemcpyHostToDevice));
    GEO_HIP(hipMemcpy(d_emap, emap.data(), sizeof(uint32_t) * emap.size(), hipMemcpyHostToDevice));
    h_blocks = (int32_t)hmap.size();
    e_blocks = (int32_t)emap.size();
    h_per_xcd = (h_blocks + geo::kNumXcd - 1) / geo::kNumXcd;
    e_per_xcd = (e_blocks + geo::kNumXcd - 1) / geo::kNumXcd;
    GEO_HIP(hipMemcpy(d_hp, host_hp.data(), sizeof(geo::HParams<T, C>) * n_runs,
                      hipMemcpyHostToDevice));
    GEO_HIP(hipMemcpy(d_ep, host_ep.data(), sizeof(geo::EParams<T, C>) * n_runs,
                      hipMemcpyHostToDevice));
    return GEO_OK;
  }

  template <bool P, bool M, int SS>
  void launch_pair(hipStream_t s) {
    hipLaunchKernelGGL((geo::h_ensemble_kernel<T, C, P, M, SS>), dim3(h_per_xcd * geo::kNumXcd),
                       dim3(geo::kBlock), 0, s, (const geo::HParams<T, C>*)d_hp,
                       (const uint32_t*)d_hmap, h_blocks, h_per_xcd);
    hipLaunchKernelGGL((geo::e_ensemble_kernel<T, C, P, M, SS>), dim3(e_per_xcd * geo::kNumXcd),
                       dim3(geo::kBlock), 0, s, (const geo::EParams<T, C>*)d_ep,
                       (const uint32_t*)d_emap, e_blocks, e_per_xcd);
  }

  template <int SS>
  void launch_variant(hipStream_t s) {
    if (plasma) {
      if (magnetized) launch_pair<true, true, SS>(s);
      else launch_pair<true, false, SS>(s);
    } else {
      launch_pair<false, false, SS>(s);
    }
  }

  void launch_step(hipStream_t s) {
    if (S == 1) launch_variant<1>(s);
    else if (S == 2) launch_variant<2>(s);
    else launch_variant<4>(s);
  }

  int ensure_graph(int32_t chunk) {
    if (graph_exec && graph_chunk == chunk) return GEO_OK;
    drop_graph();
    GEO_HIP(hipStreamBeginCapture(capture_stream, hipStreamCaptureModeThreadLocal));
    for (int32_t i = 0; i < chunk; ++i) launch_step(capture_stream);
    hipGraph_t g = nullptr;
    GEO_HIP(hipStreamEndCapture(capture_stream, &g));
    graph = g;
    GEO_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    // upload the executable graph now, so that its first replay does not
    // carry the one-off upload (the replays themselves are unchanged)
    GEO_HIP(hipGraphUpload(graph_exec, capture_stream));
    GEO_HIP(hipStreamSynchronize(capture_stream));
    graph_chunk = chunk;
    return GEO_OK;
  }

  int refresh(hipStream_t stream) {
    for (int32_t r = 0; r < n_runs; ++r)
      if (!plans[r]->bound) return fail(GEO_ERR_INVALID, "a run has no bound state");
    bool stale = false;
    for (int32_t r = 0; r < n_runs; ++r) stale = stale || seen[r] != plans[r]->generation;
    return stale ? upload_params(stream) : GEO_OK;
  }

  // currents: (n_runs, n) compute type or NULL; traces: (n_runs, n + 1, n_rx).
  int step(const void* currents, int64_t n, int32_t chunk, void* traces,
           hipStream_t stream) override {
    if (n < 0) return fail(GEO_ERR_INVALID, "step count must be non-negative");
    int rc = refresh(stream);
    if (rc != GEO_OK) return rc;
    int32_t n_rx = plans[0]->n_rx;
    if (traces != nullptr) {
      for (Plan<T, C>* p : plans)
        if (p->n_rx != n_rx || n_rx <= 0)
          return fail(GEO_ERR_INVALID, "trace recording needs the same receiver count in every run");
    }
    const int threads = 256;
    hipLaunchKernelGGL(geo::set_dynamic_runs_kernel, dim3((n_runs + threads - 1) / threads),
                       dim3(threads), 0, stream, (geo::Dynamic* const*)d_dyns, n_runs,
                       static_cast<const char*>(currents), (int64_t)(n * sizeof(C)),
                       static_cast<char*>(traces), (int64_t)((n + 1) * n_rx * sizeof(C)),
                       traces ? n + 1 : (int64_t)0, n);
    GEO_HIP(hipGetLastError());
    int64_t done = 0;
    if (chunk > 1 && n >= chunk) {
      if ((rc = ensure_graph(chunk)) != GEO_OK) return rc;
      for (; done + chunk <= n; done += chunk) GEO_HIP(hipGraphLaunch(graph_exec, stream));
    }
    for (; done < n; ++done) launch_step(stream);
    if (traces != nullptr)
      hipLaunchKernelGGL((geo::sample_runs_kernel<T, C>), dim3(n_runs), dim3(geo::kBlock), 0,
                         stream, (const geo::HParams<T, C>*)d_hp);
    GEO_HIP(hipGetLastError());
    return GEO_OK;
  }

  // Device bytes of one run's state (fields, surface-impedance memory, plasma).
  struct StateSpan { void* ptr; size_t bytes; };
  std::vector<StateSpan> state_spans() const {
    std::vector<StateSpan> v;
    for (const Plan<T, C>* p : plans) {
      v.push_back({(void*)p->hp.er, sizeof(T) * p->V * p->pn});
      v.push_back({(void*)p->hp.ht, sizeof(T) * p->E * p->pn});
      v.push_back({(void*)p->hp.et, sizeof(T) * p->E * p->pm});
      v.push_back({(void*)p->hp.hr, sizeof(T) * p->F * p->pm});
      if (p->si_on) v.push_back({(void*)p->hp.si_mem, sizeof(C) * p->hp.si_pitch * p->E});
      // the rounding step index advances with the state
      v.push_back({(void*)&p->dyn->sr_h, 2 * sizeof(uint32_t)});
      if (p->plasma)
        v.push_back({(void*)p->hp.pl.current, sizeof(C) * p->hp.pl.n_species *
                                                  p->layout.plasma_components * p->F *
                                                  p->layout.plasma_pitch});
    }
    return v;
  }

  // Time every (S, policy) candidate over `steps` source-free steps and keep
  // the fastest (the flat-mapping policies only with S = 1); the state of
  // every run is restored afterwards.
  int tune(int32_t steps, hipStream_t stream, int32_t* best_s, int32_t* best_policy,
           double* best_ms) override {
    if (steps < 1) return fail(GEO_ERR_INVALID, "tuning needs at least one step");
    int rc = refresh(stream);
    if (rc != GEO_OK) return rc;
    const std::vector<StateSpan> spans = state_spans();
    size_t total = 0;
    for (const StateSpan& sp : spans) total += round_up((int64_t)sp.bytes, 256);
    char* save = nullptr;
    GEO_HIP(hipMalloc(&save, total));
    size_t off = 0;
    for (const StateSpan& sp : spans) {
      GEO_HIP(hipMemcpyAsync(save + off, sp.ptr, sp.bytes, hipMemcpyDeviceToDevice, stream));
      off += round_up((int64_t)sp.bytes, 256);
    }
    hipEvent_t a, b;
    GEO_HIP(hipEventCreate(&a));
    GEO_HIP(hipEventCreate(&b));
    double best = 1e300;
    int32_t bs = 1, bp = 0;
    for (int32_t pol : {0, 1, 2, 3}) {
      for (int32_t s : {1, 2, 4}) {
        if ((pol & 2) && s != 1) continue;
        if ((rc = configure(s, pol)) != GEO_OK) break;
        if ((rc = step(nullptr, 2, 1, nullptr, stream)) != GEO_OK) break;  // warm-up
        GEO_HIP(hipEventRecord(a, stream));
        if ((rc = step(nullptr, steps, 1, nullptr, stream)) != GEO_OK) break;
        GEO_HIP(hipEventRecord(b, stream));
        GEO_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        GEO_HIP(hipEventElapsedTime(&ms, a, b));
        if (ms / steps < best) {
          best = ms / steps;
          bs = s;
          bp = pol;
        }
      }
      if (rc != GEO_OK) break;
    }
    off = 0;
    for (const StateSpan& sp : spans) {
      (void)hipMemcpyAsync(sp.ptr, save + off, sp.bytes, hipMemcpyDeviceToDevice, stream);
      off += round_up((int64_t)sp.bytes, 256);
    }
    (void)hipStreamSynchronize(stream);
    (void)hipFree(save);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != GEO_OK) return rc;
    if ((rc = configure(bs, bp)) != GEO_OK) return rc;
    *best_s = bs;
    *best_policy = bp;
    *best_ms = best;
    return GEO_OK;
  }
};

template <typename T, typename C>
int make_ensemble(PlanBase* const* bases, int32_t n, std::unique_ptr<EnsembleBase>* out) {
  std::vector<Plan<T, C>*> ps(n);
  for (int32_t i = 0; i < n; ++i) {
    ps[i] = dynamic_cast<Plan<T, C>*>(bases[i]);
    if (!ps[i]) return fail(GEO_ERR_INVALID, "runs of an ensemble must share the storage type");
  }
  auto e = std::make_unique<Ensemble<T, C>>();
  const int rc = e->create(ps);
  if (rc != GEO_OK) return rc;
  *out = std::move(e);
  return GEO_OK;
}

int check_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return fail(GEO_ERR_UNAVAILABLE, "no HIP device is available");
  int dev = 0;
  GEO_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  GEO_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(GEO_ERR_UNAVAILABLE,
                std::string("this library is built for gfx950; device is ") + prop.gcnArchName);
  return GEO_OK;
}

}  // namespace

int geo_internal_fail(int code, const char* msg) { return fail(code, msg); }

// ---------------------------------------------------------------------------
// Surface operators of the backend interface (numpy_backend.py:90-114): one
// thread per output value, the reference's order of additions.
// ---------------------------------------------------------------------------
template <typename V>
__global__ void mesh_op_kernel(int32_t op, const V* __restrict__ in, int64_t out_rows,
                               int64_t cols, const int32_t* __restrict__ idx,
                               const int32_t* __restrict__ sgn, int32_t width,
                               V* __restrict__ out) {
  const int64_t total = out_rows * cols;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols;
    const int64_t c = i - r * cols;
    V acc;
    if (op == GEO_OP_EDGE_DIFFERENCE || op == GEO_OP_DUAL_EDGE_DIFFERENCE) {
      // edges (E,2) = tail, head; edge faces (E,2) = left, right
      const V a = in[(int64_t)idx[2 * r] * cols + c], b = in[(int64_t)idx[2 * r + 1] * cols + c];
      acc = op == GEO_OP_EDGE_DIFFERENCE ? b - a : a - b;
    } else {
      // face_edges (F,3) with signs, or padded vertex incidence (V,width)
      acc = in[(int64_t)idx[width * r] * cols + c] * (V)sgn[width * r];
      for (int32_t s = 1; s < width; ++s)
        acc += in[(int64_t)idx[width * r + s] * cols + c] * (V)sgn[width * r + s];
    }
    out[i] = acc;
  }
}

template <typename V>
__global__ void array_diff_kernel(const V* __restrict__ in, int64_t rows, int64_t cols,
                                  int32_t axis, V* __restrict__ out) {
  const int64_t orows = axis == 0 ? rows - 1 : rows;
  const int64_t ocols = axis == 0 ? cols : cols - 1;
  const int64_t total = orows * ocols;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / ocols;
    const int64_t c = i - r * ocols;
    const int64_t a = r * cols + c;
    out[i] = in[axis == 0 ? a + cols : a + 1] - in[a];
  }
}

struct GeoMeshOpsImpl {
  int64_t V = 0, E = 0, F = 0;
  int32_t width = 0;
  DeviceArena arena;
  int32_t *edges = nullptr, *edge_faces = nullptr, *face_edges = nullptr, *face_signs = nullptr;
  int32_t *vertex_edges = nullptr, *vertex_signs = nullptr;
};

struct GeoMeshOps {
  GeoMeshOpsImpl impl;
};

struct GeoPlan {
  std::unique_ptr<PlanBase> impl;
};

struct GeoEnsemble {
  std::unique_ptr<EnsembleBase> impl;
};

extern "C" {

int geo_version(void) { return 1; }

int geo_mesh_ops_create(const GeoMeshDesc* m, GeoMeshOps** out) {
  if (!m || !out) return fail(GEO_ERR_INVALID, "null argument");
  *out = nullptr;
  int rc = check_device();
  if (rc != GEO_OK) return rc;
  if (!m->edges || !m->face_edges || !m->face_edge_signs || !m->edge_left_faces ||
      !m->edge_right_faces || !m->vertex_edges || !m->vertex_edge_signs)
    return fail(GEO_ERR_INVALID, "mesh topology arrays are required");
  const int64_t V = m->n_vertices, E = m->n_edges, F = m->n_faces, W = m->max_degree;
  if (V <= 0 || E <= 0 || F <= 0 || W < 1 || W > 16 || V > INT32_MAX || E > INT32_MAX ||
      F > INT32_MAX)
    return fail(GEO_ERR_INVALID, "mesh sizes are out of range");
  std::vector<int32_t> edges(2 * E), faces2(2 * E), fe(3 * F), fs(3 * F), ve(W * V), vs(W * V);
  for (int64_t e = 0; e < E; ++e) {
    const int64_t t = m->edges[2 * e], h = m->edges[2 * e + 1];
    const int64_t l = m->edge_left_faces[e], r = m->edge_right_faces[e];
    if (t < 0 || t >= V || h < 0 || h >= V || l < 0 || l >= F || r < 0 || r >= F)
      return fail(GEO_ERR_INVALID, "edge incidence index out of range");
    edges[2 * e] = (int32_t)t;
    edges[2 * e + 1] = (int32_t)h;
    faces2[2 * e] = (int32_t)l;
    faces2[2 * e + 1] = (int32_t)r;
  }
  for (int64_t i = 0; i < 3 * F; ++i) {
    if (m->face_edges[i] < 0 || m->face_edges[i] >= E)
      return fail(GEO_ERR_INVALID, "face edge index out of range");
    fe[i] = (int32_t)m->face_edges[i];
    fs[i] = (int32_t)m->face_edge_signs[i];
  }
  for (int64_t i = 0; i < W * V; ++i) {
    if (m->vertex_edges[i] < 0 || m->vertex_edges[i] >= E)
      return fail(GEO_ERR_INVALID, "vertex edge index out of range");
    ve[i] = (int32_t)m->vertex_edges[i];
    vs[i] = (int32_t)m->vertex_edge_signs[i];
  }
  auto ops = std::make_unique<GeoMeshOps>();
  GeoMeshOpsImpl& o = ops->impl;
  o.V = V;
  o.E = E;
  o.F = F;
  o.width = (int32_t)W;
  if ((rc = o.arena.upload(edges, &o.edges)) != GEO_OK) return rc;
  if ((rc = o.arena.upload(faces2, &o.edge_faces)) != GEO_OK) return rc;
  if ((rc = o.arena.upload(fe, &o.face_edges)) != GEO_OK) return rc;
  if ((rc = o.arena.upload(fs, &o.face_signs)) != GEO_OK) return rc;
  if ((rc = o.arena.upload(ve, &o.vertex_edges)) != GEO_OK) return rc;
  if ((rc = o.arena.upload(vs, &o.vertex_signs)) != GEO_OK) return rc;
  *out = ops.release();
  return GEO_OK;
}

int geo_mesh_op(GeoMeshOps* ops, int32_t op, int32_t value_bytes, const void* in, int64_t cols,
                void* out, void* stream) {
  if (!ops || !in || !out) return fail(GEO_ERR_INVALID, "null argument");
  if (value_bytes != 4 && value_bytes != 8) return fail(GEO_ERR_INVALID, "values must be FP32 or FP64");
  if (cols < 1) return fail(GEO_ERR_INVALID, "cols must be positive");
  const GeoMeshOpsImpl& o = ops->impl;
  int64_t rows;
  const int32_t *idx, *sgn = nullptr;
  int32_t width = 0;
  switch (op) {
    case GEO_OP_EDGE_DIFFERENCE: rows = o.E; idx = o.edges; break;
    case GEO_OP_DUAL_EDGE_DIFFERENCE: rows = o.E; idx = o.edge_faces; break;
    case GEO_OP_FACE_CIRCULATION: rows = o.F; idx = o.face_edges; sgn = o.face_signs; width = 3; break;
    case GEO_OP_DUAL_CELL_CIRCULATION:
      rows = o.V; idx = o.vertex_edges; sgn = o.vertex_signs; width = o.width; break;
    default: return fail(GEO_ERR_INVALID, "unknown mesh operator");
  }
  const int64_t total = rows * cols;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (value_bytes == 8)
    hipLaunchKernelGGL(mesh_op_kernel<double>, dim3(blocks), dim3(256), 0, s, op,
                       static_cast<const double*>(in), rows, cols, idx, sgn, width,
                       static_cast<double*>(out));
  else
    hipLaunchKernelGGL(mesh_op_kernel<float>, dim3(blocks), dim3(256), 0, s, op,
                       static_cast<const float*>(in), rows, cols, idx, sgn, width,
                       static_cast<float*>(out));
  GEO_HIP(hipGetLastError());
  return GEO_OK;
}

int geo_array_diff(int32_t value_bytes, const void* in, int64_t rows, int64_t cols, int32_t axis,
                   void* out, void* stream) {
  if (!in || !out) return fail(GEO_ERR_INVALID, "null argument");
  if (value_bytes != 4 && value_bytes != 8) return fail(GEO_ERR_INVALID, "values must be FP32 or FP64");
  if (axis != 0 && axis != 1) return fail(GEO_ERR_INVALID, "axis must be 0 or 1");
  if (rows < 1 || cols < 1 || (axis == 0 ? rows : cols) < 2)
    return fail(GEO_ERR_INVALID, "array is too small to difference");
  const int64_t total = (axis == 0 ? rows - 1 : rows) * (axis == 0 ? cols : cols - 1);
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (value_bytes == 8)
    hipLaunchKernelGGL(array_diff_kernel<double>, dim3(blocks), dim3(256), 0, s,
                       static_cast<const double*>(in), rows, cols, axis, static_cast<double*>(out));
  else
    hipLaunchKernelGGL(array_diff_kernel<float>, dim3(blocks), dim3(256), 0, s,
                       static_cast<const float*>(in), rows, cols, axis, static_cast<float*>(out));
  GEO_HIP(hipGetLastError());
  return GEO_OK;
}

void geo_mesh_ops_destroy(GeoMeshOps* ops) { delete ops; }

const char* geo_last_error(void) { return g_error.c_str(); }

int geo_plan_create(const GeoMeshDesc* mesh, const GeoRadialDesc* radial,
                    const GeoPhysicsDesc* physics, GeoPlan** out) {
  if (!mesh || !radial || !physics || !out) return fail(GEO_ERR_INVALID, "null argument");
  *out = nullptr;
  int rc = check_device();
  if (rc != GEO_OK) return rc;
  std::unique_ptr<PlanBase> impl;
  switch (physics->storage) {
    case GEO_STORAGE_F32: {
      auto p = std::make_unique<Plan<float, float>>();
      rc = p->create(mesh, radial, physics);
      impl = std::move(p);
      break;
    }
    case GEO_STORAGE_F64: {
      auto p = std::make_unique<Plan<double, double>>();
      rc = p->create(mesh, radial, physics);
      impl = std::move(p);
      break;
    }
    case GEO_STORAGE_F16: {
      auto p = std::make_unique<Plan<__half, float>>();
      rc = p->create(mesh, radial, physics);
      impl = std::move(p);
      break;
    }
    case GEO_STORAGE_BF16: {
      auto p = std::make_unique<Plan<__hip_bfloat16, float>>();
      rc = p->create(mesh, radial, physics);
      impl = std::move(p);
      break;
    }
    default:
      return fail(GEO_ERR_INVALID, "unknown storage type");
  }
  if (rc != GEO_OK) return rc;
  *out = new GeoPlan{std::move(impl)};
  return GEO_OK;
}

int geo_plan_layout(const GeoPlan* plan, GeoLayout* out) {
  if (!plan || !out) return fail(GEO_ERR_INVALID, "null argument");
  *out = plan->impl->layout;
  return GEO_OK;
}

int geo_plan_bind(GeoPlan* plan, const GeoState* state) {
  if (!plan || !state) return fail(GEO_ERR_INVALID, "null argument");
  return plan->impl->bind(state);
}

int geo_step(GeoPlan* plan, const void* currents, int64_t n_steps, int32_t chunk, void* trace,
             void* stream) {
  if (!plan) return fail(GEO_ERR_INVALID, "null plan");
  return plan->impl->step(currents, n_steps, chunk, trace, static_cast<hipStream_t>(stream));
}

int geo_plan_set_step_index(GeoPlan* plan, int64_t step_index, void* stream) {
  if (!plan) return fail(GEO_ERR_INVALID, "null plan");
  return plan->impl->set_step_index(step_index, static_cast<hipStream_t>(stream));
}

int geo_step_prepare(GeoPlan* plan, int32_t chunk) {
  if (!plan) return fail(GEO_ERR_INVALID, "null plan");
  return plan->impl->prepare(chunk);
}

int geo_set_receivers(GeoPlan* plan, const int64_t* vertices, const int64_t* layers,
                      const double* weights, int32_t n_receivers) {
  if (!plan) return fail(GEO_ERR_INVALID, "null plan");
  if (n_receivers > 0 && (!vertices || !layers || !weights))
    return fail(GEO_ERR_INVALID, "null receiver tables");
  return plan->impl->set_receivers(vertices, layers, weights, n_receivers);
}

int geo_set_h_receivers(GeoPlan* plan, const int64_t* faces, const int64_t* face_layers,
                        const double* face_weights, int32_t n_face_receivers, int32_t face_width,
                        const int64_t* edges, const int64_t* edge_layers,
                        const double* edge_weights, int32_t n_edge_receivers, int32_t edge_width) {
  if (!plan) return fail(GEO_ERR_INVALID, "null plan");
  return plan->impl->set_h_receivers(faces, face_layers, face_weights, n_face_receivers,
                                     face_width, edges, edge_layers, edge_weights,
                                     n_edge_receivers, edge_width);
}

int geo_step_h(GeoPlan* plan, const void* currents, int64_t n_steps, int32_t chunk,
               int32_t sample_every, void* hr_trace, void* ht_trace, void* stream) {
  if (!plan) return fail(GEO_ERR_INVALID, "null plan");
  return plan->impl->step_h(currents, n_steps, chunk, sample_every, hr_trace, ht_trace,
                            static_cast<hipStream_t>(stream));
}

int geo_sample_er(GeoPlan* plan, void* row, void* stream) {
  if (!plan || !row) return fail(GEO_ERR_INVALID, "null argument");
  return plan->impl->sample(row, static_cast<hipStream_t>(stream));
}

int geo_diagnostics(GeoPlan* plan, GeoDiagOut* out, void* stream) {
  if (!plan || !out) return fail(GEO_ERR_INVALID, "null argument");
  return plan->impl->diagnostics(out, static_cast<hipStream_t>(stream));
}

int geo_radial_profiles(GeoPlan* plan, const GeoProfiles* out, void* stream) {
  if (!plan || !out) return fail(GEO_ERR_INVALID, "null argument");
  return plan->impl->radial_profiles(out, static_cast<hipStream_t>(stream));
}

int geo_time_kernels(GeoPlan* plan, int64_t n_steps, void* stream, double* out_ms) {
  if (!plan || !out_ms) return fail(GEO_ERR_INVALID, "null argument");
  return plan->impl->time_kernels(n_steps, static_cast<hipStream_t>(stream), out_ms);
}

int geo_get_storage_scales(GeoPlan* plan, double* out4, void* stream) {
  if (!plan || !out4) return fail(GEO_ERR_INVALID, "null argument");
  return plan->impl->get_scales(out4, static_cast<hipStream_t>(stream));
}

int geo_set_storage_scales(GeoPlan* plan, const double* in4, void* stream) {
  if (!plan || !in4) return fail(GEO_ERR_INVALID, "null argument");
  return plan->impl->set_scales(in4, static_cast<hipStream_t>(stream));
}

void geo_plan_destroy(GeoPlan* plan) { delete plan; }

int geo_ensemble_create(GeoPlan* const* plans, int32_t n_plans, GeoEnsemble** out) {
  if (!plans || !out || n_plans < 1) return fail(GEO_ERR_INVALID, "null argument");
  *out = nullptr;
  std::vector<PlanBase*> bases(n_plans);
  for (int32_t i = 0; i < n_plans; ++i) {
    if (!plans[i]) return fail(GEO_ERR_INVALID, "null plan");
    bases[i] = plans[i]->impl.get();
  }
  // A plan listed twice would have its state advanced twice per step by
  // blocks of one launch (a race on its fields) and its step counter doubled.
  std::vector<PlanBase*> sorted(bases);
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    return fail(GEO_ERR_INVALID, "an ensemble may hold each plan only once");
  std::unique_ptr<EnsembleBase> impl;
  int rc = GEO_ERR_INVALID;
  if (dynamic_cast<Plan<float, float>*>(bases[0]))
    rc = make_ensemble<float, float>(bases.data(), n_plans, &impl);
  else if (dynamic_cast<Plan<double, double>*>(bases[0]))
    rc = make_ensemble<double, double>(bases.data(), n_plans, &impl);
  else if (dynamic_cast<Plan<__half, float>*>(bases[0]))
    rc = make_ensemble<__half, float>(bases.data(), n_plans, &impl);
  else if (dynamic_cast<Plan<__hip_bfloat16, float>*>(bases[0]))
    rc = make_ensemble<__hip_bfloat16, float>(bases.data(), n_plans, &impl);
  if (rc != GEO_OK) return rc;
  *out = new GeoEnsemble{std::move(impl)};
  return GEO_OK;
}

int geo_ensemble_configure(GeoEnsemble* ens, int32_t strips_per_thread, int32_t block_order) {
  if (!ens) return fail(GEO_ERR_INVALID, "null argument");
  return ens->impl->configure(strips_per_thread, block_order);
}

int geo_ensemble_tune(GeoEnsemble* ens, int32_t steps, void* stream, int32_t* strips_per_thread,
                      int32_t* block_order, double* ms_per_step) {
  if (!ens || !strips_per_thread || !block_order || !ms_per_step)
    return fail(GEO_ERR_INVALID, "null argument");
  return ens->impl->tune(steps, static_cast<hipStream_t>(stream), strips_per_thread, block_order,
                         ms_per_step);
}

int geo_ensemble_step(GeoEnsemble* ens, const void* currents, int64_t n_steps, int32_t chunk,
                      void* traces, void* stream) {
  if (!ens) return fail(GEO_ERR_INVALID, "null argument");
  return ens->impl->step(currents, n_steps, chunk, traces, static_cast<hipStream_t>(stream));
}

void geo_ensemble_destroy(GeoEnsemble* ens) { delete ens; }

int geo_stream_probe(void* dst, const void* src, int64_t bytes, int32_t mode, int32_t repeats,
                     void* stream, double* out_ms) {
  if (!dst || bytes <= 0 || (bytes & 15) || repeats < 1 || !out_ms)
    return fail(GEO_ERR_INVALID, "stream probe needs a 16-byte multiple and a destination");
  if ((mode & 3) != 2 && !src) return fail(GEO_ERR_INVALID, "stream probe needs a source");
  if (mode != 8 && mode != 9 && (mode & ~7)) return fail(GEO_ERR_INVALID, "unknown stream probe mode");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (mode == 8 || mode == 9) {  // edge-strip access pattern over rows of 272 bytes
    const int64_t rows = bytes / 272;
    if (rows < 3) return fail(GEO_ERR_INVALID, "pattern probe needs at least 3 rows");
    if (rows >= ((int64_t)1 << 31)) return fail(GEO_ERR_INVALID, "pattern probe rows exceed int32");
    const int64_t blocks = (rows + geo::kBlock / 17 - 1) / (geo::kBlock / 17);
    const int32_t per_xcd = (int32_t)((blocks + geo::kNumXcd - 1) / geo::kNumXcd);
    const int64_t grid = (int64_t)per_xcd * geo::kNumXcd;
    int32_t* idx = nullptr;
    if (mode == 9) {
      GEO_HIP(hipMalloc(&idx, (size_t)rows * 2 * sizeof(int32_t)));
      hipLaunchKernelGGL(pattern_index_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s,
                         idx, rows);
    }
    auto launch = [&]() {
      if (mode == 9)
        hipLaunchKernelGGL(pattern_probe_kernel<true>, dim3((unsigned)grid), dim3(geo::kBlock), 0, s,
                           static_cast<f4v*>(dst), static_cast<const f4v*>(src), idx, rows, per_xcd);
      else
        hipLaunchKernelGGL(pattern_probe_kernel<false>, dim3((unsigned)grid), dim3(geo::kBlock), 0, s,
                           static_cast<f4v*>(dst), static_cast<const f4v*>(src), idx, rows, per_xcd);
    };
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    float ms = 0.f;
    if (e == hipSuccess) {
      launch();
      e = hipEventRecord(a, s);
      for (int r = 0; r < repeats && e == hipSuccess; ++r) launch();
      if (e == hipSuccess) e = hipEventRecord(b, s);
      if (e == hipSuccess) e = hipEventSynchronize(b);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    }
    if (a != nullptr) (void)hipEventDestroy(a);
    if (b != nullptr) (void)hipEventDestroy(b);
    if (idx != nullptr) {
      const hipError_t sync = hipStreamSynchronize(s);  // the probe launches may still read idx
      if (e == hipSuccess) e = sync;
      (void)hipFree(idx);
    }
    if (e != hipSuccess)
      return fail(GEO_ERR_HIP, std::string("pattern probe failed: ") + hipGetErrorString(e));
    *out_ms = ms / repeats;
    return GEO_OK;
  }
  const int64_t n = bytes / 16;
  // one strip per thread over the whole range, as the step kernels do (a
  // resident-sized grid-stride loop measured 5.07 against 6.65 TB/s copy)
  const int grid = (int)std::min<int64_t>((n + geo::kBlock - 1) / geo::kBlock, (int64_t)1 << 30);
  hipEvent_t a, b;
  GEO_HIP(hipEventCreate(&a));
  GEO_HIP(hipEventCreate(&b));
  hipLaunchKernelGGL(stream_probe_kernel, dim3(grid), dim3(geo::kBlock), 0, s,
                     static_cast<f4v*>(dst), static_cast<const f4v*>(src), n, mode);
  GEO_HIP(hipEventRecord(a, s));
  for (int r = 0; r < repeats; ++r)
    hipLaunchKernelGGL(stream_probe_kernel, dim3(grid), dim3(geo::kBlock), 0, s,
                       static_cast<f4v*>(dst), static_cast<const f4v*>(src), n, mode);
  GEO_HIP(hipEventRecord(b, s));
  GEO_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  GEO_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *out_ms = ms / repeats;
  return GEO_OK;
}

}  // extern "C"
