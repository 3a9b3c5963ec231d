// Data-movement and reduction kernels: strided permute (the label-driven
// np.rollaxis + reshape copies of tensor.py), dtype conversion, scaling,
// FP64 sum of squares (np.linalg.norm) and the strided trace (np.trace).
// All are HBM-bound; each thread handles consecutive output elements so
// stores are coalesced.
#include "tn_common.h"

#include <mutex>
#include <vector>

namespace tn {

static thread_local std::string g_err;
static thread_local int g_code = TN_OK;

void set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  g_code = code;
}

int last_code() { return g_code; }

// ---------------------------------------------------------------- permute
// Tiled path: output dims (slow..fast).  When the source's fastest-varying
// output mode is not the last output mode, a 64x64 LDS tile transposes
// between the output's fastest mode (o) and the source's unit-stride mode (s),
// so both the gather and the store are coalesced.
template <typename T>
__global__ void __launch_bounds__(256) permute_generic(const T* __restrict__ src, Modes m,
                                                       T* __restrict__ dst, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (n <= 0xffffffffLL) {
    for (; i < n; i += stride) dst[i] = src[mode_offset32(m, (uint32_t)i)];
  } else {
    for (; i < n; i += stride) dst[i] = src[mode_offset(m, i)];
  }
}

// outer: all modes except `s` (source unit-stride mode) and `o` (last output
// mode); the tile covers a 64 (along s) x 64 (along o) block.
template <typename T>
__global__ void __launch_bounds__(256) permute_tiled(const T* __restrict__ src, Modes outer,
                                                     Modes outer_dst, int64_t ds, int64_t dsrc_s,
                                                     int64_t dst_s, int64_t do_, int64_t src_o,
                                                     T* __restrict__ dst) {
  __shared__ T tile[64][65];
  int64_t bs = (int64_t)blockIdx.x * 64;  // along s
  int64_t bo = (int64_t)blockIdx.y * 64;  // along o
  int64_t ob = blockIdx.z;
  int64_t so = mode_offset(outer, ob);
  int64_t doff = mode_offset(outer_dst, ob);
  int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
  // read: consecutive lanes along s (unit stride in src)
  for (int r = ty; r < 64; r += 4) {
    int64_t s = bs + tx, o = bo + r;
    if (s < ds && o < do_) tile[r][tx] = src[so + s * dsrc_s + o * src_o];
  }
  __syncthreads();
  // write: consecutive lanes along o (unit stride in dst)
  for (int r = ty; r < 64; r += 4) {
    int64_t s = bs + r, o = bo + tx;
    if (s < ds && o < do_) dst[doff + s * dst_s + o] = tile[tx][r];
  }
}

template <typename T>
static int permute_impl(const T* src, const Modes& m, T* dst, hipStream_t st) {
  int64_t n = modes_size(m);
  if (n == 0) return TN_OK;
  // find the source unit-stride mode
  int last = m.n - 1;
  int s_mode = -1;
  for (int i = 0; i < m.n; ++i)
    if (m.strides[i] == 1 && m.dims[i] > 1) s_mode = i;
  bool tiled = m.n >= 2 && s_mode >= 0 && s_mode != last && m.dims[last] >= 16 &&
               m.dims[s_mode] >= 16 && m.strides[last] != 1;
  if (tiled) {
    // dense destination strides
    int64_t dstr[TN_MAX_MODES];
    int64_t acc = 1;
    for (int i = m.n - 1; i >= 0; --i) { dstr[i] = acc; acc *= m.dims[i]; }
    Modes outer{}, outer_dst{};
    outer.n = 0;
    outer_dst.n = 0;
    for (int i = 0; i < m.n; ++i) {
      if (i == s_mode || i == last) continue;
      outer.dims[outer.n] = m.dims[i];
      outer.strides[outer.n] = m.strides[i];
      outer_dst.dims[outer_dst.n] = m.dims[i];
      outer_dst.strides[outer_dst.n] = dstr[i];
      outer.n++;
      outer_dst.n++;
    }
    int64_t nouter = modes_size(outer);
    if (nouter <= 65535) {
      dim3 grid((unsigned)ceil_div(m.dims[s_mode], 64), (unsigned)ceil_div(m.dims[last], 64),
                (unsigned)nouter);
      if (grid.y <= 65535) {
        hipLaunchKernelGGL(permute_tiled<T>, grid, dim3(256), 0, st, src, outer, outer_dst,
                           m.dims[s_mode], (int64_t)1, dstr[s_mode], m.dims[last],
                           m.strides[last], dst);
        TN_CHECK_LAUNCH();
        return TN_OK;
      }
    }
  }
  int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 65536);
  hipLaunchKernelGGL(permute_generic<T>, dim3((unsigned)blocks), dim3(256), 0, st, src, m, dst, n);
  TN_CHECK_LAUNCH();
  return TN_OK;
}

template <typename S, typename D>
__global__ void convert_kernel(const S* __restrict__ src, D* __restrict__ dst, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = (D)src[i];
}

template <typename T>
__global__ void scale_kernel(T* __restrict__ x, int64_t n, T alpha) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] *= alpha;
}

template <typename T>
__global__ void sqrt_kernel(T* __restrict__ x, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] = sqrt(x[i]);
}

// Two-pass FP64 sum of squares: per-block partials, then one block sums them.
template <typename T>
__global__ void __launch_bounds__(256) sumsq_partial(const T* __restrict__ x, int64_t n,
                                                     double* __restrict__ part) {
  __shared__ double red[4];
  double acc = 0.0;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    double v = (double)x[i];
    acc = fma(v, v, acc);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Counts of infinite and NaN entries (the non-finite diagnostic of a failed
// SVD); one block, so no zero-initialised counters are needed.
template <typename T>
__global__ void __launch_bounds__(256) nonfinite_count(const T* __restrict__ x, int64_t n,
                                                       unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[2][256];
  unsigned long long ninf = 0, nnan = 0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const double v = (double)x[i];
    nnan += (v != v) ? 1ull : 0ull;
    ninf += (v == v && (v - v) != (v - v)) ? 1ull : 0ull;
  }
  red[0][threadIdx.x] = ninf;
  red[1][threadIdx.x] = nnan;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 2) out[threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) sum_final(const double* __restrict__ part, int n,
                                                 double* __restrict__ out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += part[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = red[0] + red[1] + red[2] + red[3];
}

// out (n x n, row-major) = diag(1 / A_ii) with A read through its row and
// column strides; part[2b] / part[2b+1] = this block's count of nonzero
// off-diagonal entries / zero diagonal entries of A.
template <typename T>
__global__ void __launch_bounds__(256) diag_inverse_kernel(const T* __restrict__ A, int64_t n,
                                                           int64_t rs, int64_t cs,
                                                           T* __restrict__ out,
                                                           double* __restrict__ part) {
  __shared__ double red[2][4];
  double offd = 0.0, zdiag = 0.0;
  const int64_t nn = n * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e - i * n;
    const T v = A[i * rs + j * cs];
    if (i == j) {
      out[e] = (T)1 / v;
      zdiag += v == (T)0 ? 1.0 : 0.0;
    } else {
      out[e] = (T)0;
      offd += v != (T)0 ? 1.0 : 0.0;
    }
  }
  offd = wave_sum(offd);
  zdiag = wave_sum(zdiag);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = offd;
    red[1][threadIdx.x >> 6] = zdiag;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    part[2 * blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

__global__ void __launch_bounds__(256) count_final(const double* __restrict__ part, int nblk,
                                                   double* __restrict__ info) {
  __shared__ double red[2][4];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    info[0] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    info[1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) trace_kernel(const T* __restrict__ src, Modes out,
                                                    int64_t nout, int64_t dd, int64_t ds,
                                                    T* __restrict__ dst) {
  int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nout) return;
  int64_t base = mode_offset(out, o);
  double acc = 0.0;
  for (int64_t t = 0; t < dd; ++t) acc += (double)src[base + t * ds];
  dst[o] = (T)acc;
}

template <typename T>
__global__ void axpby_kernel(const T* __restrict__ x, const T* __restrict__ y, T* __restrict__ out,
                             int64_t n, double a, double b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (T)(a * (double)x[i] + b * (double)y[i]);
}

template <typename T>
__global__ void diag_embed_kernel(const double* __restrict__ s, int64_t k, T* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k * k) return;
  int64_t r = i / k, c = i - r * k;
  out[i] = (T)(r == c ? s[r] : 0.0);
}


// x /= sqrt(*sumsq) when positive; *norm_out = sqrt(*sumsq)
template <typename T>
__global__ void normalize_kernel(T* __restrict__ x, int64_t n, const double* __restrict__ sumsq,
                                 double* __restrict__ norm_out) {
  const double nrm = sqrt(*sumsq);
  if (blockIdx.x == 0 && threadIdx.x == 0 && norm_out) *norm_out = nrm;
  if (!(nrm > 0.0)) return;
  const T inv = (T)(1.0 / nrm);
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) x[i] *= inv;
}

static constexpr int kSumsqBlocks = 1024;

}  // namespace tn

using namespace tn;

extern "C" {

int tn_version(void) { return 1; }
const char* tn_last_error(void) { return tn::g_err.c_str(); }
int tn_last_error_code(void) { return tn::g_code; }

int tn_device_count(int* count) {
  TN_REQUIRE(count != nullptr, "count is NULL");
  TN_HIP(hipGetDeviceCount(count));
  return TN_OK;
}

int tn_permute(const void* src, const tn_modes* src_modes, void* dst, int dtype, void* stream) {
  TN_REQUIRE(src_modes != nullptr && src_modes->nmodes >= 0 && src_modes->nmodes <= TN_MAX_MODES,
             "tn_permute: bad modes");
  Modes m = to_modes(src_modes);
  for (int i = 0; i < m.n; ++i) TN_REQUIRE(m.dims[i] >= 0, "tn_permute: negative dim");
  if (modes_size(m) == 0) return TN_OK;
  TN_REQUIRE(src != nullptr && dst != nullptr, "tn_permute: null pointer");
  if (m.n == 0) {  // scalar
    m.n = 1; m.dims[0] = 1; m.strides[0] = 0;
  }
  if (dtype == TN_F64)
    return permute_impl((const double*)src, m, (double*)dst, as_stream(stream));
  if (dtype == TN_F32)
    return permute_impl((const float*)src, m, (float*)dst, as_stream(stream));
  set_error(TN_ERR_UNSUPPORTED, "tn_permute: unsupported dtype %d", dtype);
  return TN_ERR_UNSUPPORTED;
}

int tn_convert(const void* src, int src_dtype, void* dst, int dst_dtype, int64_t n, void* stream) {
  if (n == 0) return TN_OK;
  TN_REQUIRE(n > 0 && src && dst, "tn_convert: bad arguments");
  hipStream_t st = as_stream(stream);
  unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 256), 65536);
  if (src_dtype == TN_F64 && dst_dtype == TN_F32)
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks), dim3(256), 0, st,
                       (const double*)src, (float*)dst, n);
  else if (src_dtype == TN_F32 && dst_dtype == TN_F64)
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks), dim3(256), 0, st,
                       (const float*)src, (double*)dst, n);
  else if (src_dtype == dst_dtype && src_dtype == TN_F64)
    hipLaunchKernelGGL((convert_kernel<double, double>), dim3(blocks), dim3(256), 0, st,
                       (const double*)src, (double*)dst, n);
  else if (src_dtype == dst_dtype && src_dtype == TN_F32)
    hipLaunchKernelGGL((convert_kernel<float, float>), dim3(blocks), dim3(256), 0, st,
                       (const float*)src, (float*)dst, n);
  else {
    set_error(TN_ERR_UNSUPPORTED, "tn_convert: unsupported dtypes %d -> %d", src_dtype, dst_dtype);
    return TN_ERR_UNSUPPORTED;
  }
  TN_CHECK_LAUNCH();
  return TN_OK;
}

int tn_sqrt(void* x, int dtype, int64_t n, void* stream) {
  if (n == 0) return TN_OK;
  TN_REQUIRE(n > 0 && x, "tn_sqrt: bad arguments");
  hipStream_t st = as_stream(stream);
  unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 256), 65536);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(sqrt_kernel<double>, dim3(blocks), dim3(256), 0, st, (double*)x, n);
  else if (dtype == TN_F32)
    hipLaunchKernelGGL(sqrt_kernel<float>, dim3(blocks), dim3(256), 0, st, (float*)x, n);
  else {
    set_error(TN_ERR_UNSUPPORTED, "tn_sqrt: unsupported dtype %d", dtype);
    return TN_ERR_UNSUPPORTED;
  }
  TN_CHECK_LAUNCH();
  return TN_OK;
}

int tn_scale(void* x, int dtype, int64_t n, double alpha, void* stream) {
  if (n == 0) return TN_OK;
  TN_REQUIRE(n > 0 && x, "tn_scale: bad arguments");
  hipStream_t st = as_stream(stream);
  unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 256), 65536);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(scale_kernel<double>, dim3(blocks), dim3(256), 0, st, (double*)x, n, alpha);
  else if (dtype == TN_F32)
    hipLaunchKernelGGL(scale_kernel<float>, dim3(blocks), dim3(256), 0, st, (float*)x, n,
                       (float)alpha);
  else {
    set_error(TN_ERR_UNSUPPORTED, "tn_scale: unsupported dtype %d", dtype);
    return TN_ERR_UNSUPPORTED;
  }
  TN_CHECK_LAUNCH();
  return TN_OK;
}

int tn_count_nonfinite(const void* x, int dtype, int64_t n, unsigned long long* out,
                       void* stream) {
  TN_REQUIRE(n >= 0 && out != nullptr && (n == 0 || x != nullptr),
             "tn_count_nonfinite: bad arguments");
  hipStream_t st = as_stream(stream);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(nonfinite_count<double>, dim3(1), dim3(256), 0, st, (const double*)x, n,
                       out);
  else if (dtype == TN_F32)
    hipLaunchKernelGGL(nonfinite_count<float>, dim3(1), dim3(256), 0, st, (const float*)x, n,
                       out);
  else {
    set_error(TN_ERR_UNSUPPORTED, "tn_count_nonfinite: unsupported dtype %d", dtype);
    return TN_ERR_UNSUPPORTED;
  }
  TN_CHECK_LAUNCH();
  return TN_OK;
}

size_t tn_sumsq_workspace(int64_t n) {
  (void)n;
  return sizeof(double) * tn::kSumsqBlocks;
}

int tn_sumsq(const void* x, int dtype, int64_t n, double* out, void* ws, size_t ws_bytes,
             void* stream) {
  TN_REQUIRE(n >= 0 && out != nullptr, "tn_sumsq: bad arguments");
  TN_REQUIRE(ws != nullptr && ws_bytes >= tn_sumsq_workspace(n), "tn_sumsq: workspace too small");
  hipStream_t st = as_stream(stream);
  int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), tn::kSumsqBlocks));
  double* part = (double*)ws;
  if (n == 0) {
    hipLaunchKernelGGL(sum_final, dim3(1), dim3(256), 0, st, part, 0, out);
    TN_CHECK_LAUNCH();
    return TN_OK;
  }
  if (dtype == TN_F64)
    hipLaunchKernelGGL(sumsq_partial<double>, dim3(blocks), dim3(256), 0, st, (const double*)x, n,
                       part);
  else if (dtype == TN_F32)
    hipLaunchKernelGGL(sumsq_partial<float>, dim3(blocks), dim3(256), 0, st, (const float*)x, n,
                       part);
  else {
    set_error(TN_ERR_UNSUPPORTED, "tn_sumsq: unsupported dtype %d", dtype);
    return TN_ERR_UNSUPPORTED;
  }
  TN_CHECK_LAUNCH();
  hipLaunchKernelGGL(sum_final, dim3(1), dim3(256), 0, st, part, blocks, out);
  TN_CHECK_LAUNCH();
  return TN_OK;
}

int tn_trace(const void* src, const tn_modes* out_modes, int64_t diag_dim, int64_t diag_stride,
             void* dst, int dtype, void* stream) {
  TN_REQUIRE(out_modes && out_modes->nmodes >= 0 && out_modes->nmodes <= TN_MAX_MODES,
             "tn_trace: bad modes");
  Modes m = to_modes(out_modes);
  int64_t n = modes_size(m);
  if (n == 0) return TN_OK;
  hipStream_t st = as_stream(stream);
  unsigned blocks = (unsigned)ceil_div(n, 256);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(trace_kernel<double>, dim3(blocks), dim3(256), 0, st, (const double*)src, m,
                       n, diag_dim, diag_stride, (double*)dst);
  else if (dtype == TN_F32)
    hipLaunchKernelGGL(trace_kernel<float>, dim3(blocks), dim3(256), 0, st, (const float*)src, m, n,
                       diag_dim, diag_stride, (float*)dst);
  else {
    set_error(TN_ERR_UNSUPPORTED, "tn_trace: unsupported dtype %d", dtype);
    return TN_ERR_UNSUPPORTED;
  }
  TN_CHECK_LAUNCH();
  return TN_OK;
}

int tn_axpby(const void* x, const void* y, void* out, int dtype, int64_t n, double a, double b,
             void* stream) {
  if (n == 0) return TN_OK;
  TN_REQUIRE(n > 0 && x && y && out, "tn_axpby: bad arguments");
  hipStream_t st = as_stream(stream);
  unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n, 256), 65536);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(axpby_kernel<double>, dim3(blocks), dim3(256), 0, st, (const double*)x,
                       (const double*)y, (double*)out, n, a, b);
  else if (dtype == TN_F32)
    hipLaunchKernelGGL(axpby_kernel<float>, dim3(blocks), dim3(256), 0, st, (const float*)x,
                       (const float*)y, (float*)out, n, a, b);
  else {
    set_error(TN_ERR_UNSUPPORTED, "tn_axpby: unsupported dtype %d", dtype);
    return TN_ERR_UNSUPPORTED;
  }
  TN_CHECK_LAUNCH();
  return TN_OK;
}

int tn_diag_embed(const double* s, int64_t k, void* out, int dtype, void* stream) {
  if (k == 0) return TN_OK;
  TN_REQUIRE(k > 0 && s && out, "tn_diag_embed: bad arguments");
  hipStream_t st = as_stream(stream);
  unsigned blocks = (unsigned)ceil_div(k * k, 256);
  if (dtype == TN_F64)
    hipLaunchKernelGGL(diag_embed_kernel<double>, dim3(blocks), dim3(256), 0, st, s, k,
                       (double*)out);
  else if (dtype == TN_F32)
    hipLaunchKernelGGL(diag_embed_kernel<float>, dim3(blocks), dim3(256), 0, st, s, k,
                       (float*)out);
  else {
    set_error(TN_ERR_UNSUPPORTED, "tn_diag_embed: unsupported dtype %d", dtype);
    return TN_ERR_UNSUPPORTED;
  }
  TN_CHECK_LAUNCH();
  return TN_OK;
}

size_t tn_normalize_workspace(int64_t n) { return tn_sumsq_workspace(n) + 256; }

int tn_normalize(void* x, int dtype, int64_t n, double* norm_out, void* ws, size_t ws_bytes,
                 void* stream) {
  TN_REQUIRE(n >= 0 && x != nullptr, "tn_normalize: bad arguments");
  TN_REQUIRE(ws && ws_bytes >= tn_normalize_workspace(n), "tn_normalize: workspace too small");
  double* ss = (double*)((char*)ws + tn_sumsq_workspace(n));
  TN_TRY(tn_sumsq(x, dtype, n, ss, ws, tn_sumsq_workspace(n), stream));
  hipStream_t st = as_stream(stream);
  unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 65536));
  if (dtype == TN_F64)
    hipLaunchKernelGGL(normalize_kernel<double>, dim3(blocks), dim3(256), 0, st, (double*)x, n, ss,
                       norm_out);
  else
    hipLaunchKernelGGL(normalize_kernel<float>, dim3(blocks), dim3(256), 0, st, (float*)x, n, ss,
                       norm_out);
  TN_CHECK_LAUNCH();
  return TN_OK;
}

size_t tn_diag_inverse_workspace(int64_t n) {
  (void)n;
  return sizeof(double) * 2 * tn::kSumsqBlocks;
}

int tn_diag_inverse(const void* A, int64_t n, int64_t rs, int64_t cs, void* out, int dtype,
                    double* info, void* ws, size_t ws_bytes, void* stream) {
  TN_REQUIRE(n >= 0 && info != nullptr, "tn_diag_inverse: bad arguments");
  TN_REQUIRE(n == 0 || (A && out), "tn_diag_inverse: null pointer");
  TN_REQUIRE(ws && ws_bytes >= tn_diag_inverse_workspace(n), "tn_diag_inverse: workspace too small");
  TN_REQUIRE(dtype == TN_F64 || dtype == TN_F32, "tn_diag_inverse: unsupported dtype %d", dtype);
  hipStream_t st = as_stream(stream);
  double* part = (double*)ws;
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n * n, 256), tn::kSumsqBlocks));
  if (dtype == TN_F64)
    hipLaunchKernelGGL(diag_inverse_kernel<double>, dim3(blocks), dim3(256), 0, st,
                       (const double*)A, n, rs, cs, (double*)out, part);
  else
    hipLaunchKernelGGL(diag_inverse_kernel<float>, dim3(blocks), dim3(256), 0, st,
                       (const float*)A, n, rs, cs, (float*)out, part);
  TN_CHECK_LAUNCH();
  hipLaunchKernelGGL(count_final, dim3(1), dim3(256), 0, st, part, blocks, info);
  TN_CHECK_LAUNCH();
  return TN_OK;
}

}  // extern "C"
