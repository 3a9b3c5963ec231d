// Shared definitions for the tncontract_amd HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>

#include "../../include/tncontract_amd.h"

namespace tn {

// ---------------------------------------------------------------- errors
void set_error(int code, const char* fmt, ...);
int last_code();

#define TN_HIP(call)                                                          \
  do {                                                                        \
    hipError_t _e = (call);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::tn::set_error(TN_ERR_HIP, "%s failed: %s (%s:%d)", #call,             \
                      hipGetErrorString(_e), __FILE__, __LINE__);             \
      return TN_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

#define TN_CHECK_LAUNCH()                                                     \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess) {                                                   \
      ::tn::set_error(TN_ERR_HIP, "kernel launch failed: %s (%s:%d)",         \
                      hipGetErrorString(_e), __FILE__, __LINE__);             \
      return TN_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

#define TN_REQUIRE(cond, ...)                                                 \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ::tn::set_error(TN_ERR_INVALID, __VA_ARGS__);                           \
      return TN_ERR_INVALID;                                                  \
    }                                                                         \
  } while (0)

#define TN_TRY(expr)                                                          \
  do {                                                                        \
    int _rc = (expr);                                                         \
    if (_rc != TN_OK) return _rc;                                             \
  } while (0)

// ---------------------------------------------------------------- modes
// Device-side copy of tn_modes (passed by value in kernel arguments).
struct Modes {
  int32_t n;
  int64_t dims[TN_MAX_MODES];
  int64_t strides[TN_MAX_MODES];
};

inline Modes to_modes(const tn_modes* m) {
  Modes r{};
  if (m == nullptr) { r.n = 0; return r; }
  r.n = m->nmodes;
  for (int i = 0; i < TN_MAX_MODES; ++i) {
    r.dims[i] = i < m->nmodes ? m->dims[i] : 1;
    r.strides[i] = i < m->nmodes ? m->strides[i] : 0;
  }
  return r;
}

inline int64_t modes_size(const Modes& m) {
  int64_t s = 1;
  for (int i = 0; i < m.n; ++i) s *= m.dims[i];
  return s;
}

// Offset of linear (row-major over dims, slowest first) index `idx`.
__device__ __forceinline__ int64_t mode_offset(const Modes& m, int64_t idx) {
  int64_t off = 0;
  for (int i = m.n - 1; i >= 0; --i) {
    int64_t d = m.dims[i];
    int64_t q = idx / d;
    off += (idx - q * d) * m.strides[i];
    idx = q;
  }
  return off;
}

// Fast path when idx fits in 32 bits.
__device__ __forceinline__ int64_t mode_offset32(const Modes& m, uint32_t idx) {
  int64_t off = 0;
  for (int i = m.n - 1; i >= 0; --i) {
    uint32_t d = (uint32_t)m.dims[i];
    uint32_t q = idx / d;
    off += (int64_t)(idx - q * d) * m.strides[i];
    idx = q;
  }
  return off;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------- reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One DPP step of a wave reduction on an FP64 value: the 32-bit halves move
// with the same control; lanes of rows outside ROWMASK receive 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, ROWMASK, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the wave by DPP (VALU moves, no LDS traffic): quad butterflies,
// half-row and row mirrors give every lane its 16-lane row sum, row
// broadcasts 15 and 31 carry rows 0..2 into lane 63, whose value every lane
// then reads.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_f64<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x141, 0xF>(v);  // row_half_mirror
  v += dpp_f64<0x140, 0xF>(v);  // row_mirror
  v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)u, 63);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// The FP64 value of lane `lane` (wave-uniform) in every lane.
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)u, lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Orders a wave's LDS writes before its later LDS reads by other lanes of the
// same wave: the wave issues its LDS instructions in order, and the fences
// keep the compiler from moving accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Maximum over the wave of an FP32 value by DPP, as wave_sum_dpp (exact in
// any order, so bit-identical to a shuffle reduction).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f32(float v, float fill) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), CTRL,
                                                    ROWMASK, 0xF, false));
}
__device__ __forceinline__ float wave_max_dpp(float v) {
  v = fmaxf(v, dpp_f32<0xB1, 0xF>(v, v));
  v = fmaxf(v, dpp_f32<0x4E, 0xF>(v, v));
  v = fmaxf(v, dpp_f32<0x141, 0xF>(v, v));
  v = fmaxf(v, dpp_f32<0x140, 0xF>(v, v));
  v = fmaxf(v, dpp_f32<0x142, 0xA>(v, v));
  v = fmaxf(v, dpp_f32<0x143, 0xC>(v, v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = v > w ? v : w;
  }
  return v;
}

// Workspace carving: bump allocator over a caller-provided buffer.
struct Arena {
  char* base;
  size_t cap;
  size_t used;
  template <typename T>
  T* take(size_t count) {
    size_t off = (used + 255) & ~size_t(255);
    size_t bytes = count * sizeof(T);
    if (base == nullptr) {  // size query mode
      used = off + bytes;
      return nullptr;
    }
    if (off + bytes > cap) return nullptr;
    used = off + bytes;
    return reinterpret_cast<T*>(base + off);
  }
};

}  // namespace tn
