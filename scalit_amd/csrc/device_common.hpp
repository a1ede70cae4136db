// Device helpers shared by the operator kernels (H*X, P*X): the 16x16x4 MFMA
// wrappers with their lane layouts, storage <-> compute conversions, the
// preconditioner's per-value kernel and the per-vector activity mask.
#pragma once

#include "scalit_impl.hpp"

namespace scalit {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA: A(i,k) comes from lane i + 16 k, B(k,j) from lane j + 16 k;
// accumulator q of lane l holds D(row(l/16, q), l % 16), with the rows
// interleaved for FP64 and blocked for FP32 (checked on the device by
// tests/test_gpu_mfma_layout.py).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    using acc_t = d4_t;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int q) { return lk + 4 * q; }
};
template <>
struct Mfma<float> {
    using acc_t = f4_t;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int q) { return 4 * lk + q; }
};

// storage <-> compute conversions
__device__ __forceinline__ float bf16_to_f32(bf16_t v) { return __uint_as_float(((uint32_t)v) << 16); }
__device__ __forceinline__ bf16_t f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)(u >> 16);   // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);                                   // round to nearest even
    return (bf16_t)(u >> 16);
}
template <typename T, typename S>
__device__ __forceinline__ T ld_cvt(const S *p, long i) { return (T)p[i]; }
template <>
__device__ __forceinline__ float ld_cvt<float, bf16_t>(const bf16_t *p, long i) { return bf16_to_f32(p[i]); }
template <typename T, typename S>
__device__ __forceinline__ void st_cvt(S *p, long i, T v) { p[i] = (S)v; }
template <>
__device__ __forceinline__ void st_cvt<float, bf16_t>(bf16_t *p, long i, float v) { p[i] = f32_to_bf16(v); }

// PX_KERNEL (osb.px.h:156-211) for one value
template <typename T>
__device__ __forceinline__ T px_kernel(int nosb, T y, T eig, T e0, T de, T beta)
{
    T ade = eig - e0;
    if (nosb == SCALIT_TOSBD1) {
        if (fabs(ade) <= de) return ade < (T)0 ? -y * ((T)1 / de) : y * ((T)1 / de);
        return y / ade;
    }
    if (nosb == SCALIT_TOSBD2) {
        T b1 = (T)1 + beta, al = -beta * de, t;
        if (fabs(ade) <= de) t = ade < (T)0 ? (-al + b1 * ade) : (al + b1 * ade);
        else t = ade;
        return y / t;
    }
    return y / ade;
}

__device__ __forceinline__ bool vec_active(const int *active, int v) { return active == nullptr || active[v] != 0; }

// Targets of an address select that makes an optional load unconditional:
// an activity flag for "every vector active" and the unit factor.
static __device__ const int k_all_on = 1;
static __device__ const double k_unit_d = 1.0;
static __device__ const float k_unit_f = 1.0f;
template <typename T>
__device__ __forceinline__ const T *unit_ptr()
{
    if constexpr (sizeof(T) == sizeof(double))
        return (const T *)&k_unit_d;
    else
        return (const T *)&k_unit_f;
}

// Raw buffer access: a wave-uniform resource (base, size in bytes) plus a
// 32-bit per-lane byte offset and a wave-uniform byte offset, so a column of
// loads at a uniform stride needs one VGPR of address in all.  Callers keep
// every access inside the resource.  Dword 3 = 0x00020000: 32-bit data
// format, the value for untyped loads on gfx9-family parts.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0,
                                             (int)(bytes > 0x7fffffffL ? 0x7fffffffL : bytes), 0x00020000);
}
template <typename T, typename S>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ double buf_ld<double, double>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float buf_ld<float, float>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float buf_ld<float, bf16_t>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return bf16_to_f32((bf16_t)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
}
__device__ __forceinline__ int buf_ld_i32(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return (int)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
template <typename T, typename S>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, int voff, int soff, T v);
template <>
__device__ __forceinline__ void buf_st<double, double>(__amdgpu_buffer_rsrc_t r, int voff, int soff, double v)
{
    __builtin_amdgcn_raw_buffer_store_b64(
        __builtin_bit_cast(__typeof__(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, voff, soff, 0);
}
template <>
__device__ __forceinline__ void buf_st<float, float>(__amdgpu_buffer_rsrc_t r, int voff, int soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}
template <>
__device__ __forceinline__ void buf_st<float, bf16_t>(__amdgpu_buffer_rsrc_t r, int voff, int soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)f32_to_bf16(v), r, voff, soff, 0);
}

__host__ __device__ static inline int rup16(int n) { return (n + 15) & ~15; }

// e / d for 0 <= e < 2^22 and d >= 1, with inv = 1.0f / d: (e + 1/2) / d lies
// at least 1/(2d) from an integer, and the fp32 product is off by at most
// (e / d) 2^-23 < 1/(2d), so truncation gives the exact quotient in four
// VALU operations instead of an integer division's twenty-odd.
__device__ __forceinline__ int fdiv_small(int e, float inv) { return (int)(((float)e + 0.5f) * inv); }

// Copy count contiguous values g[0..count) into LDS, value e to slot(e), with
// U independent loads in flight per thread (NT threads): a block of tens of
// KB then arrives in a few memory latencies instead of one latency per
// NT values.
template <int NT, int U, typename T, typename Slot>
__device__ __forceinline__ void stage_lds(const T *__restrict__ g, int count, int tid, Slot slot)
{
    for (int base = 0; base < count; base += NT * U) {
        T r[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = base + u * NT + tid;
            r[u] = e < count ? g[e] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = base + u * NT + tid;
            if (e < count) slot(e) = r[u];
        }
    }
}

// XCD-aware remap of a 1-D grid: consecutive workgroup ids land on different
// XCDs (8 on MI355X, round-robin); this gives each XCD a contiguous range of
// logical ids, so that workgroups sharing data (one rotation matrix, one
// theta slab) share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk)
{
    if ((nblk & 7) != 0) return bid;
    return (bid & 7) * (nblk >> 3) + (bid >> 3);
}

}  // namespace scalit
