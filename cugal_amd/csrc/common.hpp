// Shared device helpers for the CDNA4 (gfx950, wave64) kernels of the FUGAL
// alignment pipeline, plus host-side error plumbing for the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstddef>
#include <cstdint>

namespace cg {

constexpr int kWave = 64;
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLn2 = 0.69314718055994531f;
// Finite stand-in for -inf in running maxima: keeps (x - m) free of inf-inf.
constexpr float kNegBig = -FLT_MAX;

// v_exp_f32 / v_log_f32: base-2 hardware transcendentals (about 1 ulp).
// Results below 2^-126 may flush to zero; every sum below adds terms that
// are <= 1 relative to the running maximum, so a flushed term changes a
// log-sum-exp by at most n * 2^-126 relative.
__device__ __forceinline__ float exp2_fast(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float log2_fast(float x) { return __builtin_amdgcn_logf(x); }

// Streaming loads that bypass cache retention (data read once per pass).
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_stream4(const float* p) {
    const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float ld_stream(const float* p) { return __builtin_nontemporal_load(p); }

// The same 16-byte loads at 4-byte alignment, for the rows of a matrix whose
// width is not a multiple of 4: still one dwordx4 access (global memory
// needs dword alignment only).
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ float4 ld_stream4u(const float* p) {
    const f32x4u v = __builtin_nontemporal_load(reinterpret_cast<const f32x4u*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 ld4u(const float* p) {
    const f32x4u v = *reinterpret_cast<const f32x4u*>(p);
    return make_float4(v.x, v.y, v.z, v.w);
}

// Elements col .. col + 3 of a length-n vector, `pad` past its end (scalar
// loads: nothing past n is read).
__device__ __forceinline__ float4 ld_tail4(const float* p, int col, int n, float pad) {
    return make_float4(col < n ? p[col] : pad, col + 1 < n ? p[col + 1] : pad, col + 2 < n ? p[col + 2] : pad,
                       col + 3 < n ? p[col + 3] : pad);
}

// Running (max, scaled sum) pair of a base-2 log-sum-exp.
struct Lse {
    float m;
    float s;
};

// Merge two (max, sum) pairs.  A pair with s == 0 is the identity.
__device__ __forceinline__ Lse lse_merge(Lse a, Lse b) {
    const float m = fmaxf(a.m, b.m);
    const float s = a.s * exp2_fast(a.m - m) + b.s * exp2_fast(b.m - m);
    return {m, s};
}

// Online update of (m, s) with one term x: one exp2 per term.
__device__ __forceinline__ void lse_push(float& m, float& s, float x) {
    const float d = x - m;
    const float e = exp2_fast(-fabsf(d));
    const bool up = d > 0.0f;
    s = up ? fmaf(s, e, 1.0f) : s + e;
    m = up ? x : m;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Wave-wide log-sum-exp of per-lane pairs: max first, then one rescale per
// lane and a plain sum, so only one exp2 per lane.
__device__ __forceinline__ Lse wave_lse(Lse p) {
    const float m = wave_max(p.m);
    const float s = wave_sum(p.s * exp2_fast(p.m - m));
    return {m, s};
}

// Block-wide double sum; `scratch` needs blockDim.x / 64 doubles.  Result is
// valid in thread 0 only.  Summation order is fixed, hence deterministic.
__device__ __forceinline__ double block_sum_double(double v, double* scratch) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    v = wave_sum(v);
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double total = 0.0;
    if (threadIdx.x == 0) {
        const int waves = blockDim.x / kWave;
        for (int w = 0; w < waves; ++w) total += scratch[w];
    }
    __syncthreads();
    return total;
}

// Feature rows of FD floats: 4 (degree, clustering, mean neighbour degree,
// mean neighbour clustering) or 8 (FUGAL's 7 features and a zero).  The
// width is a template parameter so the 4-float epilogues carry no extra
// address arithmetic or branch per element.
struct FeatureRow {
    float4 a0, a1;
};

template <int FD>
__device__ __forceinline__ FeatureRow feature_row(const float* __restrict__ f, int64_t i) {
    static_assert(FD == 4 || FD == 8, "feature rows hold 4 or 8 floats");
    FeatureRow r;
    r.a0 = *reinterpret_cast<const float4*>(f + FD * i);
    r.a1 = FD == 8 ? *reinterpret_cast<const float4*>(f + FD * i + 4) : make_float4(0.f, 0.f, 0.f, 0.f);
    return r;
}

// Euclidean distance between two feature rows (the direct form; the first
// four squares are summed in the same order for either width).
template <int FD>
__device__ __forceinline__ float feature_distance_rows(const FeatureRow& a, const FeatureRow& b) {
    const float d0 = a.a0.x - b.a0.x, d1 = a.a0.y - b.a0.y, d2 = a.a0.z - b.a0.z, d3 = a.a0.w - b.a0.w;
    float s = fmaf(d0, d0, fmaf(d1, d1, fmaf(d2, d2, d3 * d3)));
    if constexpr (FD == 8) {
        const float e0 = a.a1.x - b.a1.x, e1 = a.a1.y - b.a1.y, e2 = a.a1.z - b.a1.z, e3 = a.a1.w - b.a1.w;
        s += fmaf(e0, e0, fmaf(e1, e1, fmaf(e2, e2, e3 * e3)));
    }
    return sqrtf(s);
}

// Distance between row `a` and row j of ft.
template <int FD>
__device__ __forceinline__ float feature_distance(const FeatureRow& a, const float* __restrict__ ft, int64_t j) {
    return feature_distance_rows<FD>(a, feature_row<FD>(ft, j));
}

__host__ __device__ constexpr int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }
__host__ __device__ constexpr int64_t round_up(int64_t a, int64_t b) { return div_up(a, b) * b; }

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// ---------------------------------------------------------------------------
// Host-side status handling shared by every C-ABI entry point.
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);
// Checks hipGetLastError() after a launch; returns 0 or a negative code.
int check_launch(const char* what);

// Carve aligned sub-buffers out of a caller-provided workspace.
struct Carver {
    char* base;
    size_t used;
    size_t cap;
    template <typename T>
    T* take(size_t count) {
        used = static_cast<size_t>(round_up(static_cast<int64_t>(used), 256));
        T* p = reinterpret_cast<T*>(base ? base + used : nullptr);
        used += count * sizeof(T);
        return p;
    }
    bool ok() const { return used <= cap; }
};

}  // namespace cg

// Status codes returned by every entry point (0 = success).
#define CG_OK 0
#define CG_EINVAL (-1)
#define CG_EWORKSPACE (-2)
#define CG_ELAUNCH (-3)
#define CG_EUNSUPPORTED (-4)
#define CG_EHIP (-5)

#define CG_CHECK_HIP(expr)                                                        \
    do {                                                                          \
        hipError_t cg_err_ = (expr);                                              \
        if (cg_err_ != hipSuccess)                                                \
            return ::cg::fail(CG_EHIP, "%s failed: %s", #expr, hipGetErrorString(cg_err_)); \
    } while (0)
