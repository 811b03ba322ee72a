// Greedy rounding of the quasi-permutation to a permutation,
// cugal/greedy_lap.py:7-63 (GREEDY :10-19, DOUBLE_GREEDY :50-63).
//
// DOUBLE_GREEDY: rows are visited by descending row maximum (a stable LSD
// radix sort on the order-preserving key, so equal maxima keep row order;
// the reference's torch.argsort is not stable).  For each visited row r the
// reference multiplies P[r] by the 0/1 "not taken" vector in place and takes
// argmax (first index of the maximum).  Both steps are reproduced exactly:
// the taken entries of row r become P[r][j] * 0 in memory, and ties go to the
// lowest column.  The walk is sequential by definition; it runs in a single
// persistent workgroup with the taken set as a bitmap in LDS.
#include "common.hpp"
#include "cugal_hip.h"

namespace cg {
namespace {

constexpr int kLapThreads = 1024;
constexpr int kSortThreads = 1024;
constexpr int64_t kLapMaxN = int64_t(1) << 20;  // 128 KiB taken bitmap

__global__ __launch_bounds__(256) void row_max_kernel(const float* __restrict__ P, int n, float* __restrict__ rmax) {
    const int lane = threadIdx.x & 63;
    const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / 64;
    if (r >= n) return;
    const float* row = P + r * n;
    float m = -INFINITY;
    bool nan = false;
    for (int j = lane; j < n; j += 64) {
        const float x = row[j];
        nan |= (x != x);
        m = fmaxf(m, x);
    }
    m = wave_max(m);
    // torch.max propagates NaN.
    const bool any_nan = __any(nan);
    if (lane == 0) rmax[r] = any_nan ? NAN : m;
}

// Order-preserving key of a float, inverted so that an ascending sort gives
// descending values (NaN sorts first, as the largest).
__device__ __forceinline__ uint32_t desc_key(float f) {
    if (f != f) return 0u;
    if (f == 0.0f) f = 0.0f;  // -0 and +0 tie, as in a float comparison
    uint32_t b = __float_as_uint(f);
    b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ~b;
}

// Stable LSD radix sort, one bit per pass, in a single block.  Each thread
// owns a contiguous slice, so a stable split of each slice plus a block scan
// keeps the global order stable.
__global__ __launch_bounds__(kSortThreads) void stable_sort_desc(const float* __restrict__ vals, int n,
                                                                uint32_t* __restrict__ ka, int32_t* __restrict__ ia,
                                                                uint32_t* __restrict__ kb, int32_t* __restrict__ ib,
                                                                int32_t* __restrict__ order) {
    __shared__ int zeros_before[kSortThreads];
    __shared__ int total_zeros;
    const int per = static_cast<int>(div_up(n, kSortThreads));
    const int b = threadIdx.x * per;
    const int e = min(n, b + per);
    for (int i = b; i < e; ++i) {
        ka[i] = desc_key(vals[i]);
        ia[i] = i;
    }
    __syncthreads();
    uint32_t* ks = ka;
    int32_t* is = ia;
    uint32_t* kd = kb;
    int32_t* id = ib;
    for (int bit = 0; bit < 32; ++bit) {
        int z = 0;
        for (int i = b; i < e; ++i) z += ((ks[i] >> bit) & 1u) == 0u;
        zeros_before[threadIdx.x] = z;
        __syncthreads();
        if (threadIdx.x == 0) {
            int run = 0;
            for (int t = 0; t < kSortThreads; ++t) {
                const int v = zeros_before[t];
                zeros_before[t] = run;
                run += v;
            }
            total_zeros = run;
        }
        __syncthreads();
        int zpos = zeros_before[threadIdx.x];
        int opos = total_zeros + (b - zeros_before[threadIdx.x]);
        for (int i = b; i < e; ++i) {
            const uint32_t k = ks[i];
            const int dst = ((k >> bit) & 1u) ? opos++ : zpos++;
            kd[dst] = k;
            id[dst] = is[i];
        }
        __syncthreads();
        uint32_t* tk = ks;
        ks = kd;
        kd = tk;
        int32_t* ti = is;
        is = id;
        id = ti;
    }
    for (int i = b; i < e; ++i) order[i] = is[i];
}

__global__ void identity_order(int32_t* __restrict__ order, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) order[i] = i;
}

// Argmax pair: larger value wins, equal values -> lower index.
__device__ __forceinline__ void arg_better(float& v, int& i, float v2, int i2) {
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

__global__ __launch_bounds__(kLapThreads) void greedy_walk(float* __restrict__ P, int n,
                                                          const int32_t* __restrict__ order,
                                                          int32_t* __restrict__ assignment) {
    extern __shared__ uint32_t taken[];  // n bits
    __shared__ float wv[kLapThreads / kWave];
    __shared__ int wi[kLapThreads / kWave];
    const int words = static_cast<int>(div_up(n, 32));
    for (int w = threadIdx.x; w < words; w += kLapThreads) taken[w] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x / kWave;
    for (int k = 0; k < n; ++k) {
        const int r = order[k];
        float* row = P + static_cast<int64_t>(r) * n;
        float bv = -INFINITY;
        int bi = 0x7fffffff;
        for (int j = threadIdx.x; j < n; j += kLapThreads) {
            float x = row[j];
            if ((taken[j >> 5] >> (j & 31)) & 1u) {
                x = x * 0.0f;
                row[j] = x;
            }
            // NaN never wins here (the reference's argmax would pick it).
            arg_better(bv, bi, x, j);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const float ov = __shfl_xor(bv, off, kWave);
            const int oi = __shfl_xor(bi, off, kWave);
            arg_better(bv, bi, ov, oi);
        }
        if (lane == 0) {
            wv[wave] = bv;
            wi[wave] = bi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            float v = wv[0];
            int i = wi[0];
            for (int w = 1; w < kLapThreads / kWave; ++w) arg_better(v, i, wv[w], wi[w]);
            if (i == 0x7fffffff) i = 0;  // all entries NaN / -inf: argmax of torch returns index 0
            assignment[r] = i;
            taken[i >> 5] |= 1u << (i & 31);
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

size_t cg_lap_workspace(int64_t n) {
    if (n <= 0) return 0;
    Carver cv{nullptr, 0, ~size_t(0)};
    cv.take<float>(n);
    cv.take<uint32_t>(n);
    cv.take<int32_t>(n);
    cv.take<uint32_t>(n);
    cv.take<int32_t>(n);
    cv.take<int32_t>(n);
    return cv.used + 256;
}

int cg_lap_greedy(float* P, int64_t n, int method, int32_t* assignment, void* workspace, size_t workspace_bytes,
                  cg_stream_t stream) {
    if (!P || !assignment || n <= 0) return fail(CG_EINVAL, "cg_lap_greedy: bad arguments");
    if (n > kLapMaxN) return fail(CG_EUNSUPPORTED, "cg_lap_greedy: n=%lld exceeds %lld", static_cast<long long>(n),
                                  static_cast<long long>(kLapMaxN));
    if (method != 0 && method != 1) return fail(CG_EINVAL, "cg_lap_greedy: unknown method %d", method);
    if (!workspace || workspace_bytes < cg_lap_workspace(n)) return fail(CG_EWORKSPACE, "cg_lap_greedy: workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    float* rmax = cv.take<float>(n);
    uint32_t* ka = cv.take<uint32_t>(n);
    int32_t* ia = cv.take<int32_t>(n);
    uint32_t* kb = cv.take<uint32_t>(n);
    int32_t* ib = cv.take<int32_t>(n);
    int32_t* order = cv.take<int32_t>(n);
    const int ni = static_cast<int>(n);
    if (method == 1) {
        hipLaunchKernelGGL(row_max_kernel, dim3(static_cast<unsigned>(div_up(n, 4))), dim3(256), 0, s, P, ni, rmax);
        if (int rc = check_launch("row_max_kernel")) return rc;
        hipLaunchKernelGGL(stable_sort_desc, dim3(1), dim3(kSortThreads), 0, s, rmax, ni, ka, ia, kb, ib, order);
        if (int rc = check_launch("stable_sort_desc")) return rc;
    } else {
        hipLaunchKernelGGL(identity_order, dim3(static_cast<unsigned>(div_up(n, 256))), dim3(256), 0, s, order, ni);
        if (int rc = check_launch("identity_order")) return rc;
    }
    const size_t lds = sizeof(uint32_t) * static_cast<size_t>(div_up(n, 32));
    static bool attr_set = false;
    if (!attr_set && lds > 32768) {
        int dev = 0, v = 65536;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&greedy_walk),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, v);
        attr_set = true;
    }
    hipLaunchKernelGGL(greedy_walk, dim3(1), dim3(kLapThreads), lds, s, P, ni, order, assignment);
    return check_launch("greedy_walk");
}

}  // extern "C"
