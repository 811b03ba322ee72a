// Greedy rounding of the quasi-permutation to a permutation,
// cugal/greedy_lap.py:7-63 (GREEDY :10-19, DOUBLE_GREEDY :50-63).
//
// Reference semantics: rows are visited in order (GREEDY: 0..n-1;
// DOUBLE_GREEDY: by descending row maximum); each visited row r is
// multiplied in place by the 0/1 "not taken" vector and the first index of
// its maximum is taken.
//
// MI355X form (exact, not a heuristic):
//   1. row_topk: every row's 16 largest entries ordered by (value desc,
//      index asc), one wave per row, fully parallel;
//   2. DOUBLE_GREEDY only: a stable descending sort of the row maxima (the
//      reference's argsort is not stable: with tied maxima it may visit rows
//      in another order);
//   3. walk: one block visits the rows in order.  The first candidate not yet
//      taken is the argmax of the masked row whenever its value is > 0: every
//      non-candidate is after it in (value desc, index asc) order and masked
//      columns read 0.  Otherwise (all 16 taken, or a non-positive best) the
//      block scans the full masked row.  The next row's candidates are loaded
//      while the current one is decided.
//   4. mask: P[r][j] *= 0 for every column j taken before row r was visited,
//      which is what the in-place masking leaves in the matrix.
#include "common.hpp"
#include "cugal_hip.h"

namespace cg {
namespace {

constexpr int kTopK = 16;
constexpr int kSortThreads = 1024;
constexpr int64_t kLapMaxN = int64_t(1) << 20;  // 128 KiB taken bitmap

// (value, index) order: larger value first, then lower index.
__device__ __forceinline__ bool before(float v, int i, float v2, int i2) { return v > v2 || (v == v2 && i < i2); }

// ---- 1. per-row top-K ----------------------------------------------------
__global__ __launch_bounds__(256) void row_topk(const float* __restrict__ P, int n, float* __restrict__ cval,
                                               int32_t* __restrict__ cidx, float* __restrict__ rmax) {
    const int lane = threadIdx.x & 63;
    const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / 64;
    if (r >= n) return;
    const float* row = P + r * n;
    float lv[kTopK];
    int li[kTopK];
#pragma unroll
    for (int k = 0; k < kTopK; ++k) {
        lv[k] = -INFINITY;
        li[k] = 0x7fffffff;
    }
    bool nan = false;
    // A lane sees its columns in increasing index order, so a new element
    // enters only if strictly larger than the list's last (ties keep the
    // lower index already held).
    for (int j = lane; j < n; j += 64) {
        const float x = row[j];
        nan |= (x != x);
        if (x > lv[kTopK - 1]) {
            float cv = x;
            int ci = j;
#pragma unroll
            for (int k = 0; k < kTopK; ++k) {
                if (before(cv, ci, lv[k], li[k])) {
                    const float tv = lv[k];
                    const int ti = li[k];
                    lv[k] = cv;
                    li[k] = ci;
                    cv = tv;
                    ci = ti;
                }
            }
        }
    }
    // Merge the 64 lane lists: kTopK rounds of a wave-wide best head.
    int head = 0;
    for (int k = 0; k < kTopK; ++k) {
        float hv = -INFINITY;
        int hi = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < kTopK; ++q)
            if (q == head) {
                hv = lv[q];
                hi = li[q];
            }
        float bv = hv;
        int bi = hi;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const float ov = __shfl_xor(bv, off, kWave);
            const int oi = __shfl_xor(bi, off, kWave);
            if (before(ov, oi, bv, bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (hi == bi && bi != 0x7fffffff) ++head;  // the owner pops its head
        if (lane == 0) {
            cval[r * kTopK + k] = bv;
            cidx[r * kTopK + k] = bi;
        }
    }
    const bool any_nan = __any(nan);
    if (lane == 0 && rmax) rmax[r] = any_nan ? NAN : cval[r * kTopK];  // torch.max propagates NaN
}

// ---- 2. stable sort of the row maxima (descending) -----------------------
__device__ __forceinline__ uint32_t desc_key(float f) {
    if (f != f) return 0u;  // NaN first, as the largest
    if (f == 0.0f) f = 0.0f;  // -0 and +0 tie, as in a float comparison
    uint32_t b = __float_as_uint(f);
    b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ~b;
}

// LSD radix sort, one bit per pass, in one block.  Each thread owns a
// contiguous slice, so a stable split of each slice plus a block scan keeps
// the global order stable.
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

// ---- 3. the sequential walk ----------------------------------------------
// Wave 0 decides each row from its candidates; the other waves join only the
// rare full scans (all candidates taken, or a best value that is not > 0).
constexpr int kWalkThreads = 256;

__device__ __forceinline__ bool is_taken(const uint32_t* taken, int j) { return (taken[j >> 5] >> (j & 31)) & 1u; }

__global__ __launch_bounds__(kWalkThreads) void greedy_walk(const float* __restrict__ P, int n,
                                                           const int32_t* __restrict__ order,
                                                           const float* __restrict__ cval,
                                                           const int32_t* __restrict__ cidx,
                                                           int32_t* __restrict__ assignment,
                                                           int32_t* __restrict__ step_of_col,
                                                           int32_t* __restrict__ step_of_row,
                                                           int32_t* __restrict__ fallbacks) {
    extern __shared__ uint32_t taken[];  // n bits
    __shared__ int s_choice;
    __shared__ float wv[kWalkThreads / kWave];
    __shared__ int wi[kWalkThreads / kWave];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int words = static_cast<int>(div_up(n, 32));
    for (int w = tid; w < words; w += kWalkThreads) taken[w] = 0u;
    __syncthreads();
    int nfall = 0;
    // Wave 0 keeps the current row's candidates in registers and loads the
    // next row's while deciding.
    float v = -INFINITY;
    int c = 0x7fffffff;
    if (wave == 0 && lane < kTopK) {
        const int r0 = order[0];
        v = cval[static_cast<int64_t>(r0) * kTopK + lane];
        c = cidx[static_cast<int64_t>(r0) * kTopK + lane];
    }
    const bool vec = (n % 4) == 0;
    for (int k = 0; k < n; ++k) {
        const int r = order[k];
        if (wave == 0) {
            float vn = -INFINITY;
            int cn = 0x7fffffff;
            if (k + 1 < n && lane < kTopK) {
                const int rn = order[k + 1];
                vn = cval[static_cast<int64_t>(rn) * kTopK + lane];
                cn = cidx[static_cast<int64_t>(rn) * kTopK + lane];
            }
            const bool free_c = lane < kTopK && c != 0x7fffffff && !is_taken(taken, c);
            const unsigned long long m = __ballot(free_c);
            int choice = -1;
            if (m) {
                const int first = __ffsll(m) - 1;
                const float bv = __shfl(v, first, kWave);
                const int bc = __shfl(c, first, kWave);
                if (bv > 0.0f) choice = bc;
            }
            if (lane == 0) s_choice = choice;
            v = vn;
            c = cn;
        }
        __syncthreads();
        int choice = s_choice;
        if (choice < 0) {
            // Full scan of the masked row by the whole block: taken columns
            // read x * 0; first index of the maximum.
            const float* row = P + static_cast<int64_t>(r) * n;
            float bv = -INFINITY;
            int bi = 0x7fffffff;
            if (vec) {
                for (int j = tid * 4; j < n; j += kWalkThreads * 4) {
                    const float4 x4 = *reinterpret_cast<const float4*>(row + j);
                    const float xs[4] = {x4.x, x4.y, x4.z, x4.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const float x = is_taken(taken, j + e) ? xs[e] * 0.0f : xs[e];
                        if (before(x, j + e, bv, bi)) {
                            bv = x;
                            bi = j + e;
                        }
                    }
                }
            } else {
                for (int j = tid; j < n; j += kWalkThreads) {
                    const float x = is_taken(taken, j) ? row[j] * 0.0f : row[j];
                    if (before(x, j, bv, bi)) {
                        bv = x;
                        bi = j;
                    }
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const float ov = __shfl_xor(bv, off, kWave);
                const int oi = __shfl_xor(bi, off, kWave);
                if (before(ov, oi, bv, bi)) {
                    bv = ov;
                    bi = oi;
                }
            }
            if (lane == 0) {
                wv[wave] = bv;
                wi[wave] = bi;
            }
            __syncthreads();
            if (tid == 0) {
                float fv = wv[0];
                int fi = wi[0];
                for (int w = 1; w < kWalkThreads / kWave; ++w)
                    if (before(wv[w], wi[w], fv, fi)) {
                        fv = wv[w];
                        fi = wi[w];
                    }
                s_choice = fi == 0x7fffffff ? 0 : fi;  // an all-NaN row: index 0
            }
            __syncthreads();
            choice = s_choice;
            ++nfall;
        }
        if (tid == 0) {
            assignment[r] = choice;
            step_of_row[r] = k;
            if (step_of_col[choice] < 0) step_of_col[choice] = k;
            taken[choice >> 5] |= 1u << (choice & 31);
        }
        __syncthreads();
    }
    if (tid == 0 && fallbacks) *fallbacks = nfall;
}

__global__ void fill_int(int32_t* __restrict__ p, int n, int value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// ---- 4. the in-place masking the reference leaves in P -------------------
__global__ __launch_bounds__(256) void mask_taken(float* __restrict__ P, int n,
                                                 const int32_t* __restrict__ step_of_col,
                                                 const int32_t* __restrict__ step_of_row) {
    const int64_t r = blockIdx.x;
    const int s = step_of_row[r];
    float* row = P + r * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int t = step_of_col[j];
        if (t >= 0 && t < s) row[j] = row[j] * 0.0f;
    }
}

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

size_t cg_lap_workspace(int64_t n) {
    if (n <= 0) return 0;
    Carver cv{nullptr, 0, ~size_t(0)};
    cv.take<float>(n * kTopK);
    cv.take<int32_t>(n * kTopK);
    cv.take<float>(n);
    cv.take<uint32_t>(n);
    cv.take<int32_t>(n);
    cv.take<uint32_t>(n);
    cv.take<int32_t>(n);
    cv.take<int32_t>(n);
    cv.take<int32_t>(n);
    cv.take<int32_t>(n);
    cv.take<int32_t>(1);
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
    float* cval = cv.take<float>(n * kTopK);
    int32_t* cidx = cv.take<int32_t>(n * kTopK);
    float* rmax = cv.take<float>(n);
    uint32_t* ka = cv.take<uint32_t>(n);
    int32_t* ia = cv.take<int32_t>(n);
    uint32_t* kb = cv.take<uint32_t>(n);
    int32_t* ib = cv.take<int32_t>(n);
    int32_t* order = cv.take<int32_t>(n);
    int32_t* step_col = cv.take<int32_t>(n);
    int32_t* step_row = cv.take<int32_t>(n);
    int32_t* nfall = cv.take<int32_t>(1);
    const int ni = static_cast<int>(n);
    const dim3 vgrid(static_cast<unsigned>(div_up(n, 256)));
    hipLaunchKernelGGL(row_topk, dim3(static_cast<unsigned>(div_up(n, 4))), dim3(256), 0, s, P, ni, cval, cidx, rmax);
    if (int rc = check_launch("row_topk")) return rc;
    if (method == 1) {
        hipLaunchKernelGGL(stable_sort_desc, dim3(1), dim3(kSortThreads), 0, s, rmax, ni, ka, ia, kb, ib, order);
        if (int rc = check_launch("stable_sort_desc")) return rc;
    } else {
        hipLaunchKernelGGL(identity_order, vgrid, dim3(256), 0, s, order, ni);
        if (int rc = check_launch("identity_order")) return rc;
    }
    hipLaunchKernelGGL(fill_int, vgrid, dim3(256), 0, s, step_col, ni, -1);
    if (int rc = check_launch("fill_int")) return rc;
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
    hipLaunchKernelGGL(greedy_walk, dim3(1), dim3(kWalkThreads), lds, s, P, ni, order, cval, cidx, assignment,
                       step_col, step_row, nfall);
    if (int rc = check_launch("greedy_walk")) return rc;
    hipLaunchKernelGGL(mask_taken, dim3(static_cast<unsigned>(n)), dim3(256), 0, s, P, ni, step_col, step_row);
    return check_launch("mask_taken");
}

}  // extern "C"
