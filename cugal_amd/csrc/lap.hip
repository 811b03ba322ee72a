// Greedy rounding of the quasi-permutation to a permutation,
// cugal/greedy_lap.py:7-85 (GREEDY :10-19, RAND :21-32, MORE_RAND :34-48,
// DOUBLE_GREEDY :50-63, PARALLEL_GREEDY :65-85).
//
// Reference semantics: rows are visited in order (GREEDY: 0..n-1; RAND and
// MORE_RAND: a random permutation; DOUBLE_GREEDY: by descending row maximum;
// PARALLEL_GREEDY: rows whose maximum is > 0.5 take their argmax at once, the
// rest are visited by descending maximum); each visited row r is multiplied
// in place by the 0/1 "not taken" vector and the first index of its maximum
// is taken (MORE_RAND: the first index whose prefix sum reaches total * U).
//
// MI355X form (exact, not a heuristic):
//   1. row_topk: every row's 16 largest entries ordered by (value desc,
//      index asc), one wave per row, fully parallel;
//   2. DOUBLE_GREEDY / PARALLEL_GREEDY: a stable descending sort of the row
//      maxima (the reference's sorts are not stable: with tied maxima they
//      may visit rows in another order); PARALLEL_GREEDY then splits the
//      sorted rows into the parallel set and the walk order in one block;
//   3. walk: one block visits the rows in order.  The first candidate not yet
//      taken is the argmax of the masked row whenever its value is > 0: every
//      non-candidate is after it in (value desc, index asc) order and masked
//      columns read 0.  Otherwise (all 16 taken, or a non-positive best) the
//      block scans the full masked row.  The next row's candidates are loaded
//      while the current one is decided.  MORE_RAND walks with a block-wide
//      FP64 prefix sum of each masked row instead;
//   4. mask: P[r][j] *= 0 for every column j taken before row r was visited,
//      which is what the in-place masking leaves in the matrix.  Steps count
//      from 1; columns taken by PARALLEL_GREEDY's parallel set carry step 0,
//      and its parallel rows (step 0) are never masked.
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
                                                           int32_t* __restrict__ fallbacks,
                                                           const int32_t* __restrict__ walk_count,
                                                           const uint32_t* __restrict__ taken_init) {
    extern __shared__ uint32_t taken[];  // n bits
    __shared__ int s_choice;
    __shared__ float wv[kWalkThreads / kWave];
    __shared__ int wi[kWalkThreads / kWave];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int words = static_cast<int>(div_up(n, 32));
    for (int w = tid; w < words; w += kWalkThreads) taken[w] = taken_init ? taken_init[w] : 0u;
    const int count = walk_count ? *walk_count : n;
    __syncthreads();
    int nfall = 0;
    // Wave 0 keeps the current row's candidates in registers and loads the
    // next row's while deciding.
    float v = -INFINITY;
    int c = 0x7fffffff;
    if (wave == 0 && lane < kTopK && count > 0) {
        const int r0 = order[0];
        v = cval[static_cast<int64_t>(r0) * kTopK + lane];
        c = cidx[static_cast<int64_t>(r0) * kTopK + lane];
    }
    const bool vec = (n % 4) == 0;
    for (int k = 0; k < count; ++k) {
        const int r = order[k];
        if (wave == 0) {
            float vn = -INFINITY;
            int cn = 0x7fffffff;
            if (k + 1 < count && lane < kTopK) {
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
            step_of_row[r] = k + 1;
            if (step_of_col[choice] < 0) step_of_col[choice] = k + 1;
            taken[choice >> 5] |= 1u << (choice & 31);
        }
        __syncthreads();
    }
    if (tid == 0 && fallbacks) *fallbacks = nfall;
}

// ---- PARALLEL_GREEDY: the parallel set and the walk order ----------------
// One block.  Rows of the sorted order whose maximum is > 0.5 take their
// first-index argmax (the head of their candidate list) and mark it taken;
// the others keep their sorted order in walk_order (a stable compaction).
// NaN maxima are not > 0.5, so those rows are walked, as in the reference.
__global__ __launch_bounds__(kSortThreads) void split_parallel(const int32_t* __restrict__ order, int n,
                                                              const float* __restrict__ rmax,
                                                              const int32_t* __restrict__ cidx,
                                                              int32_t* __restrict__ assignment,
                                                              int32_t* __restrict__ step_of_col,
                                                              int32_t* __restrict__ step_of_row,
                                                              uint32_t* __restrict__ taken_init,
                                                              int32_t* __restrict__ walk_order,
                                                              int32_t* __restrict__ walk_count) {
    __shared__ int before_me[kSortThreads];
    const int words = static_cast<int>(div_up(n, 32));
    for (int w = threadIdx.x; w < words; w += kSortThreads) taken_init[w] = 0u;
    const int per = static_cast<int>(div_up(n, kSortThreads));
    const int b = threadIdx.x * per;
    const int e = min(n, b + per);
    int walked = 0;
    for (int k = b; k < e; ++k) walked += !(rmax[order[k]] > 0.5f);
    before_me[threadIdx.x] = walked;
    __syncthreads();  // also orders the bitmap clear before the atomics below
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < kSortThreads; ++t) {
            const int v = before_me[t];
            before_me[t] = run;
            run += v;
        }
        *walk_count = run;
    }
    __syncthreads();
    int pos = before_me[threadIdx.x];
    for (int k = b; k < e; ++k) {
        const int r = order[k];
        if (rmax[r] > 0.5f) {
            const int c = cidx[static_cast<int64_t>(r) * kTopK];
            assignment[r] = c;
            step_of_row[r] = 0;
            step_of_col[c] = 0;
            atomicOr(&taken_init[c >> 5], 1u << (c & 31));
        } else {
            walk_order[pos++] = r;
        }
    }
}

// ---- RAND / MORE_RAND input check ----------------------------------------
// order must be a permutation of 0..n-1 and every draw must lie in [0, 1].
__global__ __launch_bounds__(kSortThreads) void check_random_inputs(const int32_t* __restrict__ order, int n,
                                                                   const double* __restrict__ uniforms,
                                                                   uint32_t* __restrict__ seen,
                                                                   int32_t* __restrict__ bad) {
    const int words = static_cast<int>(div_up(n, 32));
    for (int w = threadIdx.x; w < words; w += kSortThreads) seen[w] = 0u;
    __shared__ int s_bad;
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    int my_bad = 0;
    for (int k = threadIdx.x; k < n; k += kSortThreads) {
        const int r = order[k];
        if (r < 0 || r >= n) {
            my_bad = 1;
            continue;
        }
        const uint32_t bit = 1u << (r & 31);
        if (atomicOr(&seen[r >> 5], bit) & bit) my_bad = 1;
        if (uniforms) {
            const double u = uniforms[k];
            if (!(u >= 0.0 && u <= 1.0)) my_bad = 1;
        }
    }
    if (my_bad) atomicOr(&s_bad, 1);
    __syncthreads();
    if (threadIdx.x == 0) *bad = s_bad;
}

// ---- MORE_RAND: sampling walk --------------------------------------------
// One block of 1024.  For the k-th visited row r: thread t owns a contiguous
// slice of columns and runs the FP64 prefix sum of the masked row over it,
// starting from the block-wide exclusive prefix of the slice sums; that
// running sum is the row's cumsum.  r_k = total * U_k, and the column is the
// first j with r_k <= cumsum_j: the first slice whose running maximum
// reaches r_k holds it, and its owner rescans the slice to find it.  No such
// column (only possible with negative entries, or a NaN) writes -1.
constexpr int kSampleThreads = 1024;

__global__ __launch_bounds__(kSampleThreads) void sample_walk(const float* __restrict__ P, int n,
                                                             const int32_t* __restrict__ order,
                                                             const double* __restrict__ uniforms,
                                                             int32_t* __restrict__ assignment,
                                                             int32_t* __restrict__ step_of_col,
                                                             int32_t* __restrict__ step_of_row) {
    extern __shared__ uint32_t taken[];  // n bits
    __shared__ double wtot[kSampleThreads / kWave];
    __shared__ double s_total;
    __shared__ int s_first;
    __shared__ int s_choice;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int words = static_cast<int>(div_up(n, 32));
    for (int w = tid; w < words; w += kSampleThreads) taken[w] = 0u;
    const int per = static_cast<int>(div_up(n, kSampleThreads));
    const int b = min(n, tid * per);
    const int e = min(n, b + per);
    const int last_owner = (n - 1) / per;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const int r = order[k];
        const float* row = P + static_cast<int64_t>(r) * n;
        double acc = 0.0;
        for (int j = b; j < e; ++j) {
            const float x = is_taken(taken, j) ? row[j] * 0.0f : row[j];
            acc += static_cast<double>(x);
        }
        // Block exclusive prefix of the slice sums.
        double incl = acc;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const double o = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += o;
        }
        double excl = __shfl_up(incl, 1, kWave);
        if (lane == 0) excl = 0.0;
        if (lane == kWave - 1) wtot[wave] = incl;
        if (tid == 0) s_first = 0x7fffffff;
        __syncthreads();
        double wbase = 0.0;
        for (int w = 0; w < wave; ++w) wbase += wtot[w];
        excl += wbase;
        // Running sums of the slice: the end value (the total, for the
        // owner of column n-1) and the maximum.
        double run = excl, top = -INFINITY;
        for (int j = b; j < e; ++j) {
            const float x = is_taken(taken, j) ? row[j] * 0.0f : row[j];
            run += static_cast<double>(x);
            top = fmax(top, run);
        }
        if (tid == last_owner) s_total = run;
        __syncthreads();
        const double target = s_total * uniforms[k];
        if (b < e && target <= top) atomicMin(&s_first, tid);
        __syncthreads();
        const int first = s_first;
        if (tid == first) {
            int choice = -1;
            double rs = excl;
            for (int j = b; j < e; ++j) {
                const float x = is_taken(taken, j) ? row[j] * 0.0f : row[j];
                rs += static_cast<double>(x);
                if (target <= rs) {
                    choice = j;
                    break;
                }
            }
            s_choice = choice;
        }
        __syncthreads();
        if (tid == 0) {
            const int choice = first == 0x7fffffff ? -1 : s_choice;
            assignment[r] = choice;
            step_of_row[r] = k + 1;
            if (choice >= 0) {
                if (step_of_col[choice] < 0) step_of_col[choice] = k + 1;
                taken[choice >> 5] |= 1u << (choice & 31);
            }
        }
        __syncthreads();
    }
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

namespace {

struct LapWs {
    float* cval;
    int32_t* cidx;
    float* rmax;
    uint32_t* ka;
    int32_t* ia;
    uint32_t* kb;
    int32_t* ib;
    int32_t* order;
    int32_t* step_col;
    int32_t* step_row;
    int32_t* nfall;
    int32_t* walk_order;
    uint32_t* taken_init;
    int32_t* walk_count;
    uint32_t* seen;
    int32_t* bad;
};

LapWs lap_carve(char* base, int64_t n, size_t bytes, size_t* used = nullptr) {
    Carver cv{base, 0, bytes};
    const int64_t words = div_up(n, 32);
    LapWs w;
    w.cval = cv.take<float>(n * kTopK);
    w.cidx = cv.take<int32_t>(n * kTopK);
    w.rmax = cv.take<float>(n);
    w.ka = cv.take<uint32_t>(n);
    w.ia = cv.take<int32_t>(n);
    w.kb = cv.take<uint32_t>(n);
    w.ib = cv.take<int32_t>(n);
    w.order = cv.take<int32_t>(n);
    w.step_col = cv.take<int32_t>(n);
    w.step_row = cv.take<int32_t>(n);
    w.nfall = cv.take<int32_t>(1);
    w.walk_order = cv.take<int32_t>(n);
    w.taken_init = cv.take<uint32_t>(words);
    w.walk_count = cv.take<int32_t>(1);
    w.seen = cv.take<uint32_t>(words);
    w.bad = cv.take<int32_t>(1);
    if (used) *used = cv.used;
    return w;
}

// Raise the dynamic LDS limit of a walk kernel to what its n-bit taken
// bitmap needs (the kernels also hold a little static LDS, so exactly the
// dynamic bytes are requested; the attribute only grows).
int allow_lds(const void* kernel, size_t lds, size_t& granted) {
    if (lds <= 32768 || lds <= granted) return CG_OK;
    CG_CHECK_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    granted = lds;
    return CG_OK;
}

enum class Walk { kGreedy, kDouble, kParallel, kRand, kMoreRand };

int lap_run(float* P, int64_t n, Walk method, const int32_t* order_in, const double* uniforms, int32_t* assignment,
            void* workspace, size_t workspace_bytes, hipStream_t s, const char* who) {
    if (n > kLapMaxN)
        return fail(CG_EUNSUPPORTED, "%s: n=%lld exceeds %lld", who, static_cast<long long>(n),
                    static_cast<long long>(kLapMaxN));
    if (!workspace || workspace_bytes < cg_lap_workspace(n)) return fail(CG_EWORKSPACE, "%s: workspace too small", who);
    LapWs w = lap_carve(static_cast<char*>(workspace), n, workspace_bytes);
    const int ni = static_cast<int>(n);
    const dim3 vgrid(static_cast<unsigned>(div_up(n, 256)));
    const bool random = method == Walk::kRand || method == Walk::kMoreRand;
    if (random) {
        // A bad order would index out of bounds: check it before any walk.
        hipLaunchKernelGGL(check_random_inputs, dim3(1), dim3(kSortThreads), 0, s, order_in, ni, uniforms, w.seen,
                           w.bad);
        if (int rc = check_launch("check_random_inputs")) return rc;
        int bad = 1;
        CG_CHECK_HIP(hipMemcpyAsync(&bad, w.bad, sizeof(int), hipMemcpyDeviceToHost, s));
        CG_CHECK_HIP(hipStreamSynchronize(s));
        if (bad)
            return fail(CG_EINVAL, "%s: order is not a permutation of 0..n-1 or a draw is outside [0, 1]", who);
    }
    const int32_t* order = random ? order_in : w.order;
    if (method != Walk::kMoreRand) {
        hipLaunchKernelGGL(row_topk, dim3(static_cast<unsigned>(div_up(n, 4))), dim3(256), 0, s, P, ni, w.cval, w.cidx,
                           w.rmax);
        if (int rc = check_launch("row_topk")) return rc;
    }
    if (method == Walk::kDouble || method == Walk::kParallel) {
        hipLaunchKernelGGL(stable_sort_desc, dim3(1), dim3(kSortThreads), 0, s, w.rmax, ni, w.ka, w.ia, w.kb, w.ib,
                           w.order);
        if (int rc = check_launch("stable_sort_desc")) return rc;
    } else if (method == Walk::kGreedy) {
        hipLaunchKernelGGL(identity_order, vgrid, dim3(256), 0, s, w.order, ni);
        if (int rc = check_launch("identity_order")) return rc;
    }
    hipLaunchKernelGGL(fill_int, vgrid, dim3(256), 0, s, w.step_col, ni, -1);
    if (int rc = check_launch("fill_int")) return rc;
    const size_t lds = sizeof(uint32_t) * static_cast<size_t>(div_up(n, 32));
    if (method == Walk::kMoreRand) {
        static size_t sample_attr = 0;
        if (int rc = allow_lds(reinterpret_cast<const void*>(&sample_walk), lds, sample_attr)) return rc;
        hipLaunchKernelGGL(sample_walk, dim3(1), dim3(kSampleThreads), lds, s, P, ni, order, uniforms, assignment,
                           w.step_col, w.step_row);
        if (int rc = check_launch("sample_walk")) return rc;
    } else {
        const int32_t* walk_count = nullptr;
        const uint32_t* taken_init = nullptr;
        if (method == Walk::kParallel) {
            hipLaunchKernelGGL(split_parallel, dim3(1), dim3(kSortThreads), 0, s, w.order, ni, w.rmax, w.cidx,
                               assignment, w.step_col, w.step_row, w.taken_init, w.walk_order, w.walk_count);
            if (int rc = check_launch("split_parallel")) return rc;
            order = w.walk_order;
            walk_count = w.walk_count;
            taken_init = w.taken_init;
        }
        static size_t walk_attr = 0;
        if (int rc = allow_lds(reinterpret_cast<const void*>(&greedy_walk), lds, walk_attr)) return rc;
        hipLaunchKernelGGL(greedy_walk, dim3(1), dim3(kWalkThreads), lds, s, P, ni, order, w.cval, w.cidx, assignment,
                           w.step_col, w.step_row, w.nfall, walk_count, taken_init);
        if (int rc = check_launch("greedy_walk")) return rc;
    }
    hipLaunchKernelGGL(mask_taken, dim3(static_cast<unsigned>(n)), dim3(256), 0, s, P, ni, w.step_col, w.step_row);
    return check_launch("mask_taken");
}

}  // namespace

extern "C" {

size_t cg_lap_workspace(int64_t n) {
    if (n <= 0) return 0;
    size_t used = 0;
    (void)lap_carve(nullptr, n, ~size_t(0), &used);
    return used + 256;
}

int cg_lap_greedy(float* P, int64_t n, int method, int32_t* assignment, void* workspace, size_t workspace_bytes,
                  cg_stream_t stream) {
    if (!P || !assignment || n <= 0) return fail(CG_EINVAL, "cg_lap_greedy: bad arguments");
    if (method < 0 || method > 2) return fail(CG_EINVAL, "cg_lap_greedy: unknown method %d", method);
    const Walk walk = method == 0 ? Walk::kGreedy : method == 1 ? Walk::kDouble : Walk::kParallel;
    return lap_run(P, n, walk, nullptr, nullptr, assignment, workspace, workspace_bytes,
                   reinterpret_cast<hipStream_t>(stream), "cg_lap_greedy");
}

int cg_lap_greedy_random(float* P, int64_t n, int method, const int32_t* order, const double* uniforms,
                         int32_t* assignment, void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    if (!P || !assignment || !order || n <= 0) return fail(CG_EINVAL, "cg_lap_greedy_random: bad arguments");
    if (method != 0 && method != 1) return fail(CG_EINVAL, "cg_lap_greedy_random: unknown method %d", method);
    if (method == 1 && !uniforms) return fail(CG_EINVAL, "cg_lap_greedy_random: MORE_RAND needs the draws");
    return lap_run(P, n, method == 0 ? Walk::kRand : Walk::kMoreRand, order, method == 1 ? uniforms : nullptr,
                   assignment, workspace, workspace_bytes, reinterpret_cast<hipStream_t>(stream),
                   "cg_lap_greedy_random");
}

}  // extern "C"
