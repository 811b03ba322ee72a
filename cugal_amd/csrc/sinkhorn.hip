// Log-domain Sinkhorn-Knopp for the FUGAL Frank-Wolfe loop on CDNA4.
//
// Semantics follow cugal/sinkhorn.py:90-120 (solve_log): with K = -C/reg,
//   v_j = -LSE_i(K_ij + u_i),   u_i = -LSE_j(K_ij + v_j)
// per iteration, a column-marginal L1 error every `eval_freq` iterations
// (not at iteration 0) and a stop when it drops below the threshold.  The
// kernels work in base-2 units (u2 = u*log2(e), K2 = C*(-log2(e)/reg)), so
// every exponential is one v_exp_f32.
//
// Memory traffic: C (n x n FP32, row-major) is never rescaled, transposed or
// copied.  Column log-sum-exps are formed from per-row-block partial
// (max, sum) pairs that a small combine kernel merges, so
//   * the fused path reads C once per iteration: a row block computes u_i for
//     each of its rows and, with the row still in registers, folds
//     K2_ij + u2_i + v2_j into running column pairs (the next v and the
//     column marginal of the current (u, v) both come out of the combine);
//   * the two-sweep path reads C twice per iteration (column sweep, then row
//     sweep), the same structure as the reference kernels
//     (cuda_kernels/sinkhorn_log.cu:39-66 applied to K^T and K).
// Convergence is decided on the device: the last block of a checking combine
// records the error and raises `done`; every later kernel of the solve exits
// at entry, so the host only polls once per check period and never stalls
// the queue.
#include "common.hpp"
#include "cugal_hip.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace cg {
namespace {

struct SinkCtl {
    int done;        // 1 once the threshold is met
    int iterations;  // iteration count at the stop (valid when done)
    int final_buf;   // v buffer holding the final v (valid when done)
    int nerr;        // errors recorded so far
    unsigned ticket; // arrival counter of the combine's blocks
    float momentum;  // current over-relaxation factor (1 = plain Sinkhorn)
    int pad0;
    double momentum_d;  // the same factor unrounded (the FP64 solve's)
    // followed by `double errors[capacity]`
};

__device__ __forceinline__ double* ctl_errors(SinkCtl* c) {
    return reinterpret_cast<double*>(reinterpret_cast<char*>(c) + sizeof(SinkCtl));
}

// ---------------------------------------------------------------------------
// Set-up and tear-down.
// ---------------------------------------------------------------------------
__global__ void sk_init(const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ u2,
                        float* __restrict__ v2, int n, SinkCtl* __restrict__ ctl) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) {
        u2[j] = u[j] * kLog2e;
        v2[j] = v[j] * kLog2e;
    }
    if (j == 0) {
        ctl->done = 0;
        ctl->iterations = 0;
        ctl->final_buf = 0;
        ctl->nerr = 0;
        ctl->ticket = 0u;
        ctl->momentum = 1.0f;
    ctl->momentum_d = 1.0;
        ctl->momentum_d = 1.0;
    }
}

__global__ void sk_finish(const float* __restrict__ u2, const float* __restrict__ v2a,
                          const float* __restrict__ v2b, float* __restrict__ u, float* __restrict__ v, int n,
                          const SinkCtl* __restrict__ ctl, int default_buf) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int buf = ctl->done ? ctl->final_buf : default_buf;
    u[j] = u2[j] * kLn2;
    v[j] = (buf ? v2b[j] : v2a[j]) * kLn2;
}

// ---------------------------------------------------------------------------
// Fused sweep: one read of C per iteration.
//
// Block b owns rows [b*rpb, (b+1)*rpb) and every column.  Thread t owns the
// float4 column groups g*T + t (g < G), i.e. 4G columns, so a row load is one
// coalesced float4 per group.  The current v sits in LDS.  Per row:
//   x_j  = K2_ij + v2_j                         (registers)
//   u2_i = (1-m) u2_i - m * LSE2_j(x_j)         (one block reduction)
//   s_j += 2^(x_j + u2_i)                       (= the plan entry)
// The column sums need no running maximum: the sweep always follows a
// v-update, after which every column of the plan sums to 1; the row
// normalisation then leaves every entry <= 1 and every column sum >= 1/n
// (sum_i r_i = n with row sums r_i <= n).  So s_j never overflows, and terms
// flushed below 2^-126 change a column sum by at most n * 2^-126 / (1/n)
// relative.  (With Lehmann over-relaxation, factor <= 2, the bounds widen to
// n^2, still far inside FP32.)  The next row is loaded while the current one
// is reduced (two register buffers).
// ---------------------------------------------------------------------------
// FULL: n == 4 G T, every column group exists (no bounds tests).
// KEEP: caching loads instead of non-temporal ones (see kCacheableMaxN).
// Without FULL any width works: a row of a matrix whose width is not a
// multiple of 4 starts at a 4-byte boundary only, so whole groups are
// 16-byte loads at 4-byte alignment, and the group that crosses column n
// loads its columns below n one by one (nothing past the row is read; the
// last row ends the matrix) and holds 0 above.
// Every launch has G = ceil(n / 4T) (checked on the host), so every group
// but a thread's last holds columns below n only: the bounds tests below
// are on the last group alone, and the other groups' loads are issued
// unconditionally (a branch around each would make the wave wait for every
// load before it uses the first).
template <int T, int G, bool FULL>
__device__ __forceinline__ bool group_whole(int g, int n) {
    return FULL || g + 1 < G || (g * T + static_cast<int>(threadIdx.x)) * 4 + 3 < n;
}
template <int T, int G, bool FULL>
__device__ __forceinline__ bool group_live(int g, int n) {
    return FULL || g + 1 < G || (g * T + static_cast<int>(threadIdx.x)) * 4 < n;
}

// w moved down by sh elements (0 <= sh <= 3; zeros shifted in) in two
// select stages, then the elements at or past column n set to 0: selects
// only, so the loads before it stay in flight.
__device__ __forceinline__ float4 shift_down_masked(float4 w, int sh, int col, int n) {
    const bool s1 = sh & 1, s2 = sh & 2;
    const float t0 = s1 ? w.y : w.x, t1 = s1 ? w.z : w.y, t2 = s1 ? w.w : w.z, t3 = s1 ? 0.0f : w.w;
    const float u0 = s2 ? t2 : t0, u1 = s2 ? t3 : t1, u2 = s2 ? 0.0f : t2, u3 = s2 ? 0.0f : t3;
    return make_float4(col < n ? u0 : 0.0f, col + 1 < n ? u1 : 0.0f, col + 2 < n ? u2 : 0.0f, col + 3 < n ? u3 : 0.0f);
}

template <int T, int G, bool FULL = false, bool KEEP = false>
__device__ __forceinline__ void fused_load_row(float4 (&r)[G], const float* __restrict__ row, int n) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int col = (g * T + threadIdx.x) * 4;
        if (FULL) {
            r[g] = KEEP ? *reinterpret_cast<const float4*>(row + col) : ld_stream4(row + col);
        } else if (g + 1 < G) {
            r[g] = KEEP ? ld4u(row + col) : ld_stream4u(row + col);
        } else {
            // The last group: one 16-byte load at the column clamped into the
            // row (n >= 4), its elements moved down by the clamp's shift and
            // those at or past n set to 0.  No branch: with a per-element
            // path beside the 16-byte one, the two paths' registers overlap
            // and the wave drained every earlier load of the row before
            // issuing this one (n = 4100: 29.3 against 22.8 us per LOG
            // iteration at n = 4096).
            const int cc = min(col, n - 4);
            const float4 w = KEEP ? ld4u(row + cc) : ld_stream4u(row + cc);
            // col - cc is 0 for a whole group, 1-3 for the group cut by n
            // and larger only for a group wholly past n (all masked).
            r[g] = shift_down_masked(w, min(col - cc, 3), col, n);
        }
    }
}

// Row stride of the fused sweeps' column partials: whole float4 groups, so
// the group that crosses column n stores inside its own partial row.
__host__ __device__ constexpr int64_t part_ld(int64_t n) { return round_up(n, 4); }

// Up to this width (a 576 MiB matrix) the fused sweeps read C with caching
// loads, and the LOG sweep walks its rows backwards on every other
// iteration, so a sweep starts on the rows the previous one read last,
// which the L2s and the 256 MB memory-side cache still hold.  Measured per
// LOG iteration: 11 % less time at n = 4096, 8 % at 8192, 2-3 % at 12288;
// even at 16384 and 9 % more at 32768, where non-temporal loads stream
// faster.  The STANDARD sweep keeps its forward order; caching loads alone
// take 10 % off at 4096 and 8 % at 8192, nothing at 12288.
constexpr int kCacheableMaxN = 12288;
constexpr int kStdCacheableMaxN = 8192;

template <int T, int G, bool FULL = false>
__device__ __forceinline__ void fused_row(float4 (&r)[G], float (&cs)[4 * G], int i, int n, float scale2,
                                          const float* __restrict__ sv, float2* __restrict__ red,
                                          float* __restrict__ u2, float mom) {
    constexpr int W = T / kWave;
    const int tid = threadIdx.x;
    const float u_old = u2[i];
    float lm = kNegBig;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int col = (g * T + tid) * 4;
        if (group_live<T, G, FULL>(g, n)) {
            const float4 vv = *reinterpret_cast<const float4*>(sv + col);
            r[g].x = fmaf(r[g].x, scale2, vv.x);
            r[g].y = fmaf(r[g].y, scale2, vv.y);
            r[g].z = fmaf(r[g].z, scale2, vv.z);
            r[g].w = fmaf(r[g].w, scale2, vv.w);
            lm = fmaxf(lm, fmaxf(fmaxf(r[g].x, r[g].y), fmaxf(r[g].z, r[g].w)));
        }
    }
    float ls = 0.0f;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (group_live<T, G, FULL>(g, n))
            ls += (exp2_fast(r[g].x - lm) + exp2_fast(r[g].y - lm)) + (exp2_fast(r[g].z - lm) + exp2_fast(r[g].w - lm));
    }
    const Lse w = wave_lse({lm, ls});
    float2* slot = red + (i & 1) * W;
    if ((tid & (kWave - 1)) == 0) slot[tid / kWave] = make_float2(w.m, w.s);
    __syncthreads();
    float M = kNegBig;
#pragma unroll
    for (int k = 0; k < W; ++k) M = fmaxf(M, slot[k].x);
    float S = 0.0f;
#pragma unroll
    for (int k = 0; k < W; ++k) S += slot[k].y * exp2_fast(slot[k].x - M);
    const float lse = M + log2_fast(S);
    const float ui = (mom == 1.0f) ? -lse : (1.0f - mom) * u_old - mom * lse;
    if (tid == 0) u2[i] = ui;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (group_live<T, G, FULL>(g, n)) {
            cs[4 * g + 0] += exp2_fast(r[g].x + ui);
            cs[4 * g + 1] += exp2_fast(r[g].y + ui);
            cs[4 * g + 2] += exp2_fast(r[g].z + ui);
            cs[4 * g + 3] += exp2_fast(r[g].w + ui);
        }
    }
}

// The same row step with the previous u2_i as the shift instead of the row
// maximum: the sweep follows a v-update, so every e_j = 2^(K2_ij + v2_j +
// u2_old_i) is a plan entry <= 1 and their sum S (the row mass) is at most
// n.  Then u2_i = u2_old - m log2 S, and the column terms 2^(K2 + v2 + u2_i)
// are e_j S^-m: one exponential per element instead of two, and no max
// pass.  If S leaves [2^-100, 2^100] (a row whose mass collapsed or
// exploded since the last solve; the decision is block-uniform) the row
// contributes nothing here and is queued; the block redoes it with the
// max-shifted step after its other rows.  Returns false for a queued row.
// Block barrier that orders LDS only: waits for this wave's LDS operations,
// not for its global loads, so the next row's loads stay in flight across
// the row reduction (__syncthreads would also drain them).  The memory
// clobber keeps the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <int T, int G, bool FULL>
__device__ __forceinline__ bool fused_row_shift(float4 (&r)[G], float (&cs)[4 * G], int i, int parity, int n,
                                                float scale2, const float* __restrict__ sv, float2* __restrict__ red,
                                                float* __restrict__ u2, float u_old, float mom) {
    constexpr int W = T / kWave;
    const int tid = threadIdx.x;
    float ls = 0.0f;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int col = (g * T + tid) * 4;
        if (group_live<T, G, FULL>(g, n)) {
            const float4 vv = *reinterpret_cast<const float4*>(sv + col);
            r[g].x = exp2_fast(fmaf(r[g].x, scale2, vv.x) + u_old);
            r[g].y = exp2_fast(fmaf(r[g].y, scale2, vv.y) + u_old);
            r[g].z = exp2_fast(fmaf(r[g].z, scale2, vv.z) + u_old);
            r[g].w = exp2_fast(fmaf(r[g].w, scale2, vv.w) + u_old);
            ls += (r[g].x + r[g].y) + (r[g].z + r[g].w);
        }
    }
    ls = wave_sum(ls);
    // Consecutive rows of a block alternate between two reduction slots, so
    // a wave that runs ahead never overwrites a slot still being read.
    float2* slot = red + parity * W;
    if ((tid & (kWave - 1)) == 0) slot[tid / kWave] = make_float2(0.0f, ls);
    lds_barrier();
    float S = 0.0f;
#pragma unroll
    for (int k = 0; k < W; ++k) S += slot[k].y;
    if (!(S >= 0x1p-100f && S <= 0x1p100f)) return false;
    const float l2 = log2_fast(S);
    if (tid == 0) u2[i] = u_old - mom * l2;
    const float f = (mom == 1.0f) ? 1.0f / S : exp2_fast(-mom * l2);
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (group_live<T, G, FULL>(g, n)) {
            cs[4 * g + 0] = fmaf(r[g].x, f, cs[4 * g + 0]);
            cs[4 * g + 1] = fmaf(r[g].y, f, cs[4 * g + 1]);
            cs[4 * g + 2] = fmaf(r[g].z, f, cs[4 * g + 2]);
            cs[4 * g + 3] = fmaf(r[g].w, f, cs[4 * g + 3]);
        }
    }
    return true;
}

constexpr int kFusedMaxRowsPerBlock = 128;  // n <= 32768 over 256 blocks


template <int T, int G, bool FULL, bool KEEP>
__global__ __launch_bounds__(T) void sk_fused_sweep(const float* __restrict__ C, int n, float scale2,
                                                   const float* __restrict__ v2, float* __restrict__ u2,
                                                   float* __restrict__ part, int rows_per_blk, int reverse,
                                                   const SinkCtl* __restrict__ ctl) {
    extern __shared__ float smem[];
    __shared__ int queued[kFusedMaxRowsPerBlock];
    __shared__ float su[kFusedMaxRowsPerBlock];  // the block's incoming u2
    __shared__ int nqueued;
    const int tid = threadIdx.x;
    float* sv = smem;
    float2* red = reinterpret_cast<float2*>(smem + round_up(n, 4));
    // Rows are dealt round-robin: block b takes rows b, b + nblk, b + 2 nblk,
    // ..., so at any moment the blocks read neighbouring rows (measured
    // 1.5-2 % faster than contiguous row ranges at n = 32768).
    const int nblk = static_cast<int>(gridDim.x);
    const int rows = min(rows_per_blk, (n - static_cast<int>(blockIdx.x) + nblk - 1) / nblk);
    const int64_t ld = n;
    if (ctl->done) return;
    const float mom = ctl->momentum;
    if (tid == 0) nqueued = 0;
    if constexpr (T == 256) {
        // The thread's slice of v (it reads back only its own columns) and
        // the block's incoming u, all loads in flight at once (the 1024-
        // thread sweep keeps the loop: the staging registers would spill).
        // Columns from n up to the group's end hold kNegBig: their plan
        // entries 2^(0 * scale2 + kNegBig + u) are 0.
        float4 vg[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int col = (g * T + tid) * 4;
            vg[g] = group_whole<T, G, FULL>(g, n) ? *reinterpret_cast<const float4*>(v2 + col)
                                                  : ld_tail4(v2, col, n, kNegBig);
        }
        const float u_in = tid < rows ? u2[blockIdx.x + tid * nblk] : 0.0f;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int col = (g * T + tid) * 4;
            if (group_live<T, G, FULL>(g, n)) *reinterpret_cast<float4*>(sv + col) = vg[g];
        }
        if (tid < rows) su[tid] = u_in;
    } else {
        for (int j = tid * 4; j < n; j += T * 4)
            *reinterpret_cast<float4*>(sv + j) =
                j + 3 < n ? *reinterpret_cast<const float4*>(v2 + j) : ld_tail4(v2, j, n, kNegBig);
        for (int k = tid; k < rows; k += T) su[k] = u2[blockIdx.x + k * nblk];
    }
    __syncthreads();

    float cs[4 * G];
#pragma unroll
    for (int k = 0; k < 4 * G; ++k) cs[k] = 0.0f;

    // One row buffer: a second one (loading row k+1 during the reduction of
    // row k) measured slower; the 16 waves' loads keep HBM busy.  `reverse`
    // walks the block's rows last to first (the column sums then add in
    // that order).
    for (int q = 0; q < rows; ++q) {
        const int k = reverse ? rows - 1 - q : q;
        const int i = static_cast<int>(blockIdx.x) + k * nblk;
        float4 a[G];
        fused_load_row<T, G, FULL, KEEP>(a, C + i * ld, n);
        if (!fused_row_shift<T, G, FULL>(a, cs, i, q & 1, n, scale2, sv, red, u2, su[k], mom) && tid == 0)
            queued[nqueued++] = i;
    }

    // Rows whose mass left the shifted range, with the exact max shift.
    __syncthreads();
    const int nq = nqueued;
    for (int q = 0; q < nq; ++q) {
        const int i = queued[q];
        float4 a[G];
        fused_load_row<T, G>(a, C + i * ld, n);
        fused_row<T, G>(a, cs, i, n, scale2, sv, red, u2, mom);
        __syncthreads();  // queued rows may share a reduction slot
    }

    float* out = part + static_cast<int64_t>(blockIdx.x) * part_ld(n);
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int col = (g * T + tid) * 4;
        if (col < n)
            *reinterpret_cast<float4*>(out + col) = make_float4(cs[4 * g], cs[4 * g + 1], cs[4 * g + 2], cs[4 * g + 3]);
    }
}

// ---------------------------------------------------------------------------
// Column sweep (two-sweep path, and any n): block (bx, by) folds rows
// [bx*rpb, ...) of column chunk by into per-column pairs of K2 + u2_i + v2_j.
// Scalar layout so any n and any row stride work.
// ---------------------------------------------------------------------------
template <int T, int CPT>
__global__ __launch_bounds__(T) void sk_col_sweep(const float* __restrict__ C, int n, int64_t ldc, float scale2,
                                                 const float* __restrict__ v2, const float* __restrict__ u2,
                                                 float2* __restrict__ part, int rows_per_blk,
                                                 const SinkCtl* __restrict__ ctl) {
    if (ctl->done) return;
    const int c0 = blockIdx.y * (T * CPT);
    float vj[CPT], cm[CPT], cs[CPT];
    bool ok[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int col = c0 + k * T + threadIdx.x;
        ok[k] = col < n;
        vj[k] = ok[k] ? v2[col] : 0.0f;
        cm[k] = kNegBig;
        cs[k] = 0.0f;
    }
    const int r0 = blockIdx.x * rows_per_blk;
    const int r1 = min(n, r0 + rows_per_blk);
    for (int i = r0; i < r1; ++i) {
        const float* row = C + static_cast<int64_t>(i) * ldc;
        const float ui = u2[i];
        float c[CPT];
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const int col = c0 + k * T + threadIdx.x;
            c[k] = ok[k] ? ld_stream(row + col) : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < CPT; ++k)
            if (ok[k]) lse_push(cm[k], cs[k], fmaf(c[k], scale2, vj[k]) + ui);
    }
    float2* out = part + static_cast<int64_t>(blockIdx.x) * n;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int col = c0 + k * T + threadIdx.x;
        if (ok[k]) out[col] = make_float2(cm[k], cs[k]);
    }
}

// ---------------------------------------------------------------------------
// Row sweep: one wave per row,
//   out_i = (1-m) out_i - m * oscale * LSE2_j(C_ij*cscale + add_j*ascale).
// The solver calls it in base-2 units (oscale = ascale = 1); the reference-
// compatible single step (cuda_kernels/sinkhorn_log.cu:39-66) calls it with
// K in natural units (cscale = ascale = log2 e, oscale = ln 2).
// ---------------------------------------------------------------------------
template <int W, bool VEC>
__global__ __launch_bounds__(W * kWave) void sk_row_sweep(const float* __restrict__ C, int n, int64_t ldc,
                                                         float cscale, const float* __restrict__ add, float ascale,
                                                         float oscale, float* __restrict__ out, float mom_arg,
                                                         const SinkCtl* __restrict__ ctl) {
    if (ctl && ctl->done) return;
    const float mom = ctl ? ctl->momentum : mom_arg;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * W + threadIdx.x / kWave;
    if (i >= n) return;
    const float* row = C + i * ldc;
    float m = kNegBig, s = 0.0f;
    if (VEC) {
        for (int j = lane * 4; j < n; j += kWave * 4) {
            const float4 c = ld_stream4(row + j);
            const float4 a = *reinterpret_cast<const float4*>(add + j);
            const float x0 = fmaf(c.x, cscale, a.x * ascale);
            const float x1 = fmaf(c.y, cscale, a.y * ascale);
            const float x2 = fmaf(c.z, cscale, a.z * ascale);
            const float x3 = fmaf(c.w, cscale, a.w * ascale);
            const float mn = fmaxf(m, fmaxf(fmaxf(x0, x1), fmaxf(x2, x3)));
            s = s * exp2_fast(m - mn) + ((exp2_fast(x0 - mn) + exp2_fast(x1 - mn)) + (exp2_fast(x2 - mn) + exp2_fast(x3 - mn)));
            m = mn;
        }
    } else {
        for (int j = lane; j < n; j += kWave) lse_push(m, s, fmaf(row[j], cscale, add[j] * ascale));
    }
    const Lse w = wave_lse({m, s});
    if (lane == 0) {
        const float lse = (w.m + log2_fast(w.s)) * oscale;
        out[i] = (mom == 1.0f) ? -lse : (1.0f - mom) * out[i] - mom * lse;
    }
}

// Lehmann over-relaxation, cugal/sinkhorn.py:78-88, evaluated on the device
// after `nerr` errors have been recorded.  Python's negative indexing of the
// error list is kept.
__device__ double lehmann_momentum(const double* errors, int nerr, int eval_freq, int momentum_start) {
    if (momentum_start < 0 || nerr * eval_freq < momentum_start) return 1.0;
    const int index = momentum_start / eval_freq;
    if (nerr < index - 1) return 1.0;
    const int a = index - 1 >= 0 ? index - 1 : nerr + index - 1;
    const int b = index - 2 >= 0 ? index - 2 : nerr + index - 2;
    if (a < 0 || b < 0 || a >= nerr || b >= nerr) return 1.0;
    double ratio = errors[a] / errors[b];
    if (isnan(ratio)) ratio = 0.99;
    ratio = fmin(ratio, 0.99);
    const double power = 1.0 / eval_freq;
    return 2.0 / (1.0 + sqrt(1.0 - pow(ratio, power)));
}

// ---------------------------------------------------------------------------
// Combine: merge the per-block column partials into
//   L_j = log2 sum_i plan_ij   (plan of the current (u, v)),
// then
//   v2_next_j = v2_j - L_j                       (momentum 1), or
//   raw LSE2_i(K2_ij + u2_i) = L_j - v2_j         (momentum active: the blend
//                                                  runs after the last block
//                                                  has updated m),
// and the column marginal is 2^L_j.  PAIRS: (max, sum) partials of the exact
// column sweep; otherwise plain sums of the fused sweep.  A block covers 64
// columns; its 4 waves split the partial list.  With want_err the blocks'
// FP64 error sums are totalled in fixed order by the last block to arrive.
// ---------------------------------------------------------------------------
constexpr int kCombCols = 64;
constexpr int kCombWaves = 4;
// Waves per combine block: 16 where the columns make few blocks (n <= 2048,
// at most 32 blocks), so more partial rows are in flight (n = 1174: 18.0 ->
// 14.5 us per LOG iteration); 4 above, where 16 measured equal or slower
// (32768: 637 against 630 us).
constexpr int kCombWavesNarrow = 16;
constexpr int64_t kCombNarrowMaxN = 2048;

// The error bookkeeping shared by the combines: wave 0 of each block holds
// the per-column error terms; the blocks' FP64 sums are totalled in fixed
// order by the last block to arrive, which records the error, raises `done`
// below the threshold and otherwise updates the Lehmann factor.  It also
// stores the stop decision into `host_flag` (pinned host memory, written at
// system scope) when given, so the host polls a flag instead of copying one.
template <int W>
__device__ void combine_error(double err, int want_err, int check_index, int cur_buf, double threshold,
                              int eval_freq, int momentum_start, double* __restrict__ errpart,
                              SinkCtl* __restrict__ ctl, int* __restrict__ host_flag) {
    __shared__ double dred[W];
    __shared__ int is_last;
    const int lane = threadIdx.x & (kCombCols - 1);
    const int wave = threadIdx.x / kCombCols;
    if (!want_err) return;
    if (wave == 0) {
        err = wave_sum(err);
        if (lane == 0) dred[0] = err;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        errpart[blockIdx.x] = dred[0];
        __threadfence();
        const unsigned t = atomicAdd(&ctl->ticket, 1u);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double total = 0.0;
    for (int b = threadIdx.x; b < static_cast<int>(gridDim.x); b += blockDim.x) total += errpart[b];
    // Fixed-shape reduction over the block: deterministic.
    total = wave_sum(total);
    if ((threadIdx.x & 63) == 0) dred[threadIdx.x / 64] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int w = 0; w < W; ++w) sum += dred[w];
        double* errors = ctl_errors(ctl);
        errors[ctl->nerr] = sum;
        ctl->nerr += 1;
        ctl->ticket = 0u;
        if (sum < threshold) {
            ctl->done = 1;
            ctl->iterations = check_index + 1;
            ctl->final_buf = cur_buf;
        } else {
            ctl->momentum_d = lehmann_momentum(errors, ctl->nerr, eval_freq, momentum_start);
            ctl->momentum = static_cast<float>(ctl->momentum_d);
        }
        if (host_flag) __hip_atomic_store(host_flag, sum < threshold ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

template <bool PAIRS, int W>
__global__ __launch_bounds__(kCombCols* W) void sk_combine(
    const void* __restrict__ part_raw, int nblk, int n, int64_t ldq, const float* __restrict__ v2_in,
    float* __restrict__ v2_out,
    int raw_out, int want_err, int check_index, int cur_buf, double threshold, int eval_freq, int momentum_start,
    double* __restrict__ errpart, SinkCtl* __restrict__ ctl, int* __restrict__ host_flag) {
    __shared__ float rm[W][kCombCols];
    __shared__ double rs[W][kCombCols];
    if (ctl->done) return;
    const int lane = threadIdx.x & (kCombCols - 1);
    const int wave = threadIdx.x / kCombCols;
    const int j = blockIdx.x * kCombCols + lane;
    // FP64 accumulation: the column sums feed the convergence test.
    float m = 0.0f;
    double sum = 0.0;
    if (PAIRS) {
        const float2* part = static_cast<const float2*>(part_raw);
        m = kNegBig;
        if (j < n) {
#pragma unroll 8
            for (int b = wave; b < nblk; b += W) m = fmaxf(m, part[static_cast<int64_t>(b) * ldq + j].x);
#pragma unroll 8
            for (int b = wave; b < nblk; b += W) {
                const float2 p = part[static_cast<int64_t>(b) * ldq + j];
                sum += static_cast<double>(p.y) * exp2(static_cast<double>(p.x) - static_cast<double>(m));
            }
        }
    } else {
        const float* part = static_cast<const float*>(part_raw);
        // 32 loads in flight per thread ahead of the in-order FP64 adds (the
        // order of the sums, and so the result, is that of any unroll):
        // 8.6 -> 6.5 us per combine at n = 8192.
        if (j < n) {
#pragma unroll 32
            for (int b = wave; b < nblk; b += W) sum += static_cast<double>(part[static_cast<int64_t>(b) * ldq + j]);
        }
    }
    rm[wave][lane] = m;
    rs[wave][lane] = sum;
    __syncthreads();
    double err = 0.0;
    if (wave == 0 && j < n) {
        double L;
        if (PAIRS) {
            float M = kNegBig;
            for (int w = 0; w < W; ++w) M = fmaxf(M, rm[w][lane]);
            double S = 0.0;
            for (int w = 0; w < W; ++w)
                S += rs[w][lane] * exp2(static_cast<double>(rm[w][lane]) - static_cast<double>(M));
            L = static_cast<double>(M) + log2(S);
        } else {
            double S = 0.0;
            for (int w = 0; w < W; ++w) S += rs[w][lane];
            L = log2(S);
        }
        const double vj = static_cast<double>(v2_in[j]);
        v2_out[j] = static_cast<float>(raw_out ? (L - vj) : (vj - L));
        if (want_err) err = fabs(exp2(L) - 1.0);
    }
    combine_error<W>(err, want_err, check_index, cur_buf, threshold, eval_freq, momentum_start, errpart, ctl, host_flag);
}

// v_next = (1 - m) v - m * lse  (only when Lehmann momentum is configured).
__global__ void sk_apply_momentum(const float* __restrict__ v2_in, float* __restrict__ v2_io, int n,
                                  const SinkCtl* __restrict__ ctl) {
    if (ctl->done) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const float m = ctl->momentum;
    const float lse = v2_io[j];
    v2_io[j] = (m == 1.0f) ? -lse : (1.0f - m) * v2_in[j] - m * lse;
}

// ---------------------------------------------------------------------------
// STANDARD Sinkhorn, cugal/sinkhorn.py:122-144: with K = exp(-C/reg),
//   v_j = v_j^(1-m) (1 / (sum_i u_i K_ij + 1e-16))^m,
//   u_i = u_i^(1-m) (1 / (sum_j K_ij v_j + 1e-16))^m,
// and the column-marginal error sum_j |sum_i u_i K_ij v_j - 1| at
// iterations 0, f, 2f, ...  K is recomputed as exp2(C * scale2) where it is
// read (never stored).  The fused sweep is the log one with plain sums: per
// row, s_i = sum_j K_ij v_j gives u_i, and with the row's K still in
// registers the column sums sum_i u_i K_ij of the new u accumulate.
// ---------------------------------------------------------------------------
constexpr float kStdEps = 1e-16f;  // M_EPS, cugal/sinkhorn.py:12

__device__ __forceinline__ float std_blend(float old, float fresh, float mom) {
    return mom == 1.0f ? fresh : powf(old, 1.0f - mom) * powf(fresh, mom);
}

template <int T, int G>
__device__ __forceinline__ void std_fused_row(float4 (&r)[G], float (&cs)[4 * G], int i, int n, float scale2,
                                              const float* __restrict__ sv, float* __restrict__ red,
                                              float* __restrict__ u, float mom) {
    constexpr int W = T / kWave;
    const int tid = threadIdx.x;
    const float u_old = u[i];
    float ls = 0.0f;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int col = (g * T + tid) * 4;
        if (group_live<T, G, false>(g, n)) {
            const float4 vv = *reinterpret_cast<const float4*>(sv + col);
            r[g].x = exp2_fast(r[g].x * scale2);
            r[g].y = exp2_fast(r[g].y * scale2);
            r[g].z = exp2_fast(r[g].z * scale2);
            r[g].w = exp2_fast(r[g].w * scale2);
            ls += (r[g].x * vv.x + r[g].y * vv.y) + (r[g].z * vv.z + r[g].w * vv.w);
        }
    }
    ls = wave_sum(ls);
    float* slot = red + (i & 1) * W;
    if ((tid & (kWave - 1)) == 0) slot[tid / kWave] = ls;
    __syncthreads();
    float S = 0.0f;
#pragma unroll
    for (int k = 0; k < W; ++k) S += slot[k];
    const float ui = std_blend(u_old, 1.0f / (S + kStdEps), mom);
    if (tid == 0) u[i] = ui;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (group_live<T, G, false>(g, n)) {
            cs[4 * g + 0] = fmaf(ui, r[g].x, cs[4 * g + 0]);
            cs[4 * g + 1] = fmaf(ui, r[g].y, cs[4 * g + 1]);
            cs[4 * g + 2] = fmaf(ui, r[g].z, cs[4 * g + 2]);
            cs[4 * g + 3] = fmaf(ui, r[g].w, cs[4 * g + 3]);
        }
    }
}

template <int T, int G, bool KEEP>
__global__ __launch_bounds__(T) void std_fused_sweep(const float* __restrict__ C, int n, float scale2,
                                                    const float* __restrict__ v, float* __restrict__ u,
                                                    float* __restrict__ part, int rows_per_blk,
                                                    const SinkCtl* __restrict__ ctl) {
    extern __shared__ float smem[];
    const int tid = threadIdx.x;
    float* sv = smem;
    float* red = smem + round_up(n, 4);
    const int r0 = blockIdx.x * rows_per_blk;
    const int r1 = min(n, r0 + rows_per_blk);
    const int64_t ld = n;
    // The first row and the thread's slice of v (it reads back only its own
    // columns) are loaded together before the stop flag is read, as in the
    // LOG sweep.
    float4 a[G];
    if (r0 < r1) fused_load_row<T, G, false, KEEP>(a, C + r0 * ld, n);
    float4 vg[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int col = (g * T + tid) * 4;
        // Columns from n up to the group's end hold v = 0: no share of the
        // row sums.
        vg[g] = group_whole<T, G, false>(g, n) ? *reinterpret_cast<const float4*>(v + col) : ld_tail4(v, col, n, 0.0f);
    }
    if (ctl->done) return;
    const float mom = ctl->momentum;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int col = (g * T + tid) * 4;
        if (group_live<T, G, false>(g, n)) *reinterpret_cast<float4*>(sv + col) = vg[g];
    }
    __syncthreads();
    float cs[4 * G];
#pragma unroll
    for (int k = 0; k < 4 * G; ++k) cs[k] = 0.0f;
    if constexpr (4 * G * T <= 24 * 1024) {
        float4 b[G];
        for (int i = r0; i < r1; i += 2) {
            if (i + 1 < r1) fused_load_row<T, G, false, KEEP>(b, C + (i + 1) * ld, n);
            std_fused_row<T, G>(a, cs, i, n, scale2, sv, red, u, mom);
            if (i + 1 >= r1) break;
            if (i + 2 < r1) fused_load_row<T, G, false, KEEP>(a, C + (i + 2) * ld, n);
            std_fused_row<T, G>(b, cs, i + 1, n, scale2, sv, red, u, mom);
        }
    } else {
        for (int i = r0; i < r1; ++i) {
            if (i > r0) fused_load_row<T, G, false, KEEP>(a, C + i * ld, n);
            std_fused_row<T, G>(a, cs, i, n, scale2, sv, red, u, mom);
        }
    }
    float* out = part + static_cast<int64_t>(blockIdx.x) * part_ld(n);
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int col = (g * T + tid) * 4;
        if (col < n)
            *reinterpret_cast<float4*>(out + col) = make_float4(cs[4 * g], cs[4 * g + 1], cs[4 * g + 2], cs[4 * g + 3]);
    }
}

// Column sweep: block (bx, by) sums u_i K_ij over rows [bx*rpb, ...) for
// its column chunk.
template <int T, int CPT>
__global__ __launch_bounds__(T) void std_col_sweep(const float* __restrict__ C, int n, int64_t ldc, float scale2,
                                                  const float* __restrict__ u, float* __restrict__ part,
                                                  int rows_per_blk, const SinkCtl* __restrict__ ctl) {
    if (ctl->done) return;
    const int c0 = blockIdx.y * (T * CPT);
    float cs[CPT];
    bool ok[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        ok[k] = c0 + k * T + static_cast<int>(threadIdx.x) < n;
        cs[k] = 0.0f;
    }
    const int r0 = blockIdx.x * rows_per_blk;
    const int r1 = min(n, r0 + rows_per_blk);
    for (int i = r0; i < r1; ++i) {
        const float* row = C + static_cast<int64_t>(i) * ldc;
        const float ui = u[i];
        float c[CPT];
#pragma unroll
        for (int k = 0; k < CPT; ++k) c[k] = ok[k] ? ld_stream(row + c0 + k * T + threadIdx.x) : 0.0f;
#pragma unroll
        for (int k = 0; k < CPT; ++k)
            if (ok[k]) cs[k] = fmaf(ui, exp2_fast(c[k] * scale2), cs[k]);
    }
    float* out = part + static_cast<int64_t>(blockIdx.x) * n;
#pragma unroll
    for (int k = 0; k < CPT; ++k)
        if (ok[k]) out[c0 + k * T + threadIdx.x] = cs[k];
}

// Row sweep: one wave per row, u_i = blend(u_i, 1 / (sum_j K_ij v_j + eps)).
template <int W, bool VEC>
__global__ __launch_bounds__(W * kWave) void std_row_sweep(const float* __restrict__ C, int n, int64_t ldc,
                                                          float scale2, const float* __restrict__ v,
                                                          float* __restrict__ u, const SinkCtl* __restrict__ ctl) {
    if (ctl->done) return;
    const float mom = ctl->momentum;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * W + threadIdx.x / kWave;
    if (i >= n) return;
    const float* row = C + i * ldc;
    float s = 0.0f;
    if (VEC) {
        for (int j = lane * 4; j < n; j += kWave * 4) {
            const float4 c = ld_stream4(row + j);
            const float4 a = *reinterpret_cast<const float4*>(v + j);
            s += (exp2_fast(c.x * scale2) * a.x + exp2_fast(c.y * scale2) * a.y) +
                 (exp2_fast(c.z * scale2) * a.z + exp2_fast(c.w * scale2) * a.w);
        }
    } else {
        for (int j = lane; j < n; j += kWave) s = fmaf(exp2_fast(row[j] * scale2), v[j], s);
    }
    s = wave_sum(s);
    if (lane == 0) u[i] = std_blend(u[i], 1.0f / (s + kStdEps), mom);
}

// Combine: column sums S_j = sum_i u_i K_ij (FP64 over the partials) give
// the raw next v_j = 1 / (S_j + eps) and the column marginal v_j S_j of the
// current (u, v).
template <int W>
__global__ __launch_bounds__(kCombCols* W) void std_combine(
    const float* __restrict__ part, int nblk, int n, int64_t ldq, const float* __restrict__ v_in,
    float* __restrict__ v_out, int want_err, int check_index, int cur_buf, double threshold, int eval_freq, int momentum_start,
    double* __restrict__ errpart, SinkCtl* __restrict__ ctl, int* __restrict__ host_flag) {
    __shared__ double rs[W][kCombCols];
    if (ctl->done) return;
    const int lane = threadIdx.x & (kCombCols - 1);
    const int wave = threadIdx.x / kCombCols;
    const int j = blockIdx.x * kCombCols + lane;
    double sum = 0.0;
    if (j < n) {
#pragma unroll 32
        for (int b = wave; b < nblk; b += W) sum += static_cast<double>(part[static_cast<int64_t>(b) * ldq + j]);
    }
    rs[wave][lane] = sum;
    __syncthreads();
    double err = 0.0;
    if (wave == 0 && j < n) {
        double S = 0.0;
        for (int w = 0; w < W; ++w) S += rs[w][lane];
        v_out[j] = static_cast<float>(1.0 / (S + static_cast<double>(kStdEps)));
        if (want_err) err = fabs(static_cast<double>(v_in[j]) * S - 1.0);
    }
    combine_error<W>(err, want_err, check_index, cur_buf, threshold, eval_freq, momentum_start, errpart, ctl, host_flag);
}

// v_next = v^(1-m) * raw^m once the check of this combine has set m.
__global__ void std_apply_momentum(const float* __restrict__ v_in, float* __restrict__ v_io, int n,
                                   const SinkCtl* __restrict__ ctl) {
    if (ctl->done) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    v_io[j] = std_blend(v_in[j], v_io[j], ctl->momentum);
}

__global__ void std_init(const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ uw,
                         float* __restrict__ vw, int n, SinkCtl* __restrict__ ctl) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) {
        uw[j] = u[j];
        vw[j] = v[j];
    }
    if (j == 0) {
        ctl->done = 0;
        ctl->iterations = 0;
        ctl->final_buf = 0;
        ctl->nerr = 0;
        ctl->ticket = 0u;
        ctl->momentum = 1.0f;
    ctl->momentum_d = 1.0;
        ctl->momentum_d = 1.0;
    }
}

__global__ void std_finish(const float* __restrict__ uw, const float* __restrict__ va, const float* __restrict__ vb,
                           float* __restrict__ u, float* __restrict__ v, int n, const SinkCtl* __restrict__ ctl,
                           int default_buf) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int buf = ctl->done ? ctl->final_buf : default_buf;
    u[j] = uw[j];
    v[j] = buf ? vb[j] : va[j];
}

// Row marginal error of a materialised kernel, cuda_kernels/sinkhorn_log.cu:68-98:
// out_r = |sum_c exp(K_rc + v_c + u_r) - 1|.
__global__ __launch_bounds__(256) void sk_row_marginal(const float* __restrict__ K, int n, const float* __restrict__ u,
                                                      const float* __restrict__ v, double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / 64;
    if (r >= n) return;
    const float ur = u[r];
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += exp2_fast((K[r * n + c] + v[c] + ur) * kLog2e);
    s = wave_sum(s);
    if (lane == 0) out[r] = fabs(s - 1.0);
}

__global__ void sum_doubles(const double* __restrict__ in, int n, double* __restrict__ out) {
    __shared__ double scratch[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += in[i];
    s = block_sum_double(s, scratch);
    if (threadIdx.x == 0) *out = s;
}

// ---------------------------------------------------------------------------
// STANDARD Sinkhorn in FP64 (the reference's default dtype, cugal/config.py:
// 65, for cugal/sinkhorn.py:122-144): the same iteration on an FP64 cost
// with FP64 scalings, K = exp(-C / reg) evaluated where C is read.  Two
// sweeps per iteration: the row sweep forms u from v, the column sweep the
// per-block column sums of the new u, which the combine turns into the next
// v and the column-marginal error of (u, v), exactly the two-sweep FP32
// structure.
// ---------------------------------------------------------------------------
constexpr double kStdEps64 = 1e-16;  // M_EPS, cugal/sinkhorn.py:12

__device__ __forceinline__ double std_blend64(double old, double fresh, double mom) {
    return mom == 1.0 ? fresh : pow(old, 1.0 - mom) * pow(fresh, mom);
}

template <int T, int CPT>
__global__ __launch_bounds__(T) void std64_col_sweep(const double* __restrict__ C, int n, double nrinv,
                                                    const double* __restrict__ u, double* __restrict__ part,
                                                    int rows_per_blk, const SinkCtl* __restrict__ ctl) {
    if (ctl->done) return;
    const int c0 = blockIdx.y * (T * CPT);
    double cs[CPT];
    bool ok[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        ok[k] = c0 + k * T + static_cast<int>(threadIdx.x) < n;
        cs[k] = 0.0;
    }
    const int r0 = blockIdx.x * rows_per_blk;
    const int r1 = min(n, r0 + rows_per_blk);
    for (int i = r0; i < r1; ++i) {
        const double* row = C + static_cast<int64_t>(i) * n;
        const double ui = u[i];
        double c[CPT];
#pragma unroll
        for (int k = 0; k < CPT; ++k) c[k] = ok[k] ? __builtin_nontemporal_load(row + c0 + k * T + threadIdx.x) : 0.0;
#pragma unroll
        for (int k = 0; k < CPT; ++k)
            if (ok[k]) cs[k] = fma(ui, exp(c[k] * nrinv), cs[k]);
    }
    double* out = part + static_cast<int64_t>(blockIdx.x) * n;
#pragma unroll
    for (int k = 0; k < CPT; ++k)
        if (ok[k]) out[c0 + k * T + threadIdx.x] = cs[k];
}

template <int W>
__global__ __launch_bounds__(W * kWave) void std64_row_sweep(const double* __restrict__ C, int n, double nrinv,
                                                            const double* __restrict__ v, double* __restrict__ u,
                                                            const SinkCtl* __restrict__ ctl) {
    if (ctl->done) return;
    const double mom = ctl->momentum_d;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * W + threadIdx.x / kWave;
    if (i >= n) return;
    const double* row = C + i * n;
    double s = 0.0;
    for (int j = lane; j < n; j += kWave) s = fma(exp(__builtin_nontemporal_load(row + j) * nrinv), v[j], s);
    s = wave_sum(s);
    if (lane == 0) u[i] = std_blend64(u[i], 1.0 / (s + kStdEps64), mom);
}

template <int W>
__global__ __launch_bounds__(kCombCols* W) void std64_combine(
    const double* __restrict__ part, int nblk, int n, const double* __restrict__ v_in, double* __restrict__ v_out,
    int want_err, int check_index, int cur_buf, double threshold, int eval_freq, int momentum_start,
    double* __restrict__ errpart, SinkCtl* __restrict__ ctl, int* __restrict__ host_flag) {
    __shared__ double rs[W][kCombCols];
    if (ctl->done) return;
    const int lane = threadIdx.x & (kCombCols - 1);
    const int wave = threadIdx.x / kCombCols;
    const int j = blockIdx.x * kCombCols + lane;
    double sum = 0.0;
    if (j < n) {
#pragma unroll 32
        for (int b = wave; b < nblk; b += W) sum += part[static_cast<int64_t>(b) * n + j];
    }
    rs[wave][lane] = sum;
    __syncthreads();
    double err = 0.0;
    if (wave == 0 && j < n) {
        double S = 0.0;
        for (int w = 0; w < W; ++w) S += rs[w][lane];
        v_out[j] = 1.0 / (S + kStdEps64);
        if (want_err) err = fabs(v_in[j] * S - 1.0);
    }
    combine_error<W>(err, want_err, check_index, cur_buf, threshold, eval_freq, momentum_start, errpart, ctl, host_flag);
}

__global__ void std64_apply_momentum(const double* __restrict__ v_in, double* __restrict__ v_io, int n,
                                     const SinkCtl* __restrict__ ctl) {
    if (ctl->done) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    v_io[j] = std_blend64(v_in[j], v_io[j], ctl->momentum_d);
}

__global__ void std64_init(const double* __restrict__ u, const double* __restrict__ v, double* __restrict__ uw,
                           double* __restrict__ vw, int n, SinkCtl* __restrict__ ctl) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) {
        uw[j] = u[j];
        vw[j] = v[j];
    }
    if (j == 0) {
        ctl->done = 0;
        ctl->iterations = 0;
        ctl->final_buf = 0;
        ctl->nerr = 0;
        ctl->ticket = 0u;
        ctl->momentum = 1.0f;
        ctl->momentum_d = 1.0;
    }
}

__global__ void std64_finish(const double* __restrict__ uw, const double* __restrict__ va,
                             const double* __restrict__ vb, double* __restrict__ u, double* __restrict__ v, int n,
                             const SinkCtl* __restrict__ ctl, int default_buf) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int buf = ctl->done ? ctl->final_buf : default_buf;
    u[j] = uw[j];
    v[j] = buf ? vb[j] : va[j];
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
constexpr int kFusedMaxGroups = 8;  // 1024 threads x 8 float4 groups = 32768 columns
constexpr int64_t kFusedMaxN = 1024 * 4 * kFusedMaxGroups;
constexpr int kColThreads = 256;
constexpr int kColCPT = 8;
constexpr int kRowWaves = 4;

struct Plan {
    int fused;
    int threads;       // fused: block size
    int groups;        // fused: float4 groups per thread
    int rows_per_blk;  // fused: rows per block
    int nblk;          // fused: row blocks (= partial rows)
    int col_rows;      // column sweep: rows per block
    int col_nblk;      // column sweep: row blocks
    int col_chunks;    // column sweep: chunks of kColThreads*kColCPT columns
    int comb_blocks;
};

Plan make_plan(int64_t n, int fused) {
    Plan p{};
    p.fused = fused;
    if (fused) {
        p.threads = n > 8192 ? 1024 : 256;
        const int64_t g = div_up(n, static_cast<int64_t>(p.threads) * 4);
        p.groups = static_cast<int>(std::min<int64_t>(8, std::max<int64_t>(1, g)));
        // 1024-thread blocks: one per CU (256 CUs); 256-thread blocks: two,
        // or one up to n = 2048, where the combine's read of the column sums
        // (4n bytes per block, 1/rpb of C) weighs more than the rows
        // (measured per LOG iteration: 15.6 -> 13.3 us at n = 1000, 17.7 ->
        // 16.1 us at 2048; 25.9 -> 27.4 us at 4096, so not beyond).
        const int64_t target = p.threads == 1024 ? 256 : (n <= 2048 ? 256 : 512);
        p.rows_per_blk = static_cast<int>(std::max<int64_t>(1, div_up(n, target)));
        p.nblk = static_cast<int>(div_up(n, p.rows_per_blk));
    }
    p.col_chunks = static_cast<int>(div_up(n, static_cast<int64_t>(kColThreads) * kColCPT));
    const int64_t col_target = std::max<int64_t>(1, 1024 / p.col_chunks);
    p.col_rows = static_cast<int>(std::max<int64_t>(8, div_up(n, col_target)));
    p.col_nblk = static_cast<int>(div_up(n, p.col_rows));
    p.comb_blocks = static_cast<int>(div_up(n, kCombCols));
    return p;
}

// The combines with the wave count their width calls for.
template <bool PAIRS, typename... Args>
void launch_sk_combine(int64_t n, int blocks, hipStream_t s, Args... args) {
    if (n <= kCombNarrowMaxN)
        hipLaunchKernelGGL((sk_combine<PAIRS, kCombWavesNarrow>), dim3(blocks), dim3(kCombCols * kCombWavesNarrow), 0,
                           s, args...);
    else
        hipLaunchKernelGGL((sk_combine<PAIRS, kCombWaves>), dim3(blocks), dim3(kCombCols * kCombWaves), 0, s, args...);
}
template <typename... Args>
void launch_std_combine(int64_t n, int blocks, hipStream_t s, Args... args) {
    if (n <= kCombNarrowMaxN)
        hipLaunchKernelGGL((std_combine<kCombWavesNarrow>), dim3(blocks), dim3(kCombCols * kCombWavesNarrow), 0, s,
                           args...);
    else
        hipLaunchKernelGGL((std_combine<kCombWaves>), dim3(blocks), dim3(kCombCols * kCombWaves), 0, s, args...);
}
template <typename... Args>
void launch_std64_combine(int64_t n, int blocks, hipStream_t s, Args... args) {
    if (n <= kCombNarrowMaxN)
        hipLaunchKernelGGL((std64_combine<kCombWavesNarrow>), dim3(blocks), dim3(kCombCols * kCombWavesNarrow), 0, s,
                           args...);
    else
        hipLaunchKernelGGL((std64_combine<kCombWaves>), dim3(blocks), dim3(kCombCols * kCombWaves), 0, s, args...);
}

size_t ctl_bytes(int err_capacity) { return sizeof(SinkCtl) + sizeof(double) * static_cast<size_t>(err_capacity); }

int err_capacity(int max_iterations, int eval_freq) { return max_iterations / std::max(1, eval_freq) + 4; }

struct SinkWs {
    float* u2;
    float* v2[2];
    void* part;
    double* errpart;
    SinkCtl* ctl;
};

size_t carve(Carver& cv, int64_t n, const Plan& p, int cap, SinkWs* w) {
    SinkWs t{};
    t.u2 = cv.take<float>(n);
    t.v2[0] = cv.take<float>(n);
    t.v2[1] = cv.take<float>(n);
    const size_t pair_bytes = sizeof(float2) * static_cast<size_t>(p.col_nblk) * n;
    const size_t sum_bytes = p.fused ? sizeof(float) * static_cast<size_t>(p.nblk) * part_ld(n) : 0;
    t.part = cv.take<char>(std::max(pair_bytes, sum_bytes));
    t.errpart = cv.take<double>(p.comb_blocks);
    t.ctl = reinterpret_cast<SinkCtl*>(cv.take<char>(ctl_bytes(cap)));
    if (w) *w = t;
    return cv.used;
}

int max_lds_bytes() {
    static int cached = -1;
    if (cached < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            v = 65536;
        cached = v;
    }
    return cached;
}

template <int T, int G, bool FULL, bool KEEP>
hipError_t set_fused_lds(size_t lds) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&sk_fused_sweep<T, G, FULL, KEEP>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
}

// `reverse`: the iteration's parity; the row order alternates only where
// the matrix is read with caching loads.
template <int T, int G>
int launch_fused_tg(const float* C, int n, float scale2, const float* v2, float* u2, float* part, const Plan& p,
                    const SinkCtl* ctl, hipStream_t s, int reverse) {
    if (div_up(n, 4 * T) != G) return fail(CG_EINVAL, "fused Sinkhorn sweep: %d groups of %d threads for n=%d", G, T, n);
    const size_t lds = sizeof(float) * round_up(n, 4) + sizeof(float2) * 2 * (T / kWave);
    // The kernel also holds a small static LDS queue: ask for exactly the
    // dynamic bytes this launch needs (the attribute only grows).
    static size_t attr_bytes = 0;
    if (lds + 1024 > static_cast<size_t>(max_lds_bytes()))
        return fail(CG_EUNSUPPORTED, "fused Sinkhorn sweep needs %zu bytes of LDS (device has %d)", lds,
                    max_lds_bytes());
    if (lds > attr_bytes) {
        CG_CHECK_HIP((set_fused_lds<T, G, true, false>(lds)));
        CG_CHECK_HIP((set_fused_lds<T, G, false, false>(lds)));
        CG_CHECK_HIP((set_fused_lds<T, G, true, true>(lds)));
        CG_CHECK_HIP((set_fused_lds<T, G, false, true>(lds)));
        attr_bytes = lds;
    }
    const bool keep = n <= kCacheableMaxN;
    const int rev = keep ? (reverse & 1) : 0;
    const bool full = n == 4 * G * T;
    const dim3 grid(p.nblk), block(T);
    if (full && keep)
        hipLaunchKernelGGL((sk_fused_sweep<T, G, true, true>), grid, block, lds, s, C, n, scale2, v2, u2, part,
                           p.rows_per_blk, rev, ctl);
    else if (full)
        hipLaunchKernelGGL((sk_fused_sweep<T, G, true, false>), grid, block, lds, s, C, n, scale2, v2, u2, part,
                           p.rows_per_blk, rev, ctl);
    else if (keep)
        hipLaunchKernelGGL((sk_fused_sweep<T, G, false, true>), grid, block, lds, s, C, n, scale2, v2, u2, part,
                           p.rows_per_blk, rev, ctl);
    else
        hipLaunchKernelGGL((sk_fused_sweep<T, G, false, false>), grid, block, lds, s, C, n, scale2, v2, u2, part,
                           p.rows_per_blk, rev, ctl);
    return check_launch("sk_fused_sweep");
}

int launch_fused(const float* C, int n, float scale2, const float* v2, float* u2, float* part, const Plan& p,
                 const SinkCtl* ctl, hipStream_t s, int reverse) {
    if (p.threads == 256) {
        switch (p.groups) {
            case 1: return launch_fused_tg<256, 1>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
            case 2: return launch_fused_tg<256, 2>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
            case 3: return launch_fused_tg<256, 3>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
            case 4: return launch_fused_tg<256, 4>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
            case 5: return launch_fused_tg<256, 5>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
            case 6: return launch_fused_tg<256, 6>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
            case 7: return launch_fused_tg<256, 7>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
            default: return launch_fused_tg<256, 8>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
        }
    }
    switch (p.groups) {
        case 1: return launch_fused_tg<1024, 1>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
        case 2: return launch_fused_tg<1024, 2>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
        case 3: return launch_fused_tg<1024, 3>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
        case 4: return launch_fused_tg<1024, 4>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
        case 5: return launch_fused_tg<1024, 5>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
        case 6: return launch_fused_tg<1024, 6>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
        case 7: return launch_fused_tg<1024, 7>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
        default: return launch_fused_tg<1024, 8>(C, n, scale2, v2, u2, part, p, ctl, s, reverse);
    }
}

int launch_col(const float* C, int n, int64_t ldc, float scale2, const float* v2, const float* u2, float2* part,
               const Plan& p, const SinkCtl* ctl, hipStream_t s) {
    hipLaunchKernelGGL((sk_col_sweep<kColThreads, kColCPT>), dim3(p.col_nblk, p.col_chunks), dim3(kColThreads), 0, s,
                       C, n, ldc, scale2, v2, u2, part, p.col_rows, ctl);
    return check_launch("sk_col_sweep");
}

int launch_row(const float* C, int n, int64_t ldc, float cscale, const float* add, float ascale, float oscale,
               float* out, float mom, const SinkCtl* ctl, hipStream_t s) {
    const bool vec = (n % 4 == 0) && (ldc % 4 == 0) && aligned16(C) && aligned16(add);
    const dim3 grid(static_cast<unsigned>(div_up(n, kRowWaves)));
    if (vec)
        hipLaunchKernelGGL((sk_row_sweep<kRowWaves, true>), grid, dim3(kRowWaves * kWave), 0, s, C, n, ldc, cscale,
                           add, ascale, oscale, out, mom, ctl);
    else
        hipLaunchKernelGGL((sk_row_sweep<kRowWaves, false>), grid, dim3(kRowWaves * kWave), 0, s, C, n, ldc, cscale,
                           add, ascale, oscale, out, mom, ctl);
    return check_launch("sk_row_sweep");
}

template <int T, int G>
int launch_std_fused_tg(const float* C, int n, float scale2, const float* v, float* u, float* part, const Plan& p,
                        const SinkCtl* ctl, hipStream_t s) {
    if (div_up(n, 4 * T) != G) return fail(CG_EINVAL, "fused Sinkhorn sweep: %d groups of %d threads for n=%d", G, T, n);
    const size_t lds = sizeof(float) * round_up(n, 4) + sizeof(float) * 2 * (T / kWave);
    static size_t attr_bytes = 0;
    if (lds > static_cast<size_t>(max_lds_bytes()))
        return fail(CG_EUNSUPPORTED, "fused Sinkhorn sweep needs %zu bytes of LDS (device has %d)", lds,
                    max_lds_bytes());
    if (lds > attr_bytes) {
        CG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&std_fused_sweep<T, G, false>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        CG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&std_fused_sweep<T, G, true>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        attr_bytes = lds;
    }
    if (n <= kStdCacheableMaxN)
        hipLaunchKernelGGL((std_fused_sweep<T, G, true>), dim3(p.nblk), dim3(T), lds, s, C, n, scale2, v, u, part,
                           p.rows_per_blk, ctl);
    else
        hipLaunchKernelGGL((std_fused_sweep<T, G, false>), dim3(p.nblk), dim3(T), lds, s, C, n, scale2, v, u, part,
                           p.rows_per_blk, ctl);
    return check_launch("std_fused_sweep");
}

int launch_std_fused(const float* C, int n, float scale2, const float* v, float* u, float* part, const Plan& p,
                     const SinkCtl* ctl, hipStream_t s) {
    if (p.threads == 256) {
        switch (p.groups) {
            case 1: return launch_std_fused_tg<256, 1>(C, n, scale2, v, u, part, p, ctl, s);
            case 2: return launch_std_fused_tg<256, 2>(C, n, scale2, v, u, part, p, ctl, s);
            case 3: return launch_std_fused_tg<256, 3>(C, n, scale2, v, u, part, p, ctl, s);
            case 4: return launch_std_fused_tg<256, 4>(C, n, scale2, v, u, part, p, ctl, s);
            case 5: return launch_std_fused_tg<256, 5>(C, n, scale2, v, u, part, p, ctl, s);
            case 6: return launch_std_fused_tg<256, 6>(C, n, scale2, v, u, part, p, ctl, s);
            case 7: return launch_std_fused_tg<256, 7>(C, n, scale2, v, u, part, p, ctl, s);
            default: return launch_std_fused_tg<256, 8>(C, n, scale2, v, u, part, p, ctl, s);
        }
    }
    switch (p.groups) {
        case 1: return launch_std_fused_tg<1024, 1>(C, n, scale2, v, u, part, p, ctl, s);
        case 2: return launch_std_fused_tg<1024, 2>(C, n, scale2, v, u, part, p, ctl, s);
        case 3: return launch_std_fused_tg<1024, 3>(C, n, scale2, v, u, part, p, ctl, s);
        case 4: return launch_std_fused_tg<1024, 4>(C, n, scale2, v, u, part, p, ctl, s);
        case 5: return launch_std_fused_tg<1024, 5>(C, n, scale2, v, u, part, p, ctl, s);
        case 6: return launch_std_fused_tg<1024, 6>(C, n, scale2, v, u, part, p, ctl, s);
        case 7: return launch_std_fused_tg<1024, 7>(C, n, scale2, v, u, part, p, ctl, s);
        default: return launch_std_fused_tg<1024, 8>(C, n, scale2, v, u, part, p, ctl, s);
    }
}

int launch_std_col(const float* C, int n, float scale2, const float* u, float* part, const Plan& p,
                   const SinkCtl* ctl, hipStream_t s) {
    hipLaunchKernelGGL((std_col_sweep<kColThreads, kColCPT>), dim3(p.col_nblk, p.col_chunks), dim3(kColThreads), 0,
                       s, C, n, static_cast<int64_t>(n), scale2, u, part, p.col_rows, ctl);
    return check_launch("std_col_sweep");
}

int launch_std_row(const float* C, int n, float scale2, const float* v, float* u, const SinkCtl* ctl,
                   hipStream_t s) {
    const bool vec = (n % 4 == 0) && aligned16(C) && aligned16(v);
    const dim3 grid(static_cast<unsigned>(div_up(n, kRowWaves)));
    if (vec)
        hipLaunchKernelGGL((std_row_sweep<kRowWaves, true>), grid, dim3(kRowWaves * kWave), 0, s, C, n,
                           static_cast<int64_t>(n), scale2, v, u, ctl);
    else
        hipLaunchKernelGGL((std_row_sweep<kRowWaves, false>), grid, dim3(kRowWaves * kWave), 0, s, C, n,
                           static_cast<int64_t>(n), scale2, v, u, ctl);
    return check_launch("std_row_sweep");
}

// Pinned host mirror of the control block, grown on demand, and the pinned
// slots the checking combines write their stop decisions into.  Both are
// per host thread, so solves issued from different threads (on different
// streams) do not share them; they are kept until the process ends.
struct HostMirror {
    void* ptr = nullptr;
    size_t bytes = 0;
};
thread_local HostMirror g_mirror;
thread_local int* g_done_slots = nullptr;  // pinned poll slots

int ensure_pinned(size_t bytes) {
    if (!g_done_slots) CG_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&g_done_slots), 64 * sizeof(int), hipHostMallocDefault));
    if (g_mirror.bytes >= bytes) return 0;
    if (g_mirror.ptr) CG_CHECK_HIP(hipHostFree(g_mirror.ptr));
    g_mirror.ptr = nullptr;
    g_mirror.bytes = 0;
    CG_CHECK_HIP(hipHostMalloc(&g_mirror.ptr, bytes, hipHostMallocDefault));
    g_mirror.bytes = bytes;
    return 0;
}

// Optional timing of the sweep kernels (bench roofline).
struct SweepTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    int device = -1;
};
thread_local SweepTimer g_timer;  // per host thread, like the poll state

int timer_begin(bool on) {
    g_timer.on = on;
    g_timer.used = 0;
    if (!on) return 0;
    int dev = 0;
    CG_CHECK_HIP(hipGetDevice(&dev));
    if (dev != g_timer.device) {
        for (auto e : g_timer.ev) (void)hipEventDestroy(e);
        g_timer.ev.clear();
        g_timer.device = dev;
    }
    return 0;
}

int timer_mark(hipStream_t s) {
    if (!g_timer.on) return 0;
    if (g_timer.used == g_timer.ev.size()) {
        hipEvent_t e;
        CG_CHECK_HIP(hipEventCreate(&e));
        g_timer.ev.push_back(e);
    }
    CG_CHECK_HIP(hipEventRecord(g_timer.ev[g_timer.used++], s));
    return 0;
}

// The timer samples one iteration in kTimerEvery (phase kTimerPhase, never a
// checking iteration when the check period is kTimerEvery).
constexpr int kTimerEvery = 10;
constexpr int kTimerPhase = 5;
bool timer_sampled(int it) { return it % kTimerEvery == kTimerPhase; }
int timer_samples_before(int iterations) {
    return iterations > kTimerPhase ? (iterations - kTimerPhase + kTimerEvery - 1) / kTimerEvery : 0;
}

// Sum of (end - start) over the first `pairs` recorded pairs (all when
// negative); call after the stream is synced.  Launches enqueued after the
// stop exit at once and are left out by the caller's count.
int timer_total(float* ms, int* launches, int64_t pairs) {
    float total = 0.0f;
    int count = 0;
    for (size_t k = 0; k + 1 < g_timer.used && (pairs < 0 || count < pairs); k += 2) {
        float t = 0.0f;
        CG_CHECK_HIP(hipEventElapsedTime(&t, g_timer.ev[k], g_timer.ev[k + 1]));
        total += t;
        ++count;
    }
    if (ms) *ms = total;
    if (launches) *launches = count;
    return 0;
}

__global__ void sk_ctl_reset(SinkCtl* __restrict__ ctl) {
    ctl->done = 0;
    ctl->iterations = 0;
    ctl->final_buf = 0;
    ctl->nerr = 0;
    ctl->ticket = 0u;
    ctl->momentum = 1.0f;
    ctl->momentum_d = 1.0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Unrolled segment (used by the differentiable Sinkhorn): `steps` plain
// iterations from a checkpoint, recording every potential in base-2 units.
// u_hist[0] = u_c, v_hist[0] = v_c (copied); u_hist[s], v_hist[s] are the
// potentials after iteration s.  The first v-update is the exact column
// sweep; later ones come from the fused sweep when it applies, so a segment
// recomputed from the same checkpoint reproduces the same bits.
// ---------------------------------------------------------------------------
size_t sinkhorn_segment_workspace(int64_t n) {
    const Plan p = make_plan(n, cg_sinkhorn_fused_supported(n));
    Carver cv{nullptr, 0, ~size_t(0)};
    return carve(cv, n, p, 4, nullptr) + 256;
}

int sinkhorn_segment(const float* C, int n, float scale2, const float* u_c, const float* v_c, int steps,
                     float* u_hist, float* v_hist, void* ws, size_t ws_bytes, hipStream_t s) {
    const int use_fused = cg_sinkhorn_fused_supported(n) && aligned16(C);
    const Plan p = make_plan(n, use_fused);
    Carver cv{static_cast<char*>(ws), 0, ws_bytes};
    SinkWs w{};
    carve(cv, n, p, 4, &w);
    if (!ws || !cv.ok()) return fail(CG_EWORKSPACE, "sinkhorn segment: workspace too small");
    const int64_t nn = n;
    CG_CHECK_HIP(hipMemcpyAsync(u_hist, u_c, sizeof(float) * nn, hipMemcpyDeviceToDevice, s));
    CG_CHECK_HIP(hipMemcpyAsync(v_hist, v_c, sizeof(float) * nn, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(sk_ctl_reset, dim3(1), dim3(1), 0, s, w.ctl);
    if (int rc = check_launch("sk_ctl_reset")) return rc;
    auto comb = [&](bool pairs, const float* vin, float* vout) -> int {
        if (pairs)
            launch_sk_combine<true>(n, p.comb_blocks, s, w.part,
                               p.col_nblk, n, static_cast<int64_t>(n), vin, vout, 0, 0, -1, 0, 0.0, 1, -1, w.errpart,
                               w.ctl, nullptr);
        else
            launch_sk_combine<false>(n, p.comb_blocks, s, w.part,
                               p.nblk, n, part_ld(n), vin, vout, 0, 0, -1, 0, 0.0, 1, -1, w.errpart, w.ctl, nullptr);
        return check_launch("sk_combine");
    };
    if (steps <= 0) return CG_OK;
    if (int rc = launch_col(C, n, n, scale2, v_hist, u_hist, static_cast<float2*>(w.part), p, w.ctl, s)) return rc;
    if (int rc = comb(true, v_hist, v_hist + nn)) return rc;
    for (int q = 1; q <= steps; ++q) {
        float* uq = u_hist + q * nn;
        float* vq = v_hist + q * nn;
        if (use_fused) {
            if (int rc = launch_fused(C, n, scale2, vq, uq, static_cast<float*>(w.part), p, w.ctl, s, q & 1))
                return rc;
            if (q < steps)
                if (int rc = comb(false, vq, vq + nn)) return rc;
        } else {
            if (int rc = launch_row(C, n, n, scale2, vq, 1.0f, 1.0f, uq, 1.0f, w.ctl, s)) return rc;
            if (q < steps) {
                if (int rc = launch_col(C, n, n, scale2, vq, uq, static_cast<float2*>(w.part), p, w.ctl, s))
                    return rc;
                if (int rc = comb(true, vq, vq + nn)) return rc;
            }
        }
    }
    return CG_OK;
}

}  // namespace cg

using namespace cg;

extern "C" {

size_t cg_sinkhorn_workspace(int64_t n, int fused, int max_iterations, int eval_freq) {
    if (n <= 0) return 0;
    const int use_fused = fused && cg_sinkhorn_fused_supported(n);
    const Plan p = make_plan(n, use_fused);
    Carver cv{nullptr, 0, ~size_t(0)};
    return carve(cv, n, p, err_capacity(max_iterations, eval_freq), nullptr) + 256;
}

int cg_sinkhorn_fused_supported(int64_t n) {
    // n >= 4: a row's last column group loads the row's last 4 floats.
    return n >= 4 && n <= kFusedMaxN ? 1 : 0;
}

int cg_sinkhorn_log_solve(const float* C, int64_t n, float regularization, float* u, float* v, int max_iterations,
                          double threshold, int eval_freq, int momentum_start, int fused, double* errors,
                          int errors_capacity, int* error_count, int* iterations, float* sweep_ms,
                          int* sweep_launches, void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!C || !u || !v || n <= 0 || n > (int64_t(1) << 30))
        return fail(CG_EINVAL, "cg_sinkhorn_log_solve: bad arguments (n=%lld)", static_cast<long long>(n));
    if (!(regularization > 0.0f)) return fail(CG_EINVAL, "sinkhorn regularization must be > 0");
    if (eval_freq <= 0) return fail(CG_EINVAL, "sinkhorn eval_freq must be > 0");
    if (max_iterations < 0) return fail(CG_EINVAL, "sinkhorn iterations must be >= 0");
    const int ni = static_cast<int>(n);
    const int use_fused = fused && cg_sinkhorn_fused_supported(n);
    if (use_fused && !aligned16(C)) {
        // The fused sweep loads float4 groups from a 16-byte aligned matrix
        // (rows of any width): fall back to the two-sweep structure for an
        // unaligned one (same results).
        return cg_sinkhorn_log_solve(C, n, regularization, u, v, max_iterations, threshold, eval_freq,
                                     momentum_start, 0, errors, errors_capacity, error_count, iterations, sweep_ms,
                                     sweep_launches, workspace, workspace_bytes, stream);
    }
    const Plan p = make_plan(n, use_fused);
    const int cap = err_capacity(max_iterations, eval_freq);
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    SinkWs w{};
    carve(cv, n, p, cap, &w);
    if (!workspace || !cv.ok())
        return fail(CG_EWORKSPACE, "cg_sinkhorn_log_solve: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    if (int rc = ensure_pinned(ctl_bytes(cap))) return rc;
    if (int rc = timer_begin(sweep_ms != nullptr || sweep_launches != nullptr)) return rc;

    const float scale2 = -kLog2e / regularization;
    const int64_t ldc = n;
    const bool momentum = momentum_start >= 0;
    const dim3 vgrid(static_cast<unsigned>(div_up(n, 256)));

    hipLaunchKernelGGL(sk_init, vgrid, dim3(256), 0, s, u, v, w.u2, w.v2[0], ni, w.ctl);
    if (int rc = check_launch("sk_init")) return rc;

    // Poll slots: one per outstanding check, host waits one period behind.
    struct Pending {
        hipEvent_t ev;
        int slot;
    };
    std::vector<Pending> pending;
    int next_slot = 0;
    bool stopped = false;
    // A checking combine stores its stop decision into a cleared pinned slot;
    // the host reads the slot once the event after the combine has passed.
    int check_slot = 0;
    auto check_flag = [&](int want_err) -> int* {
        if (!want_err) return nullptr;
        check_slot = next_slot++ & 63;
        g_done_slots[check_slot] = 0;
        return g_done_slots + check_slot;
    };
    auto post_check = [&]() -> int {
        hipEvent_t ev;
        CG_CHECK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        const int slot = check_slot;
        CG_CHECK_HIP(hipEventRecord(ev, s));
        pending.push_back({ev, slot});
        if (pending.size() >= 2) {
            Pending old = pending.front();
            pending.erase(pending.begin());
            CG_CHECK_HIP(hipEventSynchronize(old.ev));
            CG_CHECK_HIP(hipEventDestroy(old.ev));
            if (g_done_slots[old.slot]) stopped = true;
        }
        return 0;
    };
    // pairs: partials of the exact column sweep; else sums of the fused sweep.
    auto combine = [&](bool pairs, int cur, int want_err, int check_index) -> int {
        int* flag = check_flag(want_err);
        if (pairs)
            launch_sk_combine<true>(n, p.comb_blocks, s, w.part,
                               p.col_nblk, ni, static_cast<int64_t>(ni), w.v2[cur], w.v2[cur ^ 1], momentum ? 1 : 0,
                               want_err, check_index, cur, threshold, eval_freq, momentum_start, w.errpart, w.ctl, flag);
        else
            launch_sk_combine<false>(n, p.comb_blocks, s, w.part,
                               p.nblk, ni, part_ld(ni), w.v2[cur], w.v2[cur ^ 1], momentum ? 1 : 0, want_err,
                               check_index, cur, threshold, eval_freq, momentum_start, w.errpart, w.ctl, flag);
        if (int rc = check_launch("sk_combine")) return rc;
        if (momentum) {
            hipLaunchKernelGGL(sk_apply_momentum, vgrid, dim3(256), 0, s, w.v2[cur], w.v2[cur ^ 1], ni, w.ctl);
            if (int rc = check_launch("sk_apply_momentum")) return rc;
        }
        return 0;
    };
    auto is_check = [&](int it) { return it % eval_freq == 0 && it != 0; };

    int cur = 0;
    int final_default = 0;
    if (max_iterations > 0) {
        if (use_fused) {
            // First v-update with the incoming u: exact column sweep (no bound
            // on the plan holds before it).
            if (int rc = launch_col(C, ni, ldc, scale2, w.v2[0], w.u2, static_cast<float2*>(w.part), p, w.ctl, s))
                return rc;
            if (int rc = combine(true, 0, 0, -1)) return rc;
            cur = 1;
            for (int it = 0; it < max_iterations && !stopped; ++it) {
                const bool chk = is_check(it);
                const bool timed = timer_sampled(it);
                if (timed)
                    if (int rc = timer_mark(s)) return rc;
                if (int rc = launch_fused(C, ni, scale2, w.v2[cur], w.u2, static_cast<float*>(w.part), p, w.ctl, s,
                                          it & 1))
                    return rc;
                if (timed)
                    if (int rc = timer_mark(s)) return rc;
                if (int rc = combine(false, cur, chk ? 1 : 0, it)) return rc;
                final_default = cur;
                cur ^= 1;
                if (chk)
                    if (int rc = post_check()) return rc;
            }
        } else {
            // Two sweeps per iteration; the check of iteration it uses the
            // column pairs of iteration it+1 (same u, same v).
            for (int it = 0; it < max_iterations && !stopped; ++it) {
                const bool prev_chk = it > 0 && is_check(it - 1);
                const bool timed = timer_sampled(it);
                if (timed)
                    if (int rc = timer_mark(s)) return rc;
                if (int rc = launch_col(C, ni, ldc, scale2, w.v2[cur], w.u2, static_cast<float2*>(w.part), p, w.ctl, s))
                    return rc;
                if (timed)
                    if (int rc = timer_mark(s)) return rc;
                if (int rc = combine(true, cur, prev_chk ? 1 : 0, it - 1)) return rc;
                cur ^= 1;
                if (timed)
                    if (int rc = timer_mark(s)) return rc;
                if (int rc = launch_row(C, ni, ldc, scale2, w.v2[cur], 1.0f, 1.0f, w.u2, 1.0f, w.ctl, s)) return rc;
                if (timed)
                    if (int rc = timer_mark(s)) return rc;
                final_default = cur;
                if (prev_chk)
                    if (int rc = post_check()) return rc;
            }
            if (!stopped && is_check(max_iterations - 1)) {
                // The last iteration is a checking one: one more column sweep
                // to record its error (the returned state is unchanged).
                if (int rc = launch_col(C, ni, ldc, scale2, w.v2[cur], w.u2, static_cast<float2*>(w.part), p, w.ctl, s))
                    return rc;
                if (int rc = combine(true, cur, 1, max_iterations - 1)) return rc;
            }
        }
    }
    hipLaunchKernelGGL(sk_finish, vgrid, dim3(256), 0, s, w.u2, w.v2[0], w.v2[1], u, v, ni, w.ctl, final_default);
    if (int rc = check_launch("sk_finish")) return rc;
    CG_CHECK_HIP(hipMemcpyAsync(g_mirror.ptr, w.ctl, ctl_bytes(cap), hipMemcpyDeviceToHost, s));
    CG_CHECK_HIP(hipStreamSynchronize(s));
    for (auto& pe : pending) (void)hipEventDestroy(pe.ev);
    const SinkCtl* hc = static_cast<const SinkCtl*>(g_mirror.ptr);
    const double* herr = reinterpret_cast<const double*>(static_cast<const char*>(g_mirror.ptr) + sizeof(SinkCtl));
    if (iterations) *iterations = hc->done ? hc->iterations : max_iterations;
    const int ne = std::min(hc->nerr, errors ? errors_capacity : 0);
    if (errors) std::memcpy(errors, herr, sizeof(double) * static_cast<size_t>(std::max(0, ne)));
    if (error_count) *error_count = hc->nerr;
    // Timed sweeps that did work: those of the sampled iterations before the
    // stop (one sweep each on the fused path, two on the two-sweep path; the
    // column sweep whose combine raises the stop belongs to an iteration that
    // is a multiple of the check period, never sampled).
    int64_t worked = -1;
    if (hc->done) worked = (use_fused ? 1 : 2) * static_cast<int64_t>(timer_samples_before(hc->iterations));
    if (int rc = timer_total(sweep_ms, sweep_launches, worked)) return rc;
    return CG_OK;
}

int cg_sinkhorn_standard_solve(const float* C, int64_t n, float regularization, float* u, float* v,
                               int max_iterations, double threshold, int eval_freq, int momentum_start, int fused,
                               double* errors, int errors_capacity, int* error_count, int* iterations,
                               void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!C || !u || !v || n <= 0 || n > (int64_t(1) << 30))
        return fail(CG_EINVAL, "cg_sinkhorn_standard_solve: bad arguments (n=%lld)", static_cast<long long>(n));
    if (!(regularization > 0.0f)) return fail(CG_EINVAL, "sinkhorn regularization must be > 0");
    if (eval_freq <= 0) return fail(CG_EINVAL, "sinkhorn eval_freq must be > 0");
    if (max_iterations < 0) return fail(CG_EINVAL, "sinkhorn iterations must be >= 0");
    const int ni = static_cast<int>(n);
    const int use_fused = fused && cg_sinkhorn_fused_supported(n) && aligned16(C);
    const Plan p = make_plan(n, use_fused);
    // Checks at 0, f, 2f, ...: at most max_iterations / f + 1 <= capacity.
    const int cap = err_capacity(max_iterations, eval_freq);
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    SinkWs w{};
    carve(cv, n, p, cap, &w);
    if (!workspace || !cv.ok())
        return fail(CG_EWORKSPACE, "cg_sinkhorn_standard_solve: workspace too small (%zu < %zu)", workspace_bytes,
                    cv.used);
    if (int rc = ensure_pinned(ctl_bytes(cap))) return rc;
    const float scale2 = -kLog2e / regularization;
    const bool momentum = momentum_start >= 0;
    const dim3 vgrid(static_cast<unsigned>(div_up(n, 256)));
    float* part = static_cast<float*>(w.part);
    float* uw = w.u2;  // natural units here
    hipLaunchKernelGGL(std_init, vgrid, dim3(256), 0, s, u, v, uw, w.v2[0], ni, w.ctl);
    if (int rc = check_launch("std_init")) return rc;

    struct Pending {
        hipEvent_t ev;
        int slot;
    };
    std::vector<Pending> pending;
    int next_slot = 0;
    bool stopped = false;
    // A checking combine stores its stop decision into a cleared pinned slot;
    // the host reads the slot once the event after the combine has passed.
    int check_slot = 0;
    auto check_flag = [&](int want_err) -> int* {
        if (!want_err) return nullptr;
        check_slot = next_slot++ & 63;
        g_done_slots[check_slot] = 0;
        return g_done_slots + check_slot;
    };
    auto post_check = [&]() -> int {
        hipEvent_t ev;
        CG_CHECK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        const int slot = check_slot;
        CG_CHECK_HIP(hipEventRecord(ev, s));
        pending.push_back({ev, slot});
        if (pending.size() >= 2) {
            Pending old = pending.front();
            pending.erase(pending.begin());
            CG_CHECK_HIP(hipEventSynchronize(old.ev));
            CG_CHECK_HIP(hipEventDestroy(old.ev));
            if (g_done_slots[old.slot]) stopped = true;
        }
        return 0;
    };
    // fused: partials of the fused sweep (rows of part_ld(n)); else of the
    // column sweep (rows of n).
    auto combine = [&](bool fused_parts, int cur, int want_err, int check_index) -> int {
        int* flag = check_flag(want_err);
        launch_std_combine(n, p.comb_blocks, s, part,
                           fused_parts ? p.nblk : p.col_nblk, ni, fused_parts ? part_ld(ni) : static_cast<int64_t>(ni),
                           w.v2[cur], w.v2[cur ^ 1], want_err, check_index, cur, threshold, eval_freq, momentum_start,
                           w.errpart, w.ctl, flag);
        if (int rc = check_launch("std_combine")) return rc;
        if (momentum) {
            hipLaunchKernelGGL(std_apply_momentum, vgrid, dim3(256), 0, s, w.v2[cur], w.v2[cur ^ 1], ni, w.ctl);
            if (int rc = check_launch("std_apply_momentum")) return rc;
        }
        return 0;
    };
    // The reference checks at iteration 0 too (cugal/sinkhorn.py:138).
    auto is_check = [&](int it) { return it % eval_freq == 0; };

    int cur = 0;
    int final_default = 0;
    if (max_iterations > 0) {
        // v_1 from the incoming u, by the column sweep.
        if (int rc = launch_std_col(C, ni, scale2, uw, part, p, w.ctl, s)) return rc;
        if (int rc = combine(false, 0, 0, -1)) return rc;
        cur = 1;
        if (use_fused) {
            for (int it = 0; it < max_iterations && !stopped; ++it) {
                const bool chk = is_check(it);
                if (int rc = launch_std_fused(C, ni, scale2, w.v2[cur], uw, part, p, w.ctl, s)) return rc;
                if (int rc = combine(true, cur, chk ? 1 : 0, it)) return rc;
                final_default = cur;
                cur ^= 1;
                if (chk)
                    if (int rc = post_check()) return rc;
            }
        } else {
            // Row sweep (u from v), then the column sweep of the new u, which
            // both checks this iteration and forms the next v.
            for (int it = 0; it < max_iterations && !stopped; ++it) {
                const bool chk = is_check(it);
                if (int rc = launch_std_row(C, ni, scale2, w.v2[cur], uw, w.ctl, s)) return rc;
                if (int rc = launch_std_col(C, ni, scale2, uw, part, p, w.ctl, s)) return rc;
                if (int rc = combine(false, cur, chk ? 1 : 0, it)) return rc;
                final_default = cur;
                cur ^= 1;
                if (chk)
                    if (int rc = post_check()) return rc;
            }
        }
    }
    hipLaunchKernelGGL(std_finish, vgrid, dim3(256), 0, s, uw, w.v2[0], w.v2[1], u, v, ni, w.ctl, final_default);
    if (int rc = check_launch("std_finish")) return rc;
    CG_CHECK_HIP(hipMemcpyAsync(g_mirror.ptr, w.ctl, ctl_bytes(cap), hipMemcpyDeviceToHost, s));
    CG_CHECK_HIP(hipStreamSynchronize(s));
    for (auto& pe : pending) (void)hipEventDestroy(pe.ev);
    const SinkCtl* hc = static_cast<const SinkCtl*>(g_mirror.ptr);
    const double* herr = reinterpret_cast<const double*>(static_cast<const char*>(g_mirror.ptr) + sizeof(SinkCtl));
    if (iterations) *iterations = hc->done ? hc->iterations : max_iterations;
    const int ne = std::min(hc->nerr, errors ? errors_capacity : 0);
    if (errors) std::memcpy(errors, herr, sizeof(double) * static_cast<size_t>(std::max(0, ne)));
    if (error_count) *error_count = hc->nerr;
    return CG_OK;
}

namespace {

struct SinkWs64 {
    double* u;
    double* v[2];
    double* part;
    double* errpart;
    SinkCtl* ctl;
};

size_t carve64(Carver& cv, int64_t n, const Plan& p, int cap, SinkWs64* w) {
    SinkWs64 t{};
    t.u = cv.take<double>(n);
    t.v[0] = cv.take<double>(n);
    t.v[1] = cv.take<double>(n);
    t.part = cv.take<double>(static_cast<size_t>(p.col_nblk) * n);
    t.errpart = cv.take<double>(p.comb_blocks);
    t.ctl = reinterpret_cast<SinkCtl*>(cv.take<char>(ctl_bytes(cap)));
    if (w) *w = t;
    return cv.used;
}

}  // namespace

size_t cg_sinkhorn_standard_f64_workspace(int64_t n, int max_iterations, int eval_freq) {
    if (n <= 0) return 0;
    Carver cv{nullptr, 0, 0};
    return carve64(cv, n, make_plan(n, 0), err_capacity(max_iterations, eval_freq), nullptr) + 256;
}

int cg_sinkhorn_standard_solve_f64(const double* C, int64_t n, double regularization, double* u, double* v,
                                   int max_iterations, double threshold, int eval_freq, int momentum_start,
                                   double* errors, int errors_capacity, int* error_count, int* iterations,
                                   void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!C || !u || !v || n <= 0 || n > (int64_t(1) << 30))
        return fail(CG_EINVAL, "cg_sinkhorn_standard_solve_f64: bad arguments (n=%lld)", static_cast<long long>(n));
    if (!(regularization > 0.0)) return fail(CG_EINVAL, "sinkhorn regularization must be > 0");
    if (eval_freq <= 0) return fail(CG_EINVAL, "sinkhorn eval_freq must be > 0");
    if (max_iterations < 0) return fail(CG_EINVAL, "sinkhorn iterations must be >= 0");
    const int ni = static_cast<int>(n);
    const Plan p = make_plan(n, 0);
    const int cap = err_capacity(max_iterations, eval_freq);
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    SinkWs64 w{};
    carve64(cv, n, p, cap, &w);
    if (!workspace || !cv.ok())
        return fail(CG_EWORKSPACE, "cg_sinkhorn_standard_solve_f64: workspace too small (%zu < %zu)",
                    workspace_bytes, cv.used);
    if (int rc = ensure_pinned(ctl_bytes(cap))) return rc;
    const double nrinv = -1.0 / regularization;
    const bool momentum = momentum_start >= 0;
    const dim3 vgrid(static_cast<unsigned>(div_up(n, 256)));
    hipLaunchKernelGGL(std64_init, vgrid, dim3(256), 0, s, u, v, w.u, w.v[0], ni, w.ctl);
    if (int rc = check_launch("std64_init")) return rc;

    struct Pending {
        hipEvent_t ev;
        int slot;
    };
    std::vector<Pending> pending;
    int next_slot = 0;
    bool stopped = false;
    int check_slot = 0;
    auto post_check = [&]() -> int {
        hipEvent_t ev;
        CG_CHECK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        const int slot = check_slot;
        CG_CHECK_HIP(hipEventRecord(ev, s));
        pending.push_back({ev, slot});
        if (pending.size() >= 2) {
            Pending old = pending.front();
            pending.erase(pending.begin());
            CG_CHECK_HIP(hipEventSynchronize(old.ev));
            CG_CHECK_HIP(hipEventDestroy(old.ev));
            if (g_done_slots[old.slot]) stopped = true;
        }
        return 0;
    };
    auto col = [&]() -> int {
        hipLaunchKernelGGL((std64_col_sweep<kColThreads, kColCPT>), dim3(p.col_nblk, p.col_chunks),
                           dim3(kColThreads), 0, s, C, ni, nrinv, w.u, w.part, p.col_rows, w.ctl);
        return check_launch("std64_col_sweep");
    };
    auto combine = [&](int cur, int want_err, int check_index) -> int {
        int* flag = nullptr;
        if (want_err) {
            check_slot = next_slot++ & 63;
            g_done_slots[check_slot] = 0;
            flag = g_done_slots + check_slot;
        }
        launch_std64_combine(n, p.comb_blocks, s, w.part,
                           p.col_nblk, ni, w.v[cur], w.v[cur ^ 1], want_err, check_index, cur, threshold, eval_freq,
                           momentum_start, w.errpart, w.ctl, flag);
        if (int rc = check_launch("std64_combine")) return rc;
        if (momentum) {
            hipLaunchKernelGGL(std64_apply_momentum, vgrid, dim3(256), 0, s, w.v[cur], w.v[cur ^ 1], ni, w.ctl);
            if (int rc = check_launch("std64_apply_momentum")) return rc;
        }
        return 0;
    };
    int cur = 0;
    int final_default = 0;
    if (max_iterations > 0) {
        // v from the incoming u, then per iteration: u from v (row sweep),
        // the column sums of the new u (check of this iteration, next v).
        if (int rc = col()) return rc;
        if (int rc = combine(0, 0, -1)) return rc;
        cur = 1;
        for (int it = 0; it < max_iterations && !stopped; ++it) {
            const bool chk = it % eval_freq == 0;  // the reference checks at 0 too (cugal/sinkhorn.py:138)
            hipLaunchKernelGGL((std64_row_sweep<kRowWaves>), dim3(static_cast<unsigned>(div_up(n, kRowWaves))),
                               dim3(kRowWaves * kWave), 0, s, C, ni, nrinv, w.v[cur], w.u, w.ctl);
            if (int rc = check_launch("std64_row_sweep")) return rc;
            if (int rc = col()) return rc;
            if (int rc = combine(cur, chk ? 1 : 0, it)) return rc;
            final_default = cur;
            cur ^= 1;
            if (chk)
                if (int rc = post_check()) return rc;
        }
    }
    hipLaunchKernelGGL(std64_finish, vgrid, dim3(256), 0, s, w.u, w.v[0], w.v[1], u, v, ni, w.ctl, final_default);
    if (int rc = check_launch("std64_finish")) return rc;
    CG_CHECK_HIP(hipMemcpyAsync(g_mirror.ptr, w.ctl, ctl_bytes(cap), hipMemcpyDeviceToHost, s));
    CG_CHECK_HIP(hipStreamSynchronize(s));
    for (auto& pe : pending) (void)hipEventDestroy(pe.ev);
    const SinkCtl* hc = static_cast<const SinkCtl*>(g_mirror.ptr);
    const double* herr = reinterpret_cast<const double*>(static_cast<const char*>(g_mirror.ptr) + sizeof(SinkCtl));
    if (iterations) *iterations = hc->done ? hc->iterations : max_iterations;
    const int ne = std::min(hc->nerr, errors ? errors_capacity : 0);
    if (errors) std::memcpy(errors, herr, sizeof(double) * static_cast<size_t>(std::max(0, ne)));
    if (error_count) *error_count = hc->nerr;
    return CG_OK;
}

int cg_sinkhorn_log_step(const float* K, int64_t n, const float* add, float* out, float momentum,
                         cg_stream_t stream) {
    if (!K || !add || !out || n <= 0 || n > (int64_t(1) << 30))
        return fail(CG_EINVAL, "cg_sinkhorn_log_step: bad arguments");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const float mom = std::fabs(momentum - 1.0f) < 1e-4f ? 1.0f : momentum;
    return launch_row(K, static_cast<int>(n), n, kLog2e, add, kLog2e, kLn2, out, mom, nullptr, s);
}

size_t cg_sinkhorn_log_marginal_workspace(int64_t n) { return n > 0 ? sizeof(double) * (n + 1) + 256 : 0; }

int cg_sinkhorn_log_marginal(const float* K, int64_t n, const float* u, const float* v, double* result,
                             void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    if (!K || !u || !v || !result || n <= 0 || n > (int64_t(1) << 30))
        return fail(CG_EINVAL, "cg_sinkhorn_log_marginal: bad arguments");
    if (!workspace || workspace_bytes < cg_sinkhorn_log_marginal_workspace(n))
        return fail(CG_EWORKSPACE, "cg_sinkhorn_log_marginal: workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    double* rows = static_cast<double*>(workspace);
    double* total = rows + n;
    hipLaunchKernelGGL(sk_row_marginal, dim3(static_cast<unsigned>(div_up(n, 4))), dim3(256), 0, s, K,
                       static_cast<int>(n), u, v, rows);
    if (int rc = check_launch("sk_row_marginal")) return rc;
    hipLaunchKernelGGL(sum_doubles, dim3(1), dim3(256), 0, s, rows, static_cast<int>(n), total);
    if (int rc = check_launch("sum_doubles")) return rc;
    CG_CHECK_HIP(hipMemcpyAsync(result, total, sizeof(double), hipMemcpyDeviceToHost, s));
    CG_CHECK_HIP(hipStreamSynchronize(s));
    return CG_OK;
}

}  // extern "C"
