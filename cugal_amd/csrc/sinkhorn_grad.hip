// Differentiable unrolled log-domain Sinkhorn (the adjoint of T iterations of
// cugal/sinkhorn.py:90-111 from given duals), and the feature cost
// C = lambda * D(w), D_ij = sum_k w_k |fs_ik - ft_jk|.
//
// Forward: T iterations v^t = -LSE_i(K + u^{t-1}), u^t = -LSE_j(K + v^t) with
// K = -C/reg; only (u, v) at every k-th iteration are kept (checkpoints).
//
// With alpha^t_ij = exp(K_ij + u^{t-1}_i + v^t_j) (column-stochastic) and
// beta^t_ij = exp(K_ij + u^t_i + v^t_j) (row-stochastic), plan = beta^T, and
// G = dL/dplan, the adjoints obey
//   ubar^T = sum_j G beta^T,   vbar^T = sum_i (G - ubar^T_i) beta^T,
//   ubar^{t-1} = -sum_j vbar^t_j alpha^t,   vbar^{t-1} = -sum_i ubar^{t-1}_i beta^{t-1},
//   dL/dK = G (.) plan - sum_t (ubar^t_i beta^t_ij + vbar^t_j alpha^t_ij),
// and dL/dC = -dL/dK / reg.  The backward walks the segments from last to
// first: it recomputes the segment's potentials from its checkpoint (same
// kernels, same bits), runs the adjoint recursion (a row contraction and a
// column contraction per iteration, fused into one read of C for n <= 8192),
// then one pass adds
// the segment's terms to dL/dC -- or, for the feature cost, contracts them on
// the fly into dL/dlambda = sum C_bar D and dL/dw_k = lambda sum C_bar |df_k|
// without storing an n x n gradient.  Exponentials are base 2; sums of
// products are accumulated per thread in FP32 and across blocks in FP64.
#include "common.hpp"
#include "cugal_hip.h"

#include <vector>

namespace cg {
size_t sinkhorn_segment_workspace(int64_t n);
int sinkhorn_segment(const float* C, int n, float scale2, const float* u_c, const float* v_c, int steps,
                     float* u_hist, float* v_hist, void* ws, size_t ws_bytes, hipStream_t s);

namespace {

constexpr int kMaxSegment = 64;  // iterations per checkpoint segment

// ---- conversions -----------------------------------------------------------
__global__ void scale_vec(const float* __restrict__ in, float* __restrict__ out, int n, float f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] * f;
}

// plan = exp2(C * scale2 + u2_i + v2_j)
__global__ void plan_kernel(const float* __restrict__ C, int n, float scale2, const float* __restrict__ u2,
                            const float* __restrict__ v2, float* __restrict__ out) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(n) * n) return;
    const int64_t i = idx / n;
    const int j = static_cast<int>(idx - i * n);
    out[idx] = exp2_fast(fmaf(C[idx], scale2, u2[i] + v2[j]));
}

// C = lambda * sum_k w_k |fs_ik - ft_jk|
__global__ void feature_cost_kernel(const float* __restrict__ fs, const float* __restrict__ ft,
                                    const float* __restrict__ w, float lam, int n, float* __restrict__ C) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<int64_t>(n) * n) return;
    const int64_t i = idx / n;
    const int64_t j = idx - i * n;
    const float4 a = *reinterpret_cast<const float4*>(fs + 4 * i);
    const float4 b = *reinterpret_cast<const float4*>(ft + 4 * j);
    const float d = w[0] * fabsf(a.x - b.x) + w[1] * fabsf(a.y - b.y) + w[2] * fabsf(a.z - b.z) +
                    w[3] * fabsf(a.w - b.w);
    C[idx] = lam * d;
}

// ---- row contraction: out_i = sign * sum_j W_ij 2^(C_ij s + a_i + b_j) ------
// W_ij = w_j (vector) or G_ij (matrix).  One wave per row, FP32 lanes, the
// wave sum in FP32 of 64 partials (each at most n/64 terms).
template <bool MAT>
__global__ __launch_bounds__(256) void bw_row(const float* __restrict__ C, int n, float scale2,
                                             const float* __restrict__ a, const float* __restrict__ b,
                                             const float* __restrict__ w, const float* __restrict__ G, float sign,
                                             float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / 64;
    if (i >= n) return;
    const float ai = a[i];
    const float* row = C + i * n;
    double acc = 0.0;
    float part = 0.0f;
    int cnt = 0;
    for (int j = lane; j < n; j += 64) {
        const float e = exp2_fast(fmaf(row[j], scale2, ai + b[j]));
        part = fmaf(MAT ? G[i * n + j] : w[j], e, part);
        if (++cnt == 64) {  // flush to FP64 every 64 terms
            acc += part;
            part = 0.0f;
            cnt = 0;
        }
    }
    acc += part;
    acc = wave_sum(acc);
    if (lane == 0) out[i] = static_cast<float>(sign * acc);
}

// ---- column contraction partials ---------------------------------------------
//   part[blk][j] = sum_{i in blk} W_ij 2^(C_ij s + a_i + b_j),
// W_ij = w_i (vector) or (G_ij - w_i) (matrix).  Block (bx, by): rows
// [bx*rpb, ...), columns of chunk by.
constexpr int kColT = 256;
constexpr int kColC = 4;  // columns per thread

template <bool MAT>
__global__ __launch_bounds__(kColT) void bw_col(const float* __restrict__ C, int n, float scale2,
                                               const float* __restrict__ a, const float* __restrict__ b,
                                               const float* __restrict__ w, const float* __restrict__ G,
                                               int rows_per_blk, float* __restrict__ part) {
    const int c0 = blockIdx.y * (kColT * kColC);
    float bj[kColC], acc[kColC];
    bool ok[kColC];
#pragma unroll
    for (int q = 0; q < kColC; ++q) {
        const int j = c0 + q * kColT + threadIdx.x;
        ok[q] = j < n;
        bj[q] = ok[q] ? b[j] : 0.0f;
        acc[q] = 0.0f;
    }
    const int r0 = blockIdx.x * rows_per_blk;
    const int r1 = min(n, r0 + rows_per_blk);
    for (int i = r0; i < r1; ++i) {
        const float* row = C + static_cast<int64_t>(i) * n;
        const float ai = a[i];
        const float wi = w[i];
#pragma unroll
        for (int q = 0; q < kColC; ++q) {
            if (!ok[q]) continue;
            const int j = c0 + q * kColT + threadIdx.x;
            const float e = exp2_fast(fmaf(row[j], scale2, ai + bj[q]));
            const float wt = MAT ? (G[static_cast<int64_t>(i) * n + j] - wi) : wi;
            acc[q] = fmaf(wt, e, acc[q]);
        }
    }
    float* out = part + static_cast<int64_t>(blockIdx.x) * n;
#pragma unroll
    for (int q = 0; q < kColC; ++q)
        if (ok[q]) out[c0 + q * kColT + threadIdx.x] = acc[q];
}

// out_j = sign * sum_blk part[blk][j] in FP64.  A block covers 64 columns;
// wave w sums the partial rows b = w mod 4, then wave 0 adds the four wave
// totals in order (fixed order: deterministic).
constexpr int kSumCols = 64;
constexpr int kSumWaves = 4;

__global__ __launch_bounds__(kSumCols* kSumWaves) void bw_col_sum(const float* __restrict__ part, int nblk, int n,
                                                                 float sign, float* __restrict__ out) {
    __shared__ double ws[kSumWaves][kSumCols];
    const int lane = threadIdx.x & (kSumCols - 1);
    const int wave = threadIdx.x / kSumCols;
    const int j = blockIdx.x * kSumCols + lane;
    double s = 0.0;
    if (j < n) {
#pragma unroll 8
        for (int b = wave; b < nblk; b += kSumWaves) s += static_cast<double>(part[static_cast<int64_t>(b) * n + j]);
    }
    ws[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && j < n) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kSumWaves; ++w) t += ws[w][lane];
        out[j] = static_cast<float>(sign * t);
    }
}

// ---- fused adjoint sweep: one read of C per adjoint iteration -------------
// Row i of block b (rows dealt round-robin), with t_ij = K2_ij + a_i:
//   e_ij = 2^(t_ij + b_j),  r_i = row_sign * sum_j W_ij e_ij   -> out_row[i]
//   column partials (do_col):  part[b][j] += r_i 2^(t_ij + c_j)       (!MAT)
//                              part[b][j] += (G_ij - r_i) e_ij        (MAT)
// W_ij = w_j (!MAT) or G_ij (MAT).  With a = u^{t-1}, b = v^t, w = vbar^t,
// c = v^{t-1} and row_sign = -1 this is the pair of contractions
//   ubar^{t-1} = -sum_j vbar^t alpha^t,   vbar^{t-1} = -sum_i ubar^{t-1} beta^{t-1}
// (the second finished by bw_col_sum with sign -1); with a = u^T, b = v^T,
// W = G and row_sign = +1 it starts the recursion from dL/dplan.  b, w, c sit
// in LDS (3n floats); a thread owns the float4 column groups g*T + t.  The
// row sum is FP32 per thread (at most 4G terms) and FP64 across the block.
constexpr int kBwT = 256;
constexpr int kBwFusedMaxN = 8192;  // 3n floats of LDS

template <int G, bool FULL, bool MAT>
__global__ __launch_bounds__(kBwT) void bw_fused(const float* __restrict__ C, int n, float scale2,
                                                const float* __restrict__ a2, const float* __restrict__ b2,
                                                const float* __restrict__ wv, const float* __restrict__ Gm,
                                                const float* __restrict__ c2, float row_sign,
                                                float* __restrict__ out_row, int do_col, float* __restrict__ part,
                                                int rows_per_blk) {
    extern __shared__ float smem[];
    constexpr int W = kBwT / kWave;
    __shared__ double red[2][W];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const int np = static_cast<int>(round_up(n, 4));
    float* sb = smem;
    float* sw = smem + np;
    float* sc = smem + 2 * np;
    for (int j = tid * 4; j < n; j += kBwT * 4) {
        *reinterpret_cast<float4*>(sb + j) = *reinterpret_cast<const float4*>(b2 + j);
        if (!MAT) {
            *reinterpret_cast<float4*>(sw + j) = *reinterpret_cast<const float4*>(wv + j);
            *reinterpret_cast<float4*>(sc + j) = *reinterpret_cast<const float4*>(c2 + j);
        }
    }
    __syncthreads();
    float cs[4 * G];
#pragma unroll
    for (int k = 0; k < 4 * G; ++k) cs[k] = 0.0f;
    const int nblk = static_cast<int>(gridDim.x);
    const int rows = min(rows_per_blk, (n - static_cast<int>(blockIdx.x) + nblk - 1) / nblk);
    for (int k = 0; k < rows; ++k) {
        const int i = static_cast<int>(blockIdx.x) + k * nblk;
        const int64_t base = static_cast<int64_t>(i) * n;
        const float ai = a2[i];
        float4 t[G], e[G], gm[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int col = (g * kBwT + tid) * 4;
            const bool in = FULL || col < n;
            t[g] = in ? *reinterpret_cast<const float4*>(C + base + col) : make_float4(0.f, 0.f, 0.f, 0.f);
            if (MAT) gm[g] = in ? *reinterpret_cast<const float4*>(Gm + base + col) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        float rs = 0.0f;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int col = (g * kBwT + tid) * 4;
            if (FULL || col < n) {
                t[g].x = fmaf(t[g].x, scale2, ai);
                t[g].y = fmaf(t[g].y, scale2, ai);
                t[g].z = fmaf(t[g].z, scale2, ai);
                t[g].w = fmaf(t[g].w, scale2, ai);
                const float4 bb = *reinterpret_cast<const float4*>(sb + col);
                e[g] = make_float4(exp2_fast(t[g].x + bb.x), exp2_fast(t[g].y + bb.y), exp2_fast(t[g].z + bb.z),
                                   exp2_fast(t[g].w + bb.w));
                const float4 ww = MAT ? gm[g] : *reinterpret_cast<const float4*>(sw + col);
                rs = fmaf(ww.x, e[g].x, rs);
                rs = fmaf(ww.y, e[g].y, rs);
                rs = fmaf(ww.z, e[g].z, rs);
                rs = fmaf(ww.w, e[g].w, rs);
            }
        }
        // Rows alternate between two reduction slots: a wave that runs ahead
        // writes the other slot while this one is still being read.
        const double ws = wave_sum(static_cast<double>(rs));
        if (lane == 0) red[k & 1][wave] = ws;
        __syncthreads();
        double tot = 0.0;
#pragma unroll
        for (int w = 0; w < W; ++w) tot += red[k & 1][w];
        const float ri = static_cast<float>(static_cast<double>(row_sign) * tot);
        if (tid == 0) out_row[i] = ri;
        if (!do_col) continue;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int col = (g * kBwT + tid) * 4;
            if (FULL || col < n) {
                if (MAT) {
                    cs[4 * g + 0] = fmaf(gm[g].x - ri, e[g].x, cs[4 * g + 0]);
                    cs[4 * g + 1] = fmaf(gm[g].y - ri, e[g].y, cs[4 * g + 1]);
                    cs[4 * g + 2] = fmaf(gm[g].z - ri, e[g].z, cs[4 * g + 2]);
                    cs[4 * g + 3] = fmaf(gm[g].w - ri, e[g].w, cs[4 * g + 3]);
                } else {
                    const float4 cc = *reinterpret_cast<const float4*>(sc + col);
                    cs[4 * g + 0] = fmaf(ri, exp2_fast(t[g].x + cc.x), cs[4 * g + 0]);
                    cs[4 * g + 1] = fmaf(ri, exp2_fast(t[g].y + cc.y), cs[4 * g + 1]);
                    cs[4 * g + 2] = fmaf(ri, exp2_fast(t[g].z + cc.z), cs[4 * g + 2]);
                    cs[4 * g + 3] = fmaf(ri, exp2_fast(t[g].w + cc.w), cs[4 * g + 3]);
                }
            }
        }
    }
    if (!do_col) return;
    float* out = part + static_cast<int64_t>(blockIdx.x) * n;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int col = (g * kBwT + tid) * 4;
        if (FULL || col < n)
            *reinterpret_cast<float4*>(out + col) = make_float4(cs[4 * g], cs[4 * g + 1], cs[4 * g + 2], cs[4 * g + 3]);
    }
}

// ---- gradient pass of one segment ------------------------------------------
// For iterations q = 1..steps of the segment (potentials uh[0..steps],
// vh[0..steps], adjoints ub[1..steps], vb[1..steps], rows of length n):
//   acc_ij = sum_q ub[q]_i 2^(K2 + uh[q]_i + vh[q]_j) + vb[q]_j 2^(K2 + uh[q-1]_i + vh[q]_j)
//   Kbar_ij = (with_plan ? G_ij 2^(K2 + uT_i + vT_j) : 0) - acc_ij,   Cbar = -Kbar / reg.
// Output: Cbar written (first segment) or added (later ones), and/or the
// feature contractions reduced per block into red[blk][0..4]
// (dlambda = sum Cbar D, dw_k = lambda sum Cbar |fs_ik - ft_jk|).
// A block walks 64 x 64 tiles (grid-stride, fixed grid, so the per-block
// sums are deterministic); wave w owns rows 16w..16w+15 of the tile and lane
// l column l, so the row potentials are LDS broadcasts and each q loads the
// lane's column pair (vh[q]_j, vb[q]_j) once for its 16 rows.
constexpr int kGT = 64;       // tile edge
constexpr int kGRows = 16;    // rows per wave
constexpr int kGradBlocks = 2048;

__global__ __launch_bounds__(256) void bw_grad(const float* __restrict__ C, int n, float scale2, float inv_reg,
                                              int steps, const float* __restrict__ uh, const float* __restrict__ vh,
                                              const float* __restrict__ ub, const float* __restrict__ vb,
                                              const float* __restrict__ G, const float* __restrict__ uT,
                                              const float* __restrict__ vT, float* __restrict__ cbar, int accumulate,
                                              const float* __restrict__ fs, const float* __restrict__ ft,
                                              const float* __restrict__ w, float lam, double* __restrict__ red) {
    extern __shared__ float sm[];
    // rows: uh[q] (q = 0..steps), ub[q] (q = 1..steps); columns: (vh[q], vb[q]) (q = 1..steps)
    float2* svv = reinterpret_cast<float2*>(sm);           // steps x kGT
    float* su = sm + 2 * steps * kGT;                       // (steps+1) x kGT
    float* sub = su + (steps + 1) * kGT;                    // steps x kGT
    __shared__ double dred[5][4];
    const int tiles_c = static_cast<int>(div_up(n, kGT));
    const int tiles = tiles_c * tiles_c;
    const int64_t nn = n;
    const int lane = threadIdx.x & (kWave - 1);
    const int rbase = (threadIdx.x / kWave) * kGRows;
    double dl = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
    const float w0 = fs ? w[0] : 0.f, w1 = fs ? w[1] : 0.f, w2 = fs ? w[2] : 0.f, w3 = fs ? w[3] : 0.f;
    for (int tile = static_cast<int>(blockIdx.x); tile < tiles; tile += static_cast<int>(gridDim.x)) {
        const int r0 = (tile / tiles_c) * kGT;
        const int c0 = (tile % tiles_c) * kGT;
        __syncthreads();  // the previous tile's LDS reads are done
        for (int idx = threadIdx.x; idx < (steps + 1) * kGT; idx += blockDim.x) {
            const int q = idx / kGT;
            const int r = min(r0 + idx % kGT, n - 1);
            su[idx] = uh[q * nn + r];
            if (q >= 1) sub[idx - kGT] = ub[q * nn + r];
        }
        for (int idx = threadIdx.x; idx < steps * kGT; idx += blockDim.x) {
            const int q = idx / kGT + 1;
            const int c = min(c0 + idx % kGT, n - 1);
            svv[idx] = make_float2(vh[q * nn + c], vb[q * nn + c]);
        }
        __syncthreads();
        const int j = c0 + lane;
        const bool jin = j < n;
        float k2[kGRows], acc[kGRows];
#pragma unroll
        for (int r = 0; r < kGRows; ++r) {
            const int i = r0 + rbase + r;
            k2[r] = (jin && i < n) ? C[i * nn + j] * scale2 : 0.0f;
            acc[r] = 0.0f;
        }
        for (int q = 1; q <= steps; ++q) {
            const float2 vv = svv[(q - 1) * kGT + lane];
            const float* uq = su + q * kGT + rbase;
            const float* uq1 = su + (q - 1) * kGT + rbase;
            const float* ubq = sub + (q - 1) * kGT + rbase;
#pragma unroll
            for (int r = 0; r < kGRows; ++r) {
                acc[r] = fmaf(ubq[r], exp2_fast(k2[r] + uq[r] + vv.x), acc[r]);
                acc[r] = fmaf(vv.y, exp2_fast(k2[r] + uq1[r] + vv.x), acc[r]);
            }
        }
        float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
        if (fs && jin) b = *reinterpret_cast<const float4*>(ft + 4 * static_cast<int64_t>(j));
        const float vTj = (G && jin) ? vT[j] : 0.0f;
        float fl = 0.f, f0 = 0.f, f1 = 0.f, f2 = 0.f, f3 = 0.f;
#pragma unroll
        for (int r = 0; r < kGRows; ++r) {
            const int i = r0 + rbase + r;
            if (!jin || i >= n) continue;
            const int64_t idx = i * nn + j;
            float kbar = -acc[r];
            if (G) kbar = fmaf(G[idx], exp2_fast(k2[r] + uT[i] + vTj), kbar);
            const float cb = -kbar * inv_reg;
            if (cbar) cbar[idx] = accumulate ? cbar[idx] + cb : cb;
            if (fs) {
                const float4 a = *reinterpret_cast<const float4*>(fs + 4 * static_cast<int64_t>(i));
                const float e0 = fabsf(a.x - b.x), e1 = fabsf(a.y - b.y), e2 = fabsf(a.z - b.z), e3 = fabsf(a.w - b.w);
                fl = fmaf(cb, w0 * e0 + w1 * e1 + w2 * e2 + w3 * e3, fl);
                f0 = fmaf(cb, e0, f0);
                f1 = fmaf(cb, e1, f1);
                f2 = fmaf(cb, e2, f2);
                f3 = fmaf(cb, e3, f3);
            }
        }
        // FP32 over the tile's 16 rows, FP64 across tiles.
        dl += fl;
        d0 += f0;
        d1 += f1;
        d2 += f2;
        d3 += f3;
    }
    if (!fs) return;
    const double vals[5] = {dl, lam * d0, lam * d1, lam * d2, lam * d3};
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double s = wave_sum(vals[m]);
        if (lane == 0) dred[m][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        const int m = threadIdx.x;
        red[static_cast<int64_t>(blockIdx.x) * 5 + m] = ((dred[m][0] + dred[m][1]) + dred[m][2]) + dred[m][3];
    }
}

int grad_blocks(int64_t n) {
    const int64_t t = div_up(n, kGT);
    return static_cast<int>(std::min<int64_t>(t * t, kGradBlocks));
}

// sum over blocks of red[blk][m] (fixed order), added into out[m].
__global__ __launch_bounds__(256) void bw_red_sum(const double* __restrict__ red, int nblk, double* __restrict__ out) {
    __shared__ double scratch[4];
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
#pragma unroll
        for (int m = 0; m < 5; ++m) s[m] += red[static_cast<int64_t>(b) * 5 + m];
    for (int m = 0; m < 5; ++m) {
        const double t = block_sum_double(s[m], scratch);
        if (threadIdx.x == 0) out[m] += t;
    }
}

__global__ void zero_doubles(double* __restrict__ p, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

int col_plan_rows(int n) { return static_cast<int>(std::max<int64_t>(8, div_up(n, std::max<int64_t>(1, 1024 / div_up(n, kColT * kColC))))); }

struct GradWs {
    float* uh;      // (k+1) x n
    float* vh;      // (k+1) x n
    float* ub;      // (k+1) x n
    float* vb;      // (k+1) x n
    float* part;    // col partials
    double* red;    // gradient partials
    void* seg;      // segment runner workspace
    size_t seg_bytes;
};

// Fused adjoint sweep shape: 256-thread blocks, float4 groups per thread,
// about 512 row blocks (the partial rows of the column sums).
struct FusedShape {
    int groups;
    int rows_per_blk;
    int nblk;
};

FusedShape fused_shape(int64_t n) {
    const int64_t g = div_up(n, static_cast<int64_t>(kBwT) * 4);
    FusedShape f{};
    f.groups = g <= 1 ? 1 : g <= 2 ? 2 : g <= 4 ? 4 : 8;
    f.rows_per_blk = static_cast<int>(std::max<int64_t>(1, div_up(n, 512)));
    f.nblk = static_cast<int>(div_up(n, f.rows_per_blk));
    return f;
}

bool fused_ok(int64_t n, const float* C, const float* G) {
    return n <= kBwFusedMaxN && n % 4 == 0 && aligned16(C) && (G == nullptr || aligned16(G));
}

size_t grad_carve(Carver& cv, int64_t n, int k, GradWs* g) {
    GradWs t{};
    const int rpb = col_plan_rows(static_cast<int>(n));
    const int64_t nblk = std::max<int64_t>(div_up(n, rpb), fused_shape(n).nblk);
    t.uh = cv.take<float>((k + 1) * n);
    t.vh = cv.take<float>((k + 1) * n);
    t.ub = cv.take<float>((k + 1) * n);
    t.vb = cv.take<float>((k + 1) * n);
    t.part = cv.take<float>(nblk * n);
    t.red = cv.take<double>(static_cast<int64_t>(grad_blocks(n)) * 5);
    t.seg_bytes = sinkhorn_segment_workspace(n);
    t.seg = cv.take<char>(t.seg_bytes);
    if (g) *g = t;
    return cv.used;
}

int launch_col_sum(const float* part, int nblk, int n, float sign, float* out, hipStream_t s) {
    hipLaunchKernelGGL(bw_col_sum, dim3(static_cast<unsigned>(div_up(n, kSumCols))), dim3(kSumCols * kSumWaves), 0,
                       s, part, nblk, n, sign, out);
    return check_launch("bw_col_sum");
}

template <int G, bool MAT>
int launch_fused_g(const float* C, int n, float s2, const float* a, const float* b, const float* w, const float* Gm,
                   const float* c, float row_sign, float* out_row, int do_col, float* part, const FusedShape& f,
                   hipStream_t s) {
    const size_t lds = sizeof(float) * 3 * static_cast<size_t>(round_up(n, 4));
    static size_t attr_bytes = 0;
    if (lds > attr_bytes) {
        CG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&bw_fused<G, true, MAT>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        CG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&bw_fused<G, false, MAT>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        attr_bytes = lds;
    }
    if (n == 4 * G * kBwT)
        hipLaunchKernelGGL((bw_fused<G, true, MAT>), dim3(f.nblk), dim3(kBwT), lds, s, C, n, s2, a, b, w, Gm, c,
                           row_sign, out_row, do_col, part, f.rows_per_blk);
    else
        hipLaunchKernelGGL((bw_fused<G, false, MAT>), dim3(f.nblk), dim3(kBwT), lds, s, C, n, s2, a, b, w, Gm, c,
                           row_sign, out_row, do_col, part, f.rows_per_blk);
    return check_launch("bw_fused");
}

// One fused adjoint step: out_row (all rows) and, with do_col, the column
// contraction finished into out_col with sign col_sign.
template <bool MAT>
int fused_step(const float* C, int n, float s2, const float* a, const float* b, const float* w, const float* Gm,
               const float* c, float row_sign, float* out_row, int do_col, float col_sign, float* part,
               float* out_col, hipStream_t s) {
    const FusedShape f = fused_shape(n);
    int rc = 0;
    switch (f.groups) {
        case 1: rc = launch_fused_g<1, MAT>(C, n, s2, a, b, w, Gm, c, row_sign, out_row, do_col, part, f, s); break;
        case 2: rc = launch_fused_g<2, MAT>(C, n, s2, a, b, w, Gm, c, row_sign, out_row, do_col, part, f, s); break;
        case 4: rc = launch_fused_g<4, MAT>(C, n, s2, a, b, w, Gm, c, row_sign, out_row, do_col, part, f, s); break;
        default: rc = launch_fused_g<8, MAT>(C, n, s2, a, b, w, Gm, c, row_sign, out_row, do_col, part, f, s); break;
    }
    if (rc || !do_col) return rc;
    return launch_col_sum(part, f.nblk, n, col_sign, out_col, s);
}

int launch_row_c(bool mat, const float* C, int n, float s2, const float* a, const float* b, const float* w,
                 const float* G, float sign, float* out, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>(div_up(n, 4)));
    if (mat)
        hipLaunchKernelGGL(bw_row<true>, grid, dim3(256), 0, s, C, n, s2, a, b, w, G, sign, out);
    else
        hipLaunchKernelGGL(bw_row<false>, grid, dim3(256), 0, s, C, n, s2, a, b, w, G, sign, out);
    return check_launch("bw_row");
}

int launch_col_c(bool mat, const float* C, int n, float s2, const float* a, const float* b, const float* w,
                 const float* G, float sign, float* part, float* out, hipStream_t s) {
    const int rpb = col_plan_rows(n);
    const int nblk = static_cast<int>(div_up(n, rpb));
    const dim3 grid(nblk, static_cast<unsigned>(div_up(n, kColT * kColC)));
    if (mat)
        hipLaunchKernelGGL(bw_col<true>, grid, dim3(kColT), 0, s, C, n, s2, a, b, w, G, rpb, part);
    else
        hipLaunchKernelGGL(bw_col<false>, grid, dim3(kColT), 0, s, C, n, s2, a, b, w, G, rpb, part);
    if (int rc = check_launch("bw_col")) return rc;
    return launch_col_sum(part, nblk, n, sign, out, s);
}

bool bad_args(int64_t n, int T, int k) { return n <= 0 || n > (int64_t(1) << 20) || T < 0 || k < 1 || k > kMaxSegment; }

int segments(int T, int k) { return static_cast<int>(div_up(T, k)); }

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

size_t cg_sinkhorn_checkpoint_floats(int64_t n, int iterations, int checkpoint_every) {
    if (bad_args(n, iterations, checkpoint_every)) return 0;
    return static_cast<size_t>(2) * (segments(iterations, checkpoint_every) + 1) * static_cast<size_t>(n);
}

size_t cg_sinkhorn_unrolled_workspace(int64_t n, int checkpoint_every) {
    if (n <= 0 || checkpoint_every < 1 || checkpoint_every > kMaxSegment) return 0;
    Carver cv{nullptr, 0, ~size_t(0)};
    return grad_carve(cv, n, checkpoint_every, nullptr) + 256;
}

int cg_feature_cost(const float* source_features, const float* target_features, const float* weights, float lam,
                    int64_t n, float* cost, cg_stream_t stream) {
    if (!source_features || !target_features || !weights || !cost || n <= 0 || n > (int64_t(1) << 20))
        return fail(CG_EINVAL, "cg_feature_cost: bad arguments");
    if (!aligned16(source_features) || !aligned16(target_features))
        return fail(CG_EINVAL, "cg_feature_cost: features must be 16-byte aligned");
    hipLaunchKernelGGL(feature_cost_kernel, dim3(static_cast<unsigned>(div_up(n * n, 256))), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), source_features, target_features, weights, lam,
                       static_cast<int>(n), cost);
    return check_launch("feature_cost_kernel");
}

int cg_sinkhorn_plan(const float* C, int64_t n, float regularization, const float* u, const float* v, float* plan,
                     void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    if (!C || !u || !v || !plan || n <= 0 || n > (int64_t(1) << 20) || !(regularization > 0.0f))
        return fail(CG_EINVAL, "cg_sinkhorn_plan: bad arguments");
    if (!workspace || workspace_bytes < 2 * sizeof(float) * static_cast<size_t>(n) + 512)
        return fail(CG_EWORKSPACE, "cg_sinkhorn_plan: workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    float* u2 = cv.take<float>(n);
    float* v2 = cv.take<float>(n);
    const dim3 vg(static_cast<unsigned>(div_up(n, 256)));
    hipLaunchKernelGGL(scale_vec, vg, dim3(256), 0, s, u, u2, static_cast<int>(n), kLog2e);
    hipLaunchKernelGGL(scale_vec, vg, dim3(256), 0, s, v, v2, static_cast<int>(n), kLog2e);
    hipLaunchKernelGGL(plan_kernel, dim3(static_cast<unsigned>(div_up(n * n, 256))), dim3(256), 0, s, C,
                       static_cast<int>(n), -kLog2e / regularization, u2, v2, plan);
    return check_launch("plan_kernel");
}

int cg_sinkhorn_unrolled_forward(const float* C, int64_t n, float regularization, int iterations,
                                 int checkpoint_every, const float* u0, const float* v0, float* checkpoints,
                                 float* u, float* v, void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    if (!C || !u0 || !v0 || !checkpoints || !u || !v || bad_args(n, iterations, checkpoint_every) ||
        !(regularization > 0.0f))
        return fail(CG_EINVAL, "cg_sinkhorn_unrolled_forward: bad arguments");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int k = checkpoint_every;
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    GradWs g{};
    grad_carve(cv, n, k, &g);
    if (!workspace || !cv.ok()) return fail(CG_EWORKSPACE, "cg_sinkhorn_unrolled_forward: workspace too small");
    const int ni = static_cast<int>(n);
    const float s2 = -kLog2e / regularization;
    const int S = segments(iterations, k);
    float* cu = checkpoints;                   // (S+1) x n: u at iterations 0, k, 2k, ..., T
    float* cvv = checkpoints + (S + 1) * n;    // (S+1) x n: v at the same iterations
    const dim3 vg(static_cast<unsigned>(div_up(n, 256)));
    hipLaunchKernelGGL(scale_vec, vg, dim3(256), 0, s, u0, cu, ni, kLog2e);
    hipLaunchKernelGGL(scale_vec, vg, dim3(256), 0, s, v0, cvv, ni, kLog2e);
    if (int rc = check_launch("scale_vec")) return rc;
    for (int seg = 0; seg < S; ++seg) {
        const int steps = std::min(k, iterations - seg * k);
        if (int rc = sinkhorn_segment(C, ni, s2, cu + seg * n, cvv + seg * n, steps, g.uh, g.vh, g.seg, g.seg_bytes, s))
            return rc;
        CG_CHECK_HIP(hipMemcpyAsync(cu + (seg + 1) * n, g.uh + steps * n, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
        CG_CHECK_HIP(hipMemcpyAsync(cvv + (seg + 1) * n, g.vh + steps * n, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL(scale_vec, vg, dim3(256), 0, s, cu + S * n, u, ni, kLn2);
    hipLaunchKernelGGL(scale_vec, vg, dim3(256), 0, s, cvv + S * n, v, ni, kLn2);
    return check_launch("scale_vec");
}

int cg_sinkhorn_unrolled_backward(const float* C, int64_t n, float regularization, int iterations,
                                  int checkpoint_every, const float* checkpoints, const float* grad_plan,
                                  float* grad_cost, const float* source_features, const float* target_features,
                                  const float* weights, float lam, double* grad_lam_w, void* workspace,
                                  size_t workspace_bytes, cg_stream_t stream) {
    if (!C || !checkpoints || !grad_plan || bad_args(n, iterations, checkpoint_every) || !(regularization > 0.0f))
        return fail(CG_EINVAL, "cg_sinkhorn_unrolled_backward: bad arguments");
    const bool feats = source_features != nullptr;
    if (feats && (!target_features || !weights || !grad_lam_w || !aligned16(source_features) ||
                  !aligned16(target_features)))
        return fail(CG_EINVAL, "cg_sinkhorn_unrolled_backward: incomplete feature arguments");
    if (!feats && !grad_cost) return fail(CG_EINVAL, "cg_sinkhorn_unrolled_backward: nothing to compute");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int k = checkpoint_every;
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    GradWs g{};
    grad_carve(cv, n, k, &g);
    if (!workspace || !cv.ok()) return fail(CG_EWORKSPACE, "cg_sinkhorn_unrolled_backward: workspace too small");
    const int ni = static_cast<int>(n);
    const float s2 = -kLog2e / regularization;
    const float inv_reg = 1.0f / regularization;
    const int S = segments(iterations, k);
    const float* cu = checkpoints;
    const float* cvv = checkpoints + (S + 1) * n;
    const float* uT = cu + S * n;
    const float* vT = cvv + S * n;
    if (feats) {
        hipLaunchKernelGGL(zero_doubles, dim3(1), dim3(64), 0, s, grad_lam_w, 5);
        if (int rc = check_launch("zero_doubles")) return rc;
    }
    const int gblocks = grad_blocks(n);
    // One read of C per adjoint iteration (fused row + column contraction)
    // where the three column vectors fit in LDS; else separate row and
    // column passes.
    const bool fused = fused_ok(n, C, grad_plan);
    // Adjoints of (u^T, v^T) from G = dL/dplan.  With no iterations the
    // plan depends on C only through K: dL/dK = G (.) plan.
    float* ubar_cur = g.ub;  // scratch rows, overwritten per segment below
    float* vbar_cur = g.vb;
    if (S > 0) {
        if (fused) {
            if (int rc = fused_step<true>(C, ni, s2, uT, vT, nullptr, grad_plan, nullptr, 1.0f, ubar_cur, 1, 1.0f,
                                          g.part, vbar_cur, s))
                return rc;
        } else {
            if (int rc = launch_row_c(true, C, ni, s2, uT, vT, nullptr, grad_plan, 1.0f, ubar_cur, s)) return rc;
            if (int rc = launch_col_c(true, C, ni, s2, uT, vT, ubar_cur, grad_plan, 1.0f, g.part, vbar_cur, s))
                return rc;
        }
    }
    auto grad_pass = [&](int steps, bool first) -> int {
        const size_t lds = sizeof(float) * (2 * steps * kGT + (steps + 1) * kGT + steps * kGT);
        static size_t attr_bytes = 0;  // above 64 KiB (segments longer than 63) the launch must ask
        if (lds > attr_bytes) {
            CG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&bw_grad),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
            attr_bytes = lds;
        }
        hipLaunchKernelGGL(bw_grad, dim3(static_cast<unsigned>(gblocks)), dim3(256), lds, s, C, ni, s2, inv_reg,
                           steps, g.uh, g.vh, g.ub, g.vb, first ? grad_plan : nullptr, uT, vT, grad_cost,
                           first ? 0 : 1, source_features, target_features, weights, lam, g.red);
        if (int rc = check_launch("bw_grad")) return rc;
        if (feats) {
            hipLaunchKernelGGL(bw_red_sum, dim3(1), dim3(256), 0, s, g.red, gblocks, grad_lam_w);
            return check_launch("bw_red_sum");
        }
        return CG_OK;
    };
    if (S == 0) {
        // plan = exp(K + u0 + v0): only the G (.) plan term.
        CG_CHECK_HIP(hipMemcpyAsync(g.uh, uT, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
        CG_CHECK_HIP(hipMemcpyAsync(g.vh, vT, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
        return grad_pass(0, true);
    }
    for (int seg = S - 1; seg >= 0; --seg) {
        const int c = seg * k;
        const int steps = std::min(k, iterations - c);
        if (int rc = sinkhorn_segment(C, ni, s2, cu + seg * n, cvv + seg * n, steps, g.uh, g.vh, g.seg, g.seg_bytes, s))
            return rc;
        // ubar/vbar of iteration c + steps sit in row `steps` of the adjoint history.
        if (ubar_cur != g.ub + steps * n) {
            CG_CHECK_HIP(hipMemcpyAsync(g.ub + steps * n, ubar_cur, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
            CG_CHECK_HIP(hipMemcpyAsync(g.vb + steps * n, vbar_cur, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
        }
        for (int q = steps; q >= 1; --q) {
            float* ubq1 = g.ub + (q - 1) * n;
            // ubar^{t-1}_i = -sum_j vbar^t_j 2^(K2 + u^{t-1}_i + v^t_j), and
            // vbar^{t-1}_j = -sum_i ubar^{t-1}_i 2^(K2 + u^{t-1}_i + v^{t-1}_j) for t-1 >= 1
            const int do_col = c + q - 1 >= 1 ? 1 : 0;
            if (fused) {
                if (int rc = fused_step<false>(C, ni, s2, g.uh + (q - 1) * n, g.vh + q * n, g.vb + q * n, nullptr,
                                               g.vh + (q - 1) * n, -1.0f, ubq1, do_col, -1.0f, g.part,
                                               g.vb + (q - 1) * n, s))
                    return rc;
                continue;
            }
            if (int rc = launch_row_c(false, C, ni, s2, g.uh + (q - 1) * n, g.vh + q * n, g.vb + q * n, nullptr,
                                      -1.0f, ubq1, s))
                return rc;
            if (do_col)
                if (int rc = launch_col_c(false, C, ni, s2, g.uh + (q - 1) * n, g.vh + (q - 1) * n, ubq1, nullptr,
                                          -1.0f, g.part, g.vb + (q - 1) * n, s))
                    return rc;
        }
        if (int rc = grad_pass(steps, seg == S - 1)) return rc;
        // Carry the adjoints of iteration c to the previous segment (row 0);
        // the next segment copies them into its own last row.
        ubar_cur = g.ub;
        vbar_cur = g.vb;
        if (seg > 0) {
            const int prev_steps = std::min(k, iterations - (seg - 1) * k);
            CG_CHECK_HIP(hipMemcpyAsync(g.ub + prev_steps * n, g.ub, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
            CG_CHECK_HIP(hipMemcpyAsync(g.vb + prev_steps * n, g.vb, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
            ubar_cur = g.ub + prev_steps * n;
            vbar_cur = g.vb + prev_steps * n;
        }
    }
    return CG_OK;
}

}  // extern "C"
