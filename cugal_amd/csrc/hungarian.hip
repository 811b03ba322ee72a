// Exact linear assignment on the device: replaces dense_hungarian
// (cuda_kernels/hungarian/hungarian.cu:299-448, bound at
// cuda_kernels/kernels.cpp:69) and the cugraph dense Hungarian behind
// HungarianMethod.DENSE (cugal/pred.py:161-166).
//
// The reference runs the Munkres star/prime/cover steps over an n x n slack
// matrix, one full-matrix pass per step.  Here the solver is the shortest
// augmenting path method with FP64 dual potentials (u for rows, v for
// columns; reduced cost c_ij - u_i - v_j >= 0, 0 on assigned pairs):
//   1. col_min: v_j = min_i c_ij (parallel over the matrix);
//   2. row_min: u_i = min_j (c_ij - v_j) and its first-index argmin, one
//      wave per row; every row whose argmin column no lower row claims is
//      assigned there (a zero reduced cost), deterministically;
//   3. sap_solve: one block of 1024 threads takes the free rows in index
//      order.  For each it runs Dijkstra over the columns on reduced costs:
//      each thread keeps the tentative distances of its columns (in
//      registers up to n = 8192, else in a global array only it touches),
//      relaxes them against the newly reached row (one coalesced read of
//      that row) and a block-wide lexicographic minimum picks the next
//      column (distance, then free columns first, then lowest index).
//      On reaching a free column the duals of the scanned columns and rows
//      move by (final distance - their distance) and the path is flipped.
// Each Dijkstra step scans one new column, so a row is done in at most n
// steps; a row with no finite path (NaN or infinite costs) stops the solve
// with an error.  Maximisation negates the costs, which is exact.
#include "common.hpp"
#include "cugal_hip.h"

namespace cg {
namespace {

constexpr int kSapThreads = 1024;
constexpr int kSapMaxQ = 32;  // columns per thread: n <= 32768
constexpr int kSapRegQ = 8;   // distances in registers up to 8 per thread
constexpr int kColMinChunk = 256;  // rows per col_min block
constexpr int kNoRow = 0x7fffffff;

__global__ __launch_bounds__(256) void col_min_partial(const float* __restrict__ C, int n, float sign,
                                                      float* __restrict__ part) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int r0 = blockIdx.y * kColMinChunk;
    const int r1 = min(n, r0 + kColMinChunk);
    float m = INFINITY;
    bool nan = false;
    for (int i = r0; i < r1; ++i) {
        const float c = sign * C[static_cast<int64_t>(i) * n + j];
        nan |= (c != c);
        m = fminf(m, c);
    }
    part[static_cast<int64_t>(blockIdx.y) * n + j] = nan ? NAN : m;
}

__global__ __launch_bounds__(256) void col_min_final(const float* __restrict__ part, int n, int chunks,
                                                    double* __restrict__ v, int32_t* __restrict__ row4col,
                                                    int32_t* __restrict__ status) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    float m = INFINITY;
    bool bad = false;
    for (int k = 0; k < chunks; ++k) {
        const float x = part[static_cast<int64_t>(k) * n + j];
        bad |= (x != x);
        m = fminf(m, x);
    }
    bad |= !(m < INFINITY) || !(m > -INFINITY);  // NaN, all-+inf or a -inf entry
    v[j] = static_cast<double>(m);
    row4col[j] = kNoRow;
    if (bad) *status = 1;
}

__global__ __launch_bounds__(256) void row_min(const float* __restrict__ C, int n, float sign,
                                              const double* __restrict__ v, double* __restrict__ u,
                                              int32_t* __restrict__ arg, int32_t* __restrict__ row4col) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / kWave;
    if (i >= n) return;
    const float* row = C + i * n;
    double best = INFINITY;
    int bj = kNoRow;
    for (int j = lane; j < n; j += kWave) {
        const double r = static_cast<double>(sign * row[j]) - v[j];
        if (r < best) {  // lanes see increasing j: strict keeps the first
            best = r;
            bj = j;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off, kWave);
        const int oj = __shfl_xor(bj, off, kWave);
        if (ob < best || (ob == best && oj < bj)) {
            best = ob;
            bj = oj;
        }
    }
    if (lane == 0) {
        u[i] = best;
        arg[i] = bj;
        if (bj != kNoRow) atomicMin(&row4col[bj], static_cast<int>(i));
    }
}

// The lowest row claiming a column gets it; the other rows stay free.
__global__ __launch_bounds__(256) void claim_rows(const int32_t* __restrict__ arg, int n,
                                                 const int32_t* __restrict__ row4col, int32_t* __restrict__ col4row) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int j = arg[i];
    col4row[i] = (j != kNoRow && row4col[j] == i) ? j : -1;
}

__global__ __launch_bounds__(256) void unclaimed_to_free(int32_t* __restrict__ row4col, int n) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < n && row4col[j] == kNoRow) row4col[j] = -1;
}

// Tentative distances of a thread's columns: registers (R) or the global
// array `dist` at the column's index (each column has one owner thread, so
// no synchronisation is involved either way).
template <int Q, bool R>
struct Dist {
    double reg[R ? Q : 1];
    double* glob;
    __device__ __forceinline__ double get(int q, int j) const {
        if constexpr (R) {
            (void)j;
            return reg[q];
        } else {
            (void)q;
            return glob[j];
        }
    }
    __device__ __forceinline__ void set(int q, int j, double x) {
        if constexpr (R) {
            (void)j;
            reg[q] = x;
        } else {
            (void)q;
            glob[j] = x;
        }
    }
};

template <int Q, bool R>
__global__ __launch_bounds__(kSapThreads) void sap_solve(const float* __restrict__ C, int n, float sign,
                                                        double* __restrict__ u, double* __restrict__ v,
                                                        int32_t* __restrict__ row4col, int32_t* __restrict__ col4row,
                                                        int32_t* __restrict__ path, int32_t* __restrict__ free_rows,
                                                        double* __restrict__ dist, int32_t* __restrict__ status,
                                                        int32_t* __restrict__ stats) {
    __shared__ double red_d[2][kSapThreads / kWave];
    __shared__ int red_k[2][kSapThreads / kWave];
    __shared__ int counts[kSapThreads];
    __shared__ int s_nfree;
    constexpr int kUnroll = R ? Q : 8;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;

    // Free rows in index order (a stable block compaction).
    {
        const int per = static_cast<int>(div_up(n, kSapThreads));
        const int b = min(n, tid * per);
        const int e = min(n, b + per);
        int c = 0;
        for (int i = b; i < e; ++i) c += col4row[i] < 0;
        counts[tid] = c;
        __syncthreads();
        if (tid == 0) {
            int run = 0;
            for (int t = 0; t < kSapThreads; ++t) {
                const int x = counts[t];
                counts[t] = run;
                run += x;
            }
            s_nfree = run;
        }
        __syncthreads();
        int pos = counts[tid];
        for (int i = b; i < e; ++i)
            if (col4row[i] < 0) free_rows[pos++] = i;
        __syncthreads();
    }
    const int nfree = s_nfree;

    uint32_t freebits = 0;
#pragma unroll kUnroll
    for (int q = 0; q < Q; ++q) {
        const int j = tid + q * kSapThreads;
        if (j < n && row4col[j] < 0) freebits |= 1u << q;
    }
    Dist<Q, R> d;
    d.glob = dist;
    int it = 0;
    long long steps = 0;
    bool failed = false;
    for (int f = 0; f < nfree && !failed; ++f) {
        const int i0 = free_rows[f];
#pragma unroll kUnroll
        for (int q = 0; q < Q; ++q) {
            const int j = tid + q * kSapThreads;
            if (R || j < n) d.set(q, j, INFINITY);
        }
        uint32_t scanned = 0;
        double min_val = 0.0;
        int i = i0;
        int sink = -1;
        for (int cnt = 0; cnt <= n; ++cnt) {
            const double ui = u[i];
            const float* row = C + static_cast<int64_t>(i) * n;
            double bd = INFINITY;
            int bk = kNoRow;
#pragma unroll kUnroll
            for (int q = 0; q < Q; ++q) {
                const int j = tid + q * kSapThreads;
                if (j < n && !((scanned >> q) & 1u)) {
                    const double r = min_val + static_cast<double>(sign * row[j]) - ui - v[j];
                    double dq = d.get(q, j);
                    if (r < dq) {
                        dq = r;
                        d.set(q, j, r);
                        path[j] = i;
                    }
                    const int key = (((freebits >> q) & 1u) ? 0 : (1 << 30)) | j;
                    if (dq < bd || (dq == bd && key < bk)) {
                        bd = dq;
                        bk = key;
                    }
                }
            }
#pragma unroll kUnroll
            for (int off = 32; off > 0; off >>= 1) {
                const double od = __shfl_xor(bd, off, kWave);
                const int ok = __shfl_xor(bk, off, kWave);
                if (od < bd || (od == bd && ok < bk)) {
                    bd = od;
                    bk = ok;
                }
            }
            if (lane == 0) {
                red_d[it & 1][wave] = bd;
                red_k[it & 1][wave] = bk;
            }
            __syncthreads();
            bd = red_d[it & 1][0];
            bk = red_k[it & 1][0];
            for (int w = 1; w < kSapThreads / kWave; ++w) {
                const double od = red_d[it & 1][w];
                const int ok = red_k[it & 1][w];
                if (od < bd || (od == bd && ok < bk)) {
                    bd = od;
                    bk = ok;
                }
            }
            ++it;
            ++steps;
            if (!(bd < INFINITY) || bk == kNoRow) {
                failed = true;
                break;
            }
            min_val = bd;
            const int js = bk & ((1 << 30) - 1);
            if ((js % kSapThreads) == tid) scanned |= 1u << (js / kSapThreads);
            if (!(bk >> 30)) {
                sink = js;
                break;
            }
            i = row4col[js];
        }
        if (sink < 0) {
            failed = true;
            break;
        }
        // Dual update over the scanned columns and the rows reached through
        // them, then flip the path.
#pragma unroll kUnroll
        for (int q = 0; q < Q; ++q) {
            if ((scanned >> q) & 1u) {
                const int j = tid + q * kSapThreads;
                const double delta = min_val - d.get(q, j);
                v[j] -= delta;
                if (j != sink) u[row4col[j]] += delta;
            }
        }
        if (tid == 0) u[i0] += min_val;
        __syncthreads();
        if (tid == 0) {
            int j = sink;
            for (int hop = 0; hop <= n; ++hop) {
                const int r = path[j];
                row4col[j] = r;
                const int prev = col4row[r];
                col4row[r] = j;
                j = prev;
                if (r == i0) break;
            }
        }
        if ((sink % kSapThreads) == tid) freebits &= ~(1u << (sink / kSapThreads));
        __syncthreads();
    }
    if (tid == 0) {
        if (failed) *status = 2;
        if (stats) {
            stats[0] = nfree;
            stats[1] = static_cast<int32_t>(steps > 0x7fffffffLL ? 0x7fffffffLL : steps);
        }
    }
}

template <int Q>
void launch_sap(hipStream_t s, const float* C, int n, float sign, double* u, double* v, int32_t* row4col,
                int32_t* col4row, int32_t* path, int32_t* free_rows, double* dist, int32_t* status, int32_t* stats) {
    constexpr bool R = Q <= kSapRegQ;
    hipLaunchKernelGGL((sap_solve<Q, R>), dim3(1), dim3(kSapThreads), 0, s, C, n, sign, u, v, row4col, col4row,
                       path, free_rows, dist, status, stats);
}

struct HungWs {
    float* part;
    double* u;
    double* v;
    int32_t* arg;
    int32_t* row4col;
    int32_t* path;
    int32_t* free_rows;
    double* dist;
    int32_t* status;
};

HungWs hung_carve(char* base, int64_t n, size_t bytes, size_t* used = nullptr) {
    Carver cv{base, 0, bytes};
    HungWs w;
    w.part = cv.take<float>(div_up(n, kColMinChunk) * n);
    w.u = cv.take<double>(n);
    w.v = cv.take<double>(n);
    w.arg = cv.take<int32_t>(n);
    w.row4col = cv.take<int32_t>(n);
    w.path = cv.take<int32_t>(n);
    w.free_rows = cv.take<int32_t>(n);
    w.dist = cv.take<double>(n);
    w.status = cv.take<int32_t>(1);
    if (used) *used = cv.used;
    return w;
}

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

size_t cg_lap_hungarian_workspace(int64_t n) {
    if (n <= 0) return 0;
    size_t used = 0;
    (void)hung_carve(nullptr, n, ~size_t(0), &used);
    return used + 256;
}

int cg_lap_hungarian(const float* cost, int64_t n, int maximize, int32_t* assignment, int32_t* stats,
                     void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    if (!cost || !assignment || n <= 0) return fail(CG_EINVAL, "cg_lap_hungarian: bad arguments");
    if (n > int64_t(kSapThreads) * kSapMaxQ)
        return fail(CG_EUNSUPPORTED, "cg_lap_hungarian: n=%lld exceeds %d", static_cast<long long>(n),
                    kSapThreads * kSapMaxQ);
    if (!workspace || workspace_bytes < cg_lap_hungarian_workspace(n))
        return fail(CG_EWORKSPACE, "cg_lap_hungarian: workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    HungWs w = hung_carve(static_cast<char*>(workspace), n, workspace_bytes);
    const int ni = static_cast<int>(n);
    const float sign = maximize ? -1.0f : 1.0f;
    const unsigned vb = static_cast<unsigned>(div_up(n, 256));
    const unsigned chunks = static_cast<unsigned>(div_up(n, kColMinChunk));
    CG_CHECK_HIP(hipMemsetAsync(w.status, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(col_min_partial, dim3(vb, chunks), dim3(256), 0, s, cost, ni, sign, w.part);
    if (int rc = check_launch("col_min_partial")) return rc;
    hipLaunchKernelGGL(col_min_final, dim3(vb), dim3(256), 0, s, w.part, ni, static_cast<int>(chunks), w.v,
                       w.row4col, w.status);
    if (int rc = check_launch("col_min_final")) return rc;
    hipLaunchKernelGGL(row_min, dim3(static_cast<unsigned>(div_up(n, 4))), dim3(256), 0, s, cost, ni, sign, w.v, w.u,
                       w.arg, w.row4col);
    if (int rc = check_launch("row_min")) return rc;
    hipLaunchKernelGGL(claim_rows, dim3(vb), dim3(256), 0, s, w.arg, ni, w.row4col, assignment);
    if (int rc = check_launch("claim_rows")) return rc;
    hipLaunchKernelGGL(unclaimed_to_free, dim3(vb), dim3(256), 0, s, w.row4col, ni);
    if (int rc = check_launch("unclaimed_to_free")) return rc;
    // A NaN or infinite column stops here, before the solve.
    int status = 0;
    CG_CHECK_HIP(hipMemcpyAsync(&status, w.status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CG_CHECK_HIP(hipStreamSynchronize(s));
    if (status) return fail(CG_EINVAL, "cg_lap_hungarian: the cost matrix has a NaN or an all-infinite column");
    const int q = static_cast<int>(div_up(n, kSapThreads));
    if (q <= 1)
        launch_sap<1>(s, cost, ni, sign, w.u, w.v, w.row4col, assignment, w.path, w.free_rows, w.dist, w.status, stats);
    else if (q <= 2)
        launch_sap<2>(s, cost, ni, sign, w.u, w.v, w.row4col, assignment, w.path, w.free_rows, w.dist, w.status, stats);
    else if (q <= 4)
        launch_sap<4>(s, cost, ni, sign, w.u, w.v, w.row4col, assignment, w.path, w.free_rows, w.dist, w.status, stats);
    else if (q <= 8)
        launch_sap<8>(s, cost, ni, sign, w.u, w.v, w.row4col, assignment, w.path, w.free_rows, w.dist, w.status, stats);
    else if (q <= 16)
        launch_sap<16>(s, cost, ni, sign, w.u, w.v, w.row4col, assignment, w.path, w.free_rows, w.dist, w.status, stats);
    else
        launch_sap<32>(s, cost, ni, sign, w.u, w.v, w.row4col, assignment, w.path, w.free_rows, w.dist, w.status, stats);
    if (int rc = check_launch("sap_solve")) return rc;
    CG_CHECK_HIP(hipMemcpyAsync(&status, w.status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CG_CHECK_HIP(hipStreamSynchronize(s));
    if (status) return fail(CG_EINVAL, "cg_lap_hungarian: no finite augmenting path (infinite costs)");
    return CG_OK;
}

}  // extern "C"
