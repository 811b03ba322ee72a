// Alignment-quality counts behind the benchmark metrics of
// benchmark/experiment.py:115-141.  With S, T the 0/1 adjacencies of the
// source and target graphs (rows and columns in the graphs' node order),
// `answer` the computed alignment and `truth` the ground-truth permutation,
// the reference forms adj1 = S[answer][:, answer], adj2 = T[truth][:, truth]
// and counts
//   inter  = #{(a, b) : adj1[a, b] = 1 and adj2[a, b] = 1},
//   ones_s = #{S == 1},   ones_t = #{adj2 == 1} = #{T == 1},
// from which EC = inter / ones_s, ICS = inter / ones_t and
// S3 = inter / (ones_s + ones_t - inter).  Here S and T are CSRs with sorted
// rows; inter is summed over the distinct non-zeros (x, y) of T as
// [answer[b] in row answer[a] of S] with a = truth^-1(x), b = truth^-1(y),
// one wave per row and a binary search per entry.  Duplicate column entries
// (a self-loop listed twice) count once, as in the dense matrices.
#include "common.hpp"
#include "cugal_hip.h"

namespace cg {
namespace {

constexpr int kMetricWaves = 4;

__global__ void fill_minus_one(int32_t* __restrict__ p, int64_t count) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) p[i] = -1;
}

// inv[truth[a]] = a; flags[0] |= 2 if truth is not a permutation of [0, n_t);
// flags[0] |= 1 if an answer lies outside [-n_s, n_s) (numpy's valid indices);
// answers are normalised into [0, n_s) in `ans`.
__global__ void prepare_maps(const int32_t* __restrict__ truth, const int32_t* __restrict__ answer, int64_t m,
                             int n_s, int n_t, int32_t* __restrict__ inv, int32_t* __restrict__ ans,
                             unsigned long long* __restrict__ flags) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a >= m) return;
    const int x = truth[a];
    if (x < 0 || x >= n_t || atomicExch(&inv[x], static_cast<int32_t>(a)) != -1) atomicOr(&flags[0], 2ull);
    int s = answer[a];
    if (s < 0) s += n_s;
    if (s < 0 || s >= n_s) {
        atomicOr(&flags[0], 1ull);
        s = 0;
    }
    ans[a] = s;
}

__device__ __forceinline__ bool row_has(const int32_t* __restrict__ col, int b, int e, int y) {
    while (b < e) {
        const int mid = (b + e) >> 1;
        const int c = col[mid];
        if (c == y) return true;
        if (c < y) b = mid + 1;
        else e = mid;
    }
    return false;
}

// counts[0] += inter, counts[1] += ones_s, counts[2] += ones_t.
__global__ __launch_bounds__(kMetricWaves* kWave) void metric_counts(
    const int32_t* __restrict__ s_col, const int32_t* __restrict__ s_rp, int n_s, const int32_t* __restrict__ t_col,
    const int32_t* __restrict__ t_rp, int n_t, const int32_t* __restrict__ inv, const int32_t* __restrict__ ans,
    unsigned long long* __restrict__ counts) {
    const int lane = threadIdx.x & (kWave - 1);
    const int row = blockIdx.x * kMetricWaves + threadIdx.x / kWave;
    int inter = 0, os = 0, ot = 0;
    if (row < n_s) {
        const int b = s_rp[row], e = s_rp[row + 1];
        for (int k = b + lane; k < e; k += kWave) os += (k == b || s_col[k] != s_col[k - 1]) ? 1 : 0;
    }
    if (row < n_t) {
        const int b = t_rp[row], e = t_rp[row + 1];
        const int a = inv[row];
        const int sa = a >= 0 ? ans[a] : 0;
        const int sb0 = a >= 0 ? s_rp[sa] : 0;
        const int se0 = a >= 0 ? s_rp[sa + 1] : 0;
        // Only entries whose row and column are ground-truth images appear
        // in adj2 (all of T when truth is a permutation of its nodes).
        for (int k = b + lane; k < e && a >= 0; k += kWave) {
            const int y = t_col[k];
            if (k != b && y == t_col[k - 1]) continue;
            const int bb = inv[y];
            if (bb < 0) continue;
            ++ot;
            if (row_has(s_col, sb0, se0, ans[bb])) ++inter;
        }
    }
    inter = wave_sum(inter);
    os = wave_sum(os);
    ot = wave_sum(ot);
    if (lane == 0) {
        if (inter) atomicAdd(&counts[0], static_cast<unsigned long long>(inter));
        if (os) atomicAdd(&counts[1], static_cast<unsigned long long>(os));
        if (ot) atomicAdd(&counts[2], static_cast<unsigned long long>(ot));
    }
}

__global__ void zero_u64(unsigned long long* __restrict__ p, int count) {
    const int i = threadIdx.x;
    if (i < count) p[i] = 0ull;
}

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

size_t cg_alignment_counts_workspace(int64_t n_target, int64_t m) {
    if (n_target < 0 || m < 0) return 0;
    Carver cv{nullptr, 0, ~size_t(0)};
    cv.take<int32_t>(static_cast<size_t>(n_target) + 1);
    cv.take<int32_t>(static_cast<size_t>(m) + 1);
    return cv.used + 256;
}

int cg_alignment_counts(const int32_t* s_col, const int32_t* s_rowptr, int64_t n_source, const int32_t* t_col,
                        const int32_t* t_rowptr, int64_t n_target, const int32_t* answer, const int32_t* truth,
                        int64_t m, int64_t* counts, void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    if (!s_rowptr || !t_rowptr || !counts || n_source < 0 || n_target < 0 || m < 0 ||
        n_source > (int64_t(1) << 30) || n_target > (int64_t(1) << 30) || m > (int64_t(1) << 30) ||
        (m > 0 && (!answer || !truth)))
        return fail(CG_EINVAL, "cg_alignment_counts: bad arguments");
    if (!workspace || workspace_bytes < cg_alignment_counts_workspace(n_target, m))
        return fail(CG_EWORKSPACE, "cg_alignment_counts: workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    int32_t* inv = cv.take<int32_t>(static_cast<size_t>(n_target) + 1);
    int32_t* ans = cv.take<int32_t>(static_cast<size_t>(m) + 1);
    unsigned long long* c = reinterpret_cast<unsigned long long*>(counts);
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, s, c, 4);
    if (int rc = check_launch("zero_u64")) return rc;
    if (n_target > 0) {
        hipLaunchKernelGGL(fill_minus_one, dim3(static_cast<unsigned>(div_up(n_target, 256))), dim3(256), 0, s, inv,
                           n_target);
        if (int rc = check_launch("fill_minus_one")) return rc;
    }
    if (m > 0) {
        hipLaunchKernelGGL(prepare_maps, dim3(static_cast<unsigned>(div_up(m, 256))), dim3(256), 0, s, truth, answer, m,
                           static_cast<int>(n_source), static_cast<int>(n_target), inv, ans, c + 3);
        if (int rc = check_launch("prepare_maps")) return rc;
    }
    const int64_t rows = std::max(n_source, n_target);
    if (rows > 0) {
        hipLaunchKernelGGL(metric_counts, dim3(static_cast<unsigned>(div_up(rows, kMetricWaves))),
                           dim3(kMetricWaves * kWave), 0, s, s_col, s_rowptr, static_cast<int>(n_source), t_col,
                           t_rowptr, static_cast<int>(n_target), inv, ans, c);
        if (int rc = check_launch("metric_counts")) return rc;
    }
    return CG_OK;
}

}  // extern "C"
