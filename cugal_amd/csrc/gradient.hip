// Frank-Wolfe gradient of FUGAL for undirected graphs (A, B symmetric CSR):
//   G = -2 A P B + D + f * lambda * (1 - 2 P)
// cugal/pred.py:54-74 (sparse, f = 1) and :38-51 (dense, f = 2: the
// reference adds lambda (1 - 2P) twice there).  D[x][y] is the Euclidean
// distance of the 4-feature rows, recomputed in the epilogue instead of being
// read from an n x n matrix (cuda_kernels/distance.cu:11-33 computes the same
// quantity; cugal/feature_extraction.py:98-108).
//
// Two passes, each reading and writing n x n once:
//   1. T = P B, row-local: rows of P are staged in LDS and every output
//      column j gathers P[i][l] for l in N_B(j) (B symmetric, so the CSR row
//      j lists the non-zeros of column j).
//   2. G = epilogue(-2 A T): output row i sums the rows T[k] for k in N_A(i),
//      float4 per lane; tiles are remapped so that consecutive rows of one
//      column chunk land on the same XCD and share its L2 for common
//      neighbours.
#include "common.hpp"
#include "cugal_hip.h"

namespace cg {
namespace {

constexpr int kSpThreads = 256;
constexpr int kNumXcd = 8;

// ---- pass 1: out[i][j] = sum_{l in N(j)} X[i][l] -------------------------
template <int R>
__global__ __launch_bounds__(kSpThreads) void spmm_rowlocal_lds(const int32_t* __restrict__ bcol,
                                                               const int32_t* __restrict__ brp, int n,
                                                               const float* __restrict__ X, float* __restrict__ out) {
    extern __shared__ float sx[];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * R;
    const int rows = static_cast<int>(min(static_cast<int64_t>(R), n - r0));
    for (int r = 0; r < rows; ++r) {
        const float* src = X + (r0 + r) * n;
        for (int j = threadIdx.x; j < n; j += kSpThreads) sx[r * n + j] = src[j];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += kSpThreads) {
        const int b = brp[j];
        const int e = brp[j + 1];
        float acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = 0.0f;
        for (int k = b; k < e; ++k) {
            const int l = bcol[k];
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] += sx[r * n + l];
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r < rows) out[(r0 + r) * n + j] = acc[r];
    }
}

// Same without LDS staging (rows wider than the LDS).
__global__ __launch_bounds__(kSpThreads) void spmm_rowlocal_global(const int32_t* __restrict__ bcol,
                                                                  const int32_t* __restrict__ brp, int n,
                                                                  const float* __restrict__ X,
                                                                  float* __restrict__ out) {
    const int64_t i = blockIdx.x;
    const float* src = X + i * n;
    for (int j = threadIdx.x; j < n; j += kSpThreads) {
        float acc = 0.0f;
        for (int k = brp[j]; k < brp[j + 1]; ++k) acc += src[bcol[k]];
        out[i * n + j] = acc;
    }
}

// ---- pass 2: out[i][c] = scale * sum_{k in N(i)} X[k][c] (+ epilogue) -----
struct Epilogue {
    const float* P;   // nullptr: no regulariser term
    const float* fs;  // source features n x 4 (nullptr: no distance term)
    const float* ft;  // target features n x 4
    float rl;         // f * lambda
};

__device__ __forceinline__ float feat_dist(float4 a, const float* __restrict__ ft, int j) {
    const float4 b = *reinterpret_cast<const float4*>(ft + 4 * static_cast<int64_t>(j));
    const float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
    return sqrtf(fmaf(d0, d0, fmaf(d1, d1, fmaf(d2, d2, d3 * d3))));
}

__device__ __forceinline__ float epi_one(float acc, float scale, const Epilogue& ep, int64_t idx, float4 fsi, int j) {
    float g = acc * scale;
    if (ep.fs) g += feat_dist(fsi, ep.ft, j);
    if (ep.P) g += ep.rl - 2.0f * ep.rl * ep.P[idx];
    return g;
}

// Physical block -> logical tile such that the blocks an XCD runs together
// are consecutive tiles (blocks are dealt to the XCDs round-robin).
__device__ __forceinline__ int64_t xcd_tile(int64_t bid, int64_t tiles) {
    const int64_t per = div_up(tiles, kNumXcd);
    return (bid % kNumXcd) * per + bid / kNumXcd;
}

template <bool VEC>
__global__ __launch_bounds__(kSpThreads) void spmm_rowgather(const int32_t* __restrict__ acol,
                                                            const int32_t* __restrict__ arp, int n,
                                                            const float* __restrict__ X, float scale, Epilogue ep,
                                                            float* __restrict__ out, int64_t tiles, int chunk) {
    const int64_t t = xcd_tile(blockIdx.x, tiles);
    if (t >= tiles) return;
    const int64_t i = t % n;
    const int c0 = static_cast<int>(t / n) * chunk;
    const int b = arp[i];
    const int e = arp[i + 1];
    float4 fsi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ep.fs) fsi = *reinterpret_cast<const float4*>(ep.fs + 4 * i);
    if (VEC) {
        const int j = c0 + threadIdx.x * 4;
        if (j >= n) return;
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int k = b; k < e; ++k) {
            const float4 x = *reinterpret_cast<const float4*>(X + static_cast<int64_t>(acol[k]) * n + j);
            acc.x += x.x;
            acc.y += x.y;
            acc.z += x.z;
            acc.w += x.w;
        }
        const int64_t idx = i * n + j;
        float4 o;
        o.x = epi_one(acc.x, scale, ep, idx + 0, fsi, j + 0);
        o.y = epi_one(acc.y, scale, ep, idx + 1, fsi, j + 1);
        o.z = epi_one(acc.z, scale, ep, idx + 2, fsi, j + 2);
        o.w = epi_one(acc.w, scale, ep, idx + 3, fsi, j + 3);
        *reinterpret_cast<float4*>(out + idx) = o;
    } else {
        for (int j = c0 + threadIdx.x; j < min(n, c0 + chunk); j += kSpThreads) {
            float acc = 0.0f;
            for (int k = b; k < e; ++k) acc += X[static_cast<int64_t>(acol[k]) * n + j];
            const int64_t idx = i * n + j;
            out[idx] = epi_one(acc, scale, ep, idx, fsi, j);
        }
    }
}

// ---- elementwise reference-compatible helpers ----------------------------
__global__ void add_distance_kernel(const float* __restrict__ s, const float* __restrict__ t, int n,
                                    float* __restrict__ out) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t total = static_cast<int64_t>(n) * n;
    if (idx >= total) return;
    const int64_t x = idx / n;
    const int y = static_cast<int>(idx - x * n);
    const float4 a = *reinterpret_cast<const float4*>(s + 4 * x);
    out[idx] += feat_dist(a, t, y);
}

__global__ void regularize_kernel(float* __restrict__ g, const float* __restrict__ P, int64_t total, float lam) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    g[idx] += lam - lam * 2.0f * P[idx];
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

template <int R>
int launch_rowlocal_r(const int32_t* bcol, const int32_t* brp, int n, const float* X, float* out, hipStream_t s) {
    const size_t lds = sizeof(float) * static_cast<size_t>(R) * n;
    static bool attr_set = false;
    if (!attr_set && lds > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&spmm_rowlocal_lds<R>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, max_lds_bytes());
        attr_set = true;
    }
    hipLaunchKernelGGL(spmm_rowlocal_lds<R>, dim3(static_cast<unsigned>(div_up(n, R))), dim3(kSpThreads), lds, s,
                       bcol, brp, n, X, out);
    return check_launch("spmm_rowlocal_lds");
}

int launch_rowlocal(const int32_t* bcol, const int32_t* brp, int n, const float* X, float* out, hipStream_t s) {
    const int64_t budget = max_lds_bytes();
    const int64_t row_bytes = 4 * static_cast<int64_t>(n);
    const int64_t fit = budget / row_bytes;
    if (fit >= 8) return launch_rowlocal_r<8>(bcol, brp, n, X, out, s);
    if (fit >= 4) return launch_rowlocal_r<4>(bcol, brp, n, X, out, s);
    if (fit >= 2) return launch_rowlocal_r<2>(bcol, brp, n, X, out, s);
    if (fit >= 1) return launch_rowlocal_r<1>(bcol, brp, n, X, out, s);
    hipLaunchKernelGGL(spmm_rowlocal_global, dim3(static_cast<unsigned>(n)), dim3(kSpThreads), 0, s, bcol, brp, n,
                       X, out);
    return check_launch("spmm_rowlocal_global");
}

int launch_rowgather(const int32_t* acol, const int32_t* arp, int n, const float* X, float scale, const Epilogue& ep,
                     float* out, hipStream_t s) {
    const bool vec = (n % 4 == 0) && aligned16(X) && aligned16(out) && (!ep.P || aligned16(ep.P)) &&
                     (!ep.fs || (aligned16(ep.fs) && aligned16(ep.ft)));
    const int chunk = vec ? kSpThreads * 4 : kSpThreads * 4;
    const int64_t chunks = div_up(n, chunk);
    const int64_t tiles = chunks * n;
    const int64_t blocks = div_up(tiles, kNumXcd) * kNumXcd;
    if (blocks > 0x7fffffffLL) return fail(CG_EUNSUPPORTED, "gradient grid too large (n=%d)", n);
    if (vec)
        hipLaunchKernelGGL(spmm_rowgather<true>, dim3(static_cast<unsigned>(blocks)), dim3(kSpThreads), 0, s, acol,
                           arp, n, X, scale, ep, out, tiles, chunk);
    else
        hipLaunchKernelGGL(spmm_rowgather<false>, dim3(static_cast<unsigned>(blocks)), dim3(kSpThreads), 0, s, acol,
                           arp, n, X, scale, ep, out, tiles, chunk);
    return check_launch("spmm_rowgather");
}

bool bad_n(int64_t n) { return n <= 0 || n > (int64_t(1) << 20); }

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

int cg_adjacency_matmul(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, const float* matrix,
                        float* out, int negate, cg_stream_t stream) {
    if (!col_indices || !row_pointers || !matrix || !out || bad_n(n))
        return fail(CG_EINVAL, "cg_adjacency_matmul: bad arguments");
    Epilogue ep{nullptr, nullptr, nullptr, 0.0f};
    return launch_rowgather(col_indices, row_pointers, static_cast<int>(n), matrix, negate ? -1.0f : 1.0f, ep, out,
                            reinterpret_cast<hipStream_t>(stream));
}

int cg_add_distance(const float* source, const float* target, int64_t n, float* out, cg_stream_t stream) {
    if (!source || !target || !out || bad_n(n)) return fail(CG_EINVAL, "cg_add_distance: bad arguments");
    if (!aligned16(source) || !aligned16(target)) return fail(CG_EINVAL, "cg_add_distance: features must be 16-byte aligned");
    const int64_t total = n * n;
    hipLaunchKernelGGL(add_distance_kernel, dim3(static_cast<unsigned>(div_up(total, 256))), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), source, target, static_cast<int>(n), out);
    return check_launch("add_distance_kernel");
}

int cg_regularize(float* gradient, const float* P, int64_t n, int iteration, cg_stream_t stream) {
    if (!gradient || !P || bad_n(n)) return fail(CG_EINVAL, "cg_regularize: bad arguments");
    const int64_t total = n * n;
    hipLaunchKernelGGL(regularize_kernel, dim3(static_cast<unsigned>(div_up(total, 256))), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), gradient, P, total, static_cast<float>(iteration));
    return check_launch("regularize_kernel");
}

size_t cg_gradient_workspace(int64_t n) { return n > 0 ? sizeof(float) * static_cast<size_t>(n) * n + 256 : 0; }

int cg_gradient(const int32_t* a_col, const int32_t* a_rowptr, const int32_t* b_col, const int32_t* b_rowptr,
                int64_t n, const float* P, const float* source_features, const float* target_features, int lambda_,
                float regularizer_factor, float* gradient, void* workspace, size_t workspace_bytes,
                cg_stream_t stream) {
    if (!a_col || !a_rowptr || !b_col || !b_rowptr || !P || !gradient || bad_n(n))
        return fail(CG_EINVAL, "cg_gradient: bad arguments");
    if ((source_features == nullptr) != (target_features == nullptr))
        return fail(CG_EINVAL, "cg_gradient: give both feature matrices or neither");
    if (source_features && (!aligned16(source_features) || !aligned16(target_features)))
        return fail(CG_EINVAL, "cg_gradient: feature matrices must be 16-byte aligned");
    if (!workspace || workspace_bytes < cg_gradient_workspace(n))
        return fail(CG_EWORKSPACE, "cg_gradient: workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    float* T = static_cast<float*>(workspace);
    const int ni = static_cast<int>(n);
    if (int rc = launch_rowlocal(b_col, b_rowptr, ni, P, T, s)) return rc;
    Epilogue ep{P, source_features, target_features, regularizer_factor * static_cast<float>(lambda_)};
    return launch_rowgather(a_col, a_rowptr, ni, T, -2.0f, ep, gradient, s);
}

}  // extern "C"
