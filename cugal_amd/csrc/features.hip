// Graph construction and structural node features.
//
// CSR build (replaces cuda_kernels/adjacency.cu:152-259): degree histogram,
// single-block exclusive scan, atomic scatter, then a per-row rank sort, so
// the columns of each row come out ascending whatever order the scatter
// produced; rows with no entries get empty ranges.
//
// Features (replaces cuda_kernels/feature_extraction.cu:6-125 and the
// networkx features of cugal/feature_extraction.py:41-64):
//   t(v)   = sum_{u in N(v)} |N(v) & N(u)|   (= 2 * triangles through v)
//   deg(v) = |N(v)|,  clustering c(v) = t / (d (d - 1))  (0 if t = 0 or d < 2)
//   mean neighbour degree / clustering over N(v) (0 if d = 0).
// Integer sums are exact; the ratios are formed in FP64 and rounded once to
// FP32, i.e. they equal the FP32 rounding of networkx's FP64 values.
#include "common.hpp"
#include "cugal_hip.h"

namespace cg {
namespace {

constexpr int kScanThreads = 1024;

__global__ void zero_ints(int32_t* __restrict__ p, int64_t count) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

__global__ void count_degrees(const int32_t* __restrict__ edges, int64_t m, int n, int sym, int32_t* __restrict__ cnt,
                              int* __restrict__ bad) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int x = edges[2 * e];
    const int y = edges[2 * e + 1];
    if (x < 0 || x >= n || y < 0 || y >= n) {
        atomicExch(bad, 1);
        return;
    }
    atomicAdd(cnt + y, 1);
    if (sym) atomicAdd(cnt + x, 1);
}

// Exclusive scan of cnt[0..n) into rp[0..n]; one block, sequential chunks.
__global__ __launch_bounds__(kScanThreads) void scan_single_block(const int32_t* __restrict__ cnt, int n,
                                                                 int32_t* __restrict__ rp) {
    __shared__ int64_t part[kScanThreads];
    const int per = static_cast<int>(div_up(n, kScanThreads));
    const int b = threadIdx.x * per;
    const int e = min(n, b + per);
    int64_t s = 0;
    for (int i = b; i < e; ++i) s += cnt[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (int k = 0; k < kScanThreads; ++k) {
            const int64_t v = part[k];
            part[k] = run;
            run += v;
        }
    }
    __syncthreads();
    int64_t run = part[threadIdx.x];
    for (int i = b; i < e; ++i) {
        rp[i] = static_cast<int32_t>(run);
        run += cnt[i];
    }
    if (threadIdx.x == kScanThreads - 1) rp[n] = static_cast<int32_t>(run);
}

__global__ void scatter_edges(const int32_t* __restrict__ edges, int64_t m, int sym, int32_t* __restrict__ cursor,
                              int32_t* __restrict__ tcol, int32_t* __restrict__ trow) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int x = edges[2 * e];
    const int y = edges[2 * e + 1];
    int pos = atomicAdd(cursor + y, 1);
    tcol[pos] = x;
    trow[pos] = y;
    if (sym) {
        pos = atomicAdd(cursor + x, 1);
        tcol[pos] = y;
        trow[pos] = x;
    }
}

// Stable rank within the row: #smaller + #equal at earlier positions.
__global__ void rank_sort_rows(const int32_t* __restrict__ tcol, const int32_t* __restrict__ trow,
                               const int32_t* __restrict__ rp, int64_t nnz, int32_t* __restrict__ col) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const int r = trow[p];
    const int b = rp[r];
    const int e = rp[r + 1];
    const int val = tcol[p];
    int rank = 0;
    for (int k = b; k < e; ++k) {
        const int w = tcol[k];
        rank += (w < val) || (w == val && k < p);
    }
    col[b + rank] = val;
}

// |sorted a & sorted b| by merging.
__device__ __forceinline__ int intersect_count(const int32_t* __restrict__ a, int na, const int32_t* __restrict__ b,
                                               int nb) {
    int i = 0, j = 0, c = 0;
    while (i < na && j < nb) {
        const int x = a[i];
        const int y = b[j];
        c += (x == y);
        i += (x <= y);
        j += (y <= x);
    }
    return c;
}

// One wave per vertex; lanes split the neighbour list.
__global__ __launch_bounds__(256) void vertex_triangles(const int32_t* __restrict__ col, const int32_t* __restrict__ rp,
                                                       int n, int32_t* __restrict__ tri, float* __restrict__ clustering,
                                                       float* __restrict__ degrees) {
    const int lane = threadIdx.x & 63;
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / 64;
    if (v >= n) return;
    const int b = rp[v];
    const int e = rp[v + 1];
    int t = 0;
    for (int k = b + lane; k < e; k += 64) {
        const int u = col[k];
        t += intersect_count(col + b, e - b, col + rp[u], rp[u + 1] - rp[u]);
    }
    t = wave_sum(t);
    if (lane == 0) {
        const int d = e - b;
        if (tri) tri[v] = t;
        if (degrees) degrees[v] = static_cast<float>(d);
        if (clustering)
            clustering[v] = (t == 0 || d < 2)
                                ? 0.0f
                                : static_cast<float>(static_cast<double>(t) /
                                                     (static_cast<double>(d) * static_cast<double>(d - 1)));
    }
}

// Common neighbours of every CSR entry (v, u): |N(v) & N(u)|, one wave per
// vertex, lanes over its neighbours.
__global__ __launch_bounds__(256) void entry_common_neighbours(const int32_t* __restrict__ col,
                                                              const int32_t* __restrict__ rp, int n,
                                                              int32_t* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / 64;
    if (v >= n) return;
    const int b = rp[v];
    const int e = rp[v + 1];
    for (int k = b + lane; k < e; k += 64) {
        const int u = col[k];
        out[k] = intersect_count(col + b, e - b, col + rp[u], rp[u + 1] - rp[u]);
    }
}

__device__ __forceinline__ double clustering_of(const int32_t* __restrict__ rp, const int32_t* __restrict__ tri,
                                                int u) {
    const int d = rp[u + 1] - rp[u];
    const int t = tri[u];
    return (t == 0 || d < 2) ? 0.0 : static_cast<double>(t) / (static_cast<double>(d) * static_cast<double>(d - 1));
}

// Mean neighbour degree and clustering from exact integers; writes the
// n x 4 feature rows scaled by mu.
__global__ __launch_bounds__(256) void vertex_features4(const int32_t* __restrict__ col, const int32_t* __restrict__ rp,
                                                       const int32_t* __restrict__ tri, int n, float mu,
                                                       float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / 64;
    if (v >= n) return;
    const int b = rp[v];
    const int e = rp[v + 1];
    double sd = 0.0, sc = 0.0;
    for (int k = b + lane; k < e; k += 64) {
        const int u = col[k];
        sd += static_cast<double>(rp[u + 1] - rp[u]);
        sc += clustering_of(rp, tri, u);
    }
    sd = wave_sum(sd);
    sc = wave_sum(sc);
    if (lane == 0) {
        const int d = e - b;
        const float deg = static_cast<float>(d);
        const float clus = static_cast<float>(clustering_of(rp, tri, static_cast<int>(v)));
        const float nd = d == 0 ? 0.0f : static_cast<float>(sd / d);
        const float nc = d == 0 ? 0.0f : static_cast<float>(sc / d);
        *reinterpret_cast<float4*>(out + 4 * v) = make_float4(deg * mu, clus * mu, nd * mu, nc * mu);
    }
}

// averages[v] = sum_{u in N(v)} f[u] / d(v), FP64 accumulation.
__global__ __launch_bounds__(256) void vertex_average(const int32_t* __restrict__ col, const int32_t* __restrict__ rp,
                                                     int n, const float* __restrict__ f, float* __restrict__ avg) {
    const int lane = threadIdx.x & 63;
    const int64_t v = static_cast<int64_t>(blockIdx.x) * 4 + threadIdx.x / 64;
    if (v >= n) return;
    const int b = rp[v];
    const int e = rp[v + 1];
    double s = 0.0;
    for (int k = b + lane; k < e; k += 64) s += static_cast<double>(f[col[k]]);
    s = wave_sum(s);
    if (lane == 0) avg[v] = (e == b) ? 0.0f : static_cast<float>(s / (e - b));
}

__global__ void egonet_kernel(const int32_t* __restrict__ rp, const int32_t* __restrict__ tri, int n,
                              int32_t* __restrict__ out) {
    const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    out[v] = (rp[v + 1] - rp[v]) + tri[v] / 2;
}


// The three extra node features of FUGAL's 7-feature set (FUGAL/pred.py:
// 24-110 with simple=False) for a simple graph, next to the 4 of
// vertex_features4 (already scaled by mu), as 8-float rows
// (deg, clustering, mean neighbour degree, mean neighbour clustering,
//  ego-net edges, edges leaving the ego net, distinct nodes at distance 2, 0):
//   ego(v)  = deg(v) + t(v)                  (t = triangles at v),
//   out(v)  = sum_{u in N(v)} deg(u) - deg(v) - 2 t(v)
//             (each ego-net edge is counted twice by the degree sum of the
//              closed neighbourhood, each leaving edge once),
//   hop2(v) = |N(N(v)) \ N[v]|              (0 for an isolated vertex).
// One block per vertex (grid-stride); two bitmaps of n bits in LDS hold the
// closed neighbourhood and the distance-2 set, which is counted at first
// insertion; both are cleared by revisiting what was set, so a vertex costs
// O(sum of its neighbours' degrees).
constexpr int kFugalThreads = 256;

__global__ __launch_bounds__(kFugalThreads) void fugal_extra(const int32_t* __restrict__ col,
                                                            const int32_t* __restrict__ rp,
                                                            const int32_t* __restrict__ tri,
                                                            const float* __restrict__ f4, int n, float mu,
                                                            float* __restrict__ out8) {
    extern __shared__ uint32_t bits[];
    __shared__ int wsum[2][kFugalThreads / kWave];
    const int words = (n + 31) >> 5;
    uint32_t* nb = bits;          // closed neighbourhood
    uint32_t* hb = bits + words;  // distance-2 set
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    for (int w = tid; w < 2 * words; w += kFugalThreads) bits[w] = 0u;
    __syncthreads();
    for (int v = blockIdx.x; v < n; v += gridDim.x) {
        const int b = rp[v];
        const int e = rp[v + 1];
        int s1 = 0;
        for (int k = b + tid; k < e; k += kFugalThreads) {
            const int u = col[k];
            atomicOr(&nb[u >> 5], 1u << (u & 31));
            s1 += rp[u + 1] - rp[u];
        }
        if (tid == 0) atomicOr(&nb[v >> 5], 1u << (v & 31));
        __syncthreads();
        int hop = 0;
        for (int k = b; k < e; ++k) {
            const int u = col[k];
            const int ue = rp[u + 1];
            for (int q = rp[u] + tid; q < ue; q += kFugalThreads) {
                const int x = col[q];
                const uint32_t bit = 1u << (x & 31);
                if (!(nb[x >> 5] & bit) && !(atomicOr(&hb[x >> 5], bit) & bit)) ++hop;
            }
        }
        s1 = wave_sum(s1);
        hop = wave_sum(hop);
        if (lane == 0) {
            wsum[0][wave] = s1;
            wsum[1][wave] = hop;
        }
        __syncthreads();  // every read of nb and insertion into hb is done
        for (int k = b; k < e; ++k) {
            const int u = col[k];
            const int ue = rp[u + 1];
            for (int q = rp[u] + tid; q < ue; q += kFugalThreads) hb[col[q] >> 5] = 0u;
        }
        for (int k = b + tid; k < e; k += kFugalThreads) nb[col[k] >> 5] = 0u;
        if (tid == 0) {
            nb[v >> 5] = 0u;
            int S1 = 0, H = 0;
            for (int w = 0; w < kFugalThreads / kWave; ++w) {
                S1 += wsum[0][w];
                H += wsum[1][w];
            }
            const int d = e - b;
            const int t = tri[v] / 2;
            const float4 a = *reinterpret_cast<const float4*>(f4 + 4 * static_cast<int64_t>(v));
            float* o = out8 + 8 * static_cast<int64_t>(v);
            *reinterpret_cast<float4*>(o) = a;
            *reinterpret_cast<float4*>(o + 4) =
                d == 0 ? make_float4(0.f, 0.f, 0.f, 0.f)
                       : make_float4(mu * static_cast<float>(d + t), mu * static_cast<float>(S1 - d - 2 * t),
                                     mu * static_cast<float>(H), 0.0f);
        }
        __syncthreads();  // bitmaps clear and wsum read before the next vertex
    }
}

bool bad_n(int64_t n) { return n <= 0 || n >= (int64_t(1) << 30); }

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

size_t cg_csr_build_workspace(int64_t edge_count, int64_t node_count, int symmetrize) {
    const int64_t nnz = edge_count * (symmetrize ? 2 : 1);
    Carver cv{nullptr, 0, ~size_t(0)};
    cv.take<int32_t>(node_count + 1);  // counts / cursor
    cv.take<int32_t>(nnz);             // unsorted columns
    cv.take<int32_t>(nnz);             // row of each entry
    cv.take<int>(1);                   // bad-edge flag
    return cv.used + 256;
}

int cg_csr_build(const int32_t* edges, int64_t edge_count, int64_t node_count, int symmetrize, int32_t* col_indices,
                 int32_t* row_pointers, void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    if (bad_n(node_count) || edge_count < 0 || !row_pointers || (edge_count > 0 && (!edges || !col_indices)))
        return fail(CG_EINVAL, "cg_csr_build: bad arguments");
    const int64_t nnz = edge_count * (symmetrize ? 2 : 1);
    if (nnz >= (int64_t(1) << 31)) return fail(CG_EUNSUPPORTED, "cg_csr_build: too many entries for int32 CSR");
    if (!workspace || workspace_bytes < cg_csr_build_workspace(edge_count, node_count, symmetrize))
        return fail(CG_EWORKSPACE, "cg_csr_build: workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    int32_t* cnt = cv.take<int32_t>(node_count + 1);
    int32_t* tcol = cv.take<int32_t>(nnz);
    int32_t* trow = cv.take<int32_t>(nnz);
    int* bad = cv.take<int>(1);
    const int n = static_cast<int>(node_count);
    hipLaunchKernelGGL(zero_ints, dim3(static_cast<unsigned>(div_up(node_count + 2, 256))), dim3(256), 0, s, cnt,
                       node_count + 1);
    if (int rc = check_launch("zero_ints")) return rc;
    hipLaunchKernelGGL(zero_ints, dim3(1), dim3(64), 0, s, reinterpret_cast<int32_t*>(bad), int64_t(1));
    if (int rc = check_launch("zero_ints")) return rc;
    const unsigned eg = static_cast<unsigned>(std::max<int64_t>(1, div_up(edge_count, 256)));
    if (edge_count > 0) {
        hipLaunchKernelGGL(count_degrees, dim3(eg), dim3(256), 0, s, edges, edge_count, n, symmetrize, cnt, bad);
        if (int rc = check_launch("count_degrees")) return rc;
        int hbad = 0;
        CG_CHECK_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
        CG_CHECK_HIP(hipStreamSynchronize(s));
        if (hbad) return fail(CG_EINVAL, "cg_csr_build: an edge endpoint is outside [0, %d)", n);
    }
    hipLaunchKernelGGL(scan_single_block, dim3(1), dim3(kScanThreads), 0, s, cnt, n, row_pointers);
    if (int rc = check_launch("scan_single_block")) return rc;
    if (edge_count == 0) return CG_OK;
    // cursor = row_pointers[0..n)
    CG_CHECK_HIP(hipMemcpyAsync(cnt, row_pointers, sizeof(int32_t) * node_count, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(scatter_edges, dim3(eg), dim3(256), 0, s, edges, edge_count, symmetrize, cnt, tcol, trow);
    if (int rc = check_launch("scatter_edges")) return rc;
    hipLaunchKernelGGL(rank_sort_rows, dim3(static_cast<unsigned>(div_up(nnz, 256))), dim3(256), 0, s, tcol, trow,
                       row_pointers, nnz, col_indices);
    return check_launch("rank_sort_rows");
}

int cg_graph_features(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, float* clustering,
                      float* degrees, int32_t* triangles, cg_stream_t stream) {
    if (!col_indices || !row_pointers || bad_n(n)) return fail(CG_EINVAL, "cg_graph_features: bad arguments");
    hipLaunchKernelGGL(vertex_triangles, dim3(static_cast<unsigned>(div_up(n, 4))), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), col_indices, row_pointers, static_cast<int>(n),
                       triangles, clustering, degrees);
    return check_launch("vertex_triangles");
}

int cg_average_neighbor_features(const int32_t* col_indices, const int32_t* row_pointers, int64_t n,
                                 const float* features, float* averages, cg_stream_t stream) {
    if (!col_indices || !row_pointers || !features || !averages || bad_n(n))
        return fail(CG_EINVAL, "cg_average_neighbor_features: bad arguments");
    hipLaunchKernelGGL(vertex_average, dim3(static_cast<unsigned>(div_up(n, 4))), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), col_indices, row_pointers, static_cast<int>(n), features,
                       averages);
    return check_launch("vertex_average");
}

size_t cg_node_features_workspace(int64_t n) { return n > 0 ? sizeof(int32_t) * static_cast<size_t>(n) + 256 : 0; }

int cg_node_features(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, float mu, float* out,
                     int32_t* triangles, void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    if (!col_indices || !row_pointers || !out || bad_n(n)) return fail(CG_EINVAL, "cg_node_features: bad arguments");
    if (!aligned16(out)) return fail(CG_EINVAL, "cg_node_features: out must be 16-byte aligned");
    int32_t* tri = triangles;
    if (!tri) {
        if (!workspace || workspace_bytes < cg_node_features_workspace(n))
            return fail(CG_EWORKSPACE, "cg_node_features: workspace too small");
        tri = static_cast<int32_t*>(workspace);
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const dim3 grid(static_cast<unsigned>(div_up(n, 4)));
    hipLaunchKernelGGL(vertex_triangles, grid, dim3(256), 0, s, col_indices, row_pointers, static_cast<int>(n), tri,
                       static_cast<float*>(nullptr), static_cast<float*>(nullptr));
    if (int rc = check_launch("vertex_triangles")) return rc;
    hipLaunchKernelGGL(vertex_features4, grid, dim3(256), 0, s, col_indices, row_pointers, tri, static_cast<int>(n),
                       mu, out);
    return check_launch("vertex_features4");
}

int cg_egonet_edges(const int32_t* row_pointers, const int32_t* triangles, int64_t n, int32_t* out,
                    cg_stream_t stream) {
    if (!row_pointers || !triangles || !out || bad_n(n)) return fail(CG_EINVAL, "cg_egonet_edges: bad arguments");
    hipLaunchKernelGGL(egonet_kernel, dim3(static_cast<unsigned>(div_up(n, 256))), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), row_pointers, triangles, static_cast<int>(n), out);
    return check_launch("egonet_kernel");
}

size_t cg_fugal_features_workspace(int64_t n) {
    if (n <= 0) return 0;
    Carver cv{nullptr, 0, ~size_t(0)};
    cv.take<int32_t>(static_cast<size_t>(n));
    cv.take<float>(4 * static_cast<size_t>(n));
    return cv.used + 256;
}

int cg_fugal_features(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, float mu, float* out,
                      void* workspace, size_t workspace_bytes, cg_stream_t stream) {
    if (!row_pointers || !out || bad_n(n)) return fail(CG_EINVAL, "cg_fugal_features: bad arguments");
    if (!aligned16(out)) return fail(CG_EINVAL, "cg_fugal_features: out must be 16-byte aligned");
    if (!workspace || workspace_bytes < cg_fugal_features_workspace(n))
        return fail(CG_EWORKSPACE, "cg_fugal_features: workspace too small");
    const size_t lds = sizeof(uint32_t) * 2 * static_cast<size_t>(div_up(n, 32));
    if (lds > 160 * 1024) return fail(CG_EUNSUPPORTED, "cg_fugal_features: n = %lld exceeds the LDS bitmaps",
                                      static_cast<long long>(n));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    int32_t* tri = cv.take<int32_t>(static_cast<size_t>(n));
    float* f4 = cv.take<float>(4 * static_cast<size_t>(n));
    const int ni = static_cast<int>(n);
    const dim3 grid(static_cast<unsigned>(div_up(n, 4)));
    hipLaunchKernelGGL(vertex_triangles, grid, dim3(256), 0, s, col_indices, row_pointers, ni, tri,
                       static_cast<float*>(nullptr), static_cast<float*>(nullptr));
    if (int rc = check_launch("vertex_triangles")) return rc;
    hipLaunchKernelGGL(vertex_features4, grid, dim3(256), 0, s, col_indices, row_pointers, tri, ni, mu, f4);
    if (int rc = check_launch("vertex_features4")) return rc;
    static size_t attr = 0;
    if (lds > attr) {
        CG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fugal_extra),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        attr = lds;
    }
    const unsigned blocks = static_cast<unsigned>(std::min<int64_t>(n, 2048));
    hipLaunchKernelGGL(fugal_extra, dim3(blocks), dim3(kFugalThreads), lds, s, col_indices, row_pointers, tri, f4, ni,
                       mu, out);
    return check_launch("fugal_extra");
}

int cg_edge_common_neighbours(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, int32_t* out,
                              cg_stream_t stream) {
    if (!row_pointers || !out || n <= 0 || n > (int64_t(1) << 30))
        return fail(CG_EINVAL, "cg_edge_common_neighbours: bad arguments");
    hipLaunchKernelGGL(entry_common_neighbours, dim3(static_cast<unsigned>(div_up(n, 4))), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), col_indices, row_pointers, static_cast<int>(n), out);
    return check_launch("entry_common_neighbours");
}

}  // extern "C"
