// Frank-Wolfe gradient of FUGAL for undirected graphs (A, B symmetric CSR):
//   G = -2 A P B + D + f * lambda * (1 - 2 P)
// cugal/pred.py:54-74 (sparse, f = 1) and :38-51 (dense, f = 2: the
// reference adds lambda (1 - 2P) twice there).  D[x][y] is the Euclidean
// distance of the 4-feature rows, recomputed in the epilogue instead of being
// read from an n x n matrix (cuda_kernels/distance.cu:11-33 computes the same
// quantity; cugal/feature_extraction.py:98-108).
//
// Both sparse products are row gathers (out row r = sum of the input rows
// listed by CSR row r), which read whole 128-byte row segments; the column
// side is turned into a row side by transposing through LDS:
//   1. Y = (A P)^T : 64 x 32 tiles of A P gathered, stored transposed;
//   2. G = epilogue(-2 (B Y)^T): tiles of B Y gathered, transposed in LDS
//      and written as rows of G with the distance and regulariser terms
//      (B symmetric, so (B Y)^T = (A P) B).
// Tiles are dealt so that the blocks one XCD runs together cover
// consecutive row tiles of one 32-column strip, which fits that XCD's L2.
// For undirected pairs both products run gather_win_t, which serves the
// neighbours near a tile (most of them, on locally ordered labels) from a
// window of rows staged in LDS.
#include <mutex>

#include "common.hpp"
#include "cugal_hip.h"

namespace cg {
namespace {

constexpr int kSpThreads = 256;
constexpr int kNumXcd = 8;

// ---- row gather out[i][c] = scale * sum_{k in N(i)} X[k][c] (+ epilogue) --
struct Epilogue {
    const float* P;   // nullptr: no regulariser term
    int64_t ldp;      // row stride of P
    const float* fs;  // source features n x fdim (nullptr: no distance term)
    const float* ft;  // target features n x fdim
    float rl;         // f * lambda
    int fdim;         // 4 or 8
    int pvec;         // P 16-byte aligned with a row stride divisible by 4
};

__device__ __forceinline__ float feat_dist(float4 a, const float* __restrict__ ft, int j) {
    const float4 b = *reinterpret_cast<const float4*>(ft + 4 * static_cast<int64_t>(j));
    const float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
    return sqrtf(fmaf(d0, d0, fmaf(d1, d1, fmaf(d2, d2, d3 * d3))));
}

template <int FD>
__device__ __forceinline__ float epi_one(float acc, float scale, const Epilogue& ep, int64_t idx, const FeatureRow& fsi,
                                         int j) {
    float g = acc * scale;
    if (ep.fs) g += feature_distance<FD>(fsi, ep.ft, j);
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
    FeatureRow fsi{};
    if (ep.fs) fsi = feature_row<4>(ep.fs, i);
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
        o.x = epi_one<4>(acc.x, scale, ep, idx + 0, fsi, j + 0);
        o.y = epi_one<4>(acc.y, scale, ep, idx + 1, fsi, j + 1);
        o.z = epi_one<4>(acc.z, scale, ep, idx + 2, fsi, j + 2);
        o.w = epi_one<4>(acc.w, scale, ep, idx + 3, fsi, j + 3);
        *reinterpret_cast<float4*>(out + idx) = o;
    } else {
        for (int j = c0 + threadIdx.x; j < min(n, c0 + chunk); j += kSpThreads) {
            float acc = 0.0f;
            for (int k = b; k < e; ++k) acc += X[static_cast<int64_t>(acol[k]) * n + j];
            const int64_t idx = i * n + j;
            out[idx] = epi_one<4>(acc, scale, ep, idx, fsi, j);
        }
    }
}

// ---- tiled gather with a transposed store ---------------------------------
// For the 64 x TC tile (rows r0.., columns c0..) of Z = M X (M given by CSR):
//   Z[r][c] = sum_{k in N(r)} X[k][c],
// write out[c][r] = Z[r][c] (EPI == false), or, with EPI, treat (c, r) as
// (i, j) of the gradient: out[c][r] = scale * Z[r][c] + D(c, r) + rl (1 - 2 P[c][r]).
// A 32-column strip of X is one 128-byte line per row, 4 MiB at n = 32768:
// the L2 of one XCD.  Tiles are ordered so that an XCD sweeps all row tiles
// of one strip before the next, so each line is fetched once and reused by
// every row that lists it; outputs are stored non-temporally so they do not
// evict the strip.
constexpr int kTileR = 64;

__device__ __forceinline__ void st_stream4(float* p, float4 v) {
    f32x4 w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<f32x4*>(p));
}

// Sum of the segments X[col[k]][cq..cq+3] for k in [k0, k1) of the staged
// index chunk (sidx[k - cs] = col[k]), in neighbour order; eight loads in
// flight per step.
__device__ __forceinline__ void gather_range(float4& acc, const int* __restrict__ sidx, int cs, int k0, int k1,
                                             const float* __restrict__ X, int64_t ldx, int cq) {
    int k = k0;
    for (; k + 8 <= k1; k += 8) {
        float4 x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = *reinterpret_cast<const float4*>(X + sidx[k - cs + q] * ldx + cq);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            acc.x += x[q].x;
            acc.y += x[q].y;
            acc.z += x[q].z;
            acc.w += x[q].w;
        }
    }
    for (; k + 4 <= k1; k += 4) {
        float4 x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = *reinterpret_cast<const float4*>(X + sidx[k - cs + q] * ldx + cq);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc.x += x[q].x;
            acc.y += x[q].y;
            acc.z += x[q].z;
            acc.w += x[q].w;
        }
    }
    for (; k < k1; ++k) {
        const float4 x = *reinterpret_cast<const float4*>(X + sidx[k - cs] * ldx + cq);
        acc.x += x.x;
        acc.y += x.y;
        acc.z += x.z;
        acc.w += x.w;
    }
}

// Column indices of a tile's rows are staged in LDS in chunks of this many.
constexpr int kIdxChunk = 2048;

// A second neighbour list gathered into the same sums (the directed
// gradient adds B Y1 and B^T Y2 tile by tile); rp == nullptr: none.
struct SecondList {
    const int32_t* rp;
    const int32_t* col;
    const float* X;
};

// Writes the gathered 64 x TC tile (tile[r][c] = Z[r0 + r][c0 + c])
// transposed: out[c][r] = Z[r][c], or with EPI the gradient terms
// scale * Z + D(c, r) + rl (1 - 2 P[c][r]).  Callers synchronise first.
template <bool EPI, bool VEC, int FD, int TC>
__device__ __forceinline__ void store_tile_t(const float (*tile)[TC + 1], int r0, int c0, int n, float scale,
                                             const Epilogue& ep, float* __restrict__ out, int64_t ldo) {
    const int g = threadIdx.x >> 4;  // 16 output rows per pass
    const int q = threadIdx.x & 15;  // four consecutive output columns
    const int rq = r0 + 4 * q;
    // A thread's four output columns are fixed: their target feature rows
    // are loaded once, not once per output row.
    FeatureRow ftj[4];
    if (EPI && ep.fs) {
#pragma unroll
        for (int e = 0; e < 4; ++e) ftj[e] = rq + e < n ? feature_row<FD>(ep.ft, rq + e) : FeatureRow{};
    }
    // Every load of a thread's rows is issued before the store that precedes
    // its use: loads and stores count on one counter, so a load issued after
    // a store would also wait for that store.  P (read from HBM) is loaded
    // for all rows at once, and so are 4-float source feature rows; an
    // 8-float row is loaded one row ahead, after the current row's
    // arithmetic (all of them at once would cost the kernel a wave per SIMD
    // in registers).
    constexpr int KR = TC / 16;
    constexpr int FH = FD == 4 ? KR : 1;  // source feature rows held at once
    float v[KR][4];
    float pv[KR][4];
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        const int cc = g + 16 * k;
        const int c = c0 + cc;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v[k][e] = tile[4 * q + e][cc];
            pv[k][e] = 0.0f;
        }
        if (EPI && ep.P && c < n) {
            // P read as one 16-byte load where aligned.
            const int64_t pbase = static_cast<int64_t>(c) * ep.ldp;
            if (VEC && ep.pvec && rq + 3 < n) {
                const float4 p4 = ld_stream4(ep.P + pbase + rq);
                pv[k][0] = p4.x;
                pv[k][1] = p4.y;
                pv[k][2] = p4.z;
                pv[k][3] = p4.w;
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (rq + e < n) pv[k][e] = ep.P[pbase + rq + e];
            }
        }
    }
    FeatureRow fsi[FH];
#pragma unroll
    for (int k = 0; k < FH; ++k) {
        const int c = c0 + g + 16 * k;
        fsi[k] = EPI && ep.fs && c < n ? feature_row<FD>(ep.fs, c) : FeatureRow{};
    }
#pragma unroll
    for (int k = 0; k < KR; ++k) {
        const int c = c0 + g + 16 * k;
        if (c >= n) break;
        const int64_t base = static_cast<int64_t>(c) * ldo;
        if (EPI) {
            // Same terms in the same order as epi_one: scale * sum, + D,
            // + rl (1 - 2 P).
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float gv = v[k][e] * scale;
                if (ep.fs) gv += feature_distance_rows<FD>(fsi[FH == KR ? k : 0], ftj[e]);
                if (ep.P) gv += ep.rl - 2.0f * ep.rl * pv[k][e];
                v[k][e] = gv;
            }
            if (FH == 1 && ep.fs && k + 1 < KR && c + 16 < n) fsi[0] = feature_row<FD>(ep.fs, c + 16);
        }
        if (VEC && rq + 3 < n) {
            st_stream4(out + base + rq, make_float4(v[k][0], v[k][1], v[k][2], v[k][3]));
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (rq + e < n) out[base + rq + e] = v[k][e];
        }
    }
}

template <bool EPI, bool VEC, int FD, int TC>
__global__ __launch_bounds__(kSpThreads) void gather_tile_t(const int32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ col, int n,
                                                           const float* __restrict__ X, int64_t ldx, float scale,
                                                           Epilogue ep, float* __restrict__ out, int64_t ldo,
                                                           int tiles_r, int64_t tiles, SecondList second) {
    constexpr int LPT = TC / 4;              // threads per TC-float segment
    constexpr int RG = kSpThreads / LPT;     // row groups
    constexpr int RPT = kTileR / RG;         // rows per thread
    __shared__ float tile[kTileR][TC + 1];
    __shared__ int sidx[VEC ? kIdxChunk : 1];
    __shared__ int srp[kTileR + 1];
    const int64_t t = xcd_tile(blockIdx.x, tiles);
    if (t >= tiles) return;
    const int r0 = static_cast<int>(t % tiles_r) * kTileR;
    const int c0 = static_cast<int>(t / tiles_r) * TC;
    if (VEC) {
        // The tile's index lists, staged through LDS with coalesced loads, so
        // each gather step issues its row loads without a dependent index
        // load first (measured 11 % faster per gradient at n = 32768).
        const int g = threadIdx.x / LPT;  // row group
        const int q = threadIdx.x % LPT;  // float4 slot within the segment
        const int cq = c0 + 4 * q;
        float4 acc[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) acc[r] = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int list = 0; list < (second.rp ? 2 : 1); ++list) {
            const int32_t* __restrict__ lrp = list ? second.rp : rp;
            const int32_t* __restrict__ lcol = list ? second.col : col;
            const float* __restrict__ lX = list ? second.X : X;
            if (threadIdx.x <= kTileR) srp[threadIdx.x] = lrp[min(r0 + static_cast<int>(threadIdx.x), n)];
            __syncthreads();
            const int e_lo = srp[0];
            const int e_hi = srp[kTileR];
            for (int cs = e_lo; cs < e_hi; cs += kIdxChunk) {
                const int ce = min(e_hi, cs + kIdxChunk);
                for (int k = cs + static_cast<int>(threadIdx.x); k < ce; k += kSpThreads) sidx[k - cs] = lcol[k];
                __syncthreads();
                if (cq < n) {
#pragma unroll
                    for (int r = 0; r < RPT; ++r) {
                        const int rr = g + r * RG;
                        gather_range(acc[r], sidx, cs, max(srp[rr], cs), min(srp[rr + 1], ce), lX, ldx, cq);
                    }
                }
                __syncthreads();
            }
        }
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            tile[g + r * RG][4 * q + 0] = acc[r].x;
            tile[g + r * RG][4 * q + 1] = acc[r].y;
            tile[g + r * RG][4 * q + 2] = acc[r].z;
            tile[g + r * RG][4 * q + 3] = acc[r].w;
        }
    } else {
        const int g = threadIdx.x / LPT;
        const int q = threadIdx.x % LPT;
        const int cq = c0 + 4 * q;
        for (int rr = g; rr < kTileR; rr += RG) {
            const int r = r0 + rr;
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r < n) {
                for (int list = 0; list < (second.rp ? 2 : 1); ++list) {
                    const int32_t* __restrict__ lrp = list ? second.rp : rp;
                    const int32_t* __restrict__ lcol = list ? second.col : col;
                    const float* __restrict__ lX = list ? second.X : X;
                    for (int k = lrp[r]; k < lrp[r + 1]; ++k) {
                        const float* xr = lX + static_cast<int64_t>(lcol[k]) * ldx;
                        if (cq + 0 < n) acc.x += xr[cq + 0];
                        if (cq + 1 < n) acc.y += xr[cq + 1];
                        if (cq + 2 < n) acc.z += xr[cq + 2];
                        if (cq + 3 < n) acc.w += xr[cq + 3];
                    }
                }
            }
            tile[rr][4 * q + 0] = acc.x;
            tile[rr][4 * q + 1] = acc.y;
            tile[rr][4 * q + 2] = acc.z;
            tile[rr][4 * q + 3] = acc.w;
        }
    }
    __syncthreads();
    store_tile_t<EPI, VEC, FD, TC>(tile, r0, c0, n, scale, ep, out, ldo);
}

// ---- tiled gather through an LDS window ----------------------------------
// Where the labels are local (reverse Cuthill-McKee order, a ring lattice's
// own order), most neighbours of a 64-row tile lie within kHalo rows of it.
// gather_win_t stages the tile's window of X, rows [r0 - kHalo, r0 + 64 +
// kHalo) x 32 columns (12 KiB), in LDS with buffer loads, and gathers those
// neighbours from LDS; only the others ("far") are loaded from global
// memory.  The texture addresser, which bounds the plain gather (one
// 16-byte lane load per neighbour and column quad: 79 % busy at n = 32768),
// then handles the window's 96 row segments plus the far ones instead of
// every neighbour's.  The small window keeps 7 blocks per CU, which hides
// the tile's dependent load phases (a 192 x 64 window, 3 blocks per CU, was
// twice as slow as the plain gather).  A row's local entries are one run
// [lb, le) of its sorted column list, found once per call by window_split.
constexpr int kHalo = 16;
constexpr int kWinR = kTileR + 2 * kHalo;  // 96 window rows
constexpr int kWinTC = 32;                 // columns per tile: 8 lanes x 16 bytes per row
constexpr int kWinIdxChunk = 2048;         // staged column indices: a whole tile's lists at degree 32
constexpr int kFarWindow = 4;              // far loads in flight per thread
constexpr int kLongRow = 64;               // a graph with a longer list (a hub) takes gather_tile_t
constexpr int kRsrcWord3 = 0x00020000;     // raw buffer descriptor word 3 on CDNA

// split[r] = (lb, le): the entries [lb, le) of row r whose columns lie in the
// window of its row tile, [64 (r / 64) - kHalo, 64 (r / 64) + 64 + kHalo).
// Binary searches assume a sorted row; the run is then checked, and a row
// whose run holds a column outside the window (an unsorted row) gets an
// empty run, so all its entries are gathered from global memory.
__global__ __launch_bounds__(kSpThreads) void window_split(const int32_t* __restrict__ rp,
                                                          const int32_t* __restrict__ col, int n, int halo,
                                                          int2* __restrict__ split) {
    const int r = static_cast<int>(blockIdx.x) * kSpThreads + static_cast<int>(threadIdx.x);
    if (r >= n) return;
    const int wlo = (r / kTileR) * kTileR - halo;
    const int whi = wlo + kTileR + 2 * halo;
    const int b = rp[r], e = rp[r + 1];
    int lo = b, hi = e;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < wlo) lo = mid + 1; else hi = mid;
    }
    const int lb = lo;
    hi = e;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < whi) lo = mid + 1; else hi = mid;
    }
    int le = lo;
    for (int k = lb; k < le; ++k) {
        const int c = col[k];
        if (c < wlo || c >= whi) {
            le = lb;
            break;
        }
    }
    split[r] = make_int2(lb, le);
}

// *hub = 1 when some row lists more than kLongRow entries (one block).  The
// window kernel walks a thread's lists kFarWindow entries per round, which
// suits short lists; a hub's long list would hold its tile for hundreds of
// rounds (Barabasi-Albert n = 8192: 1.25 ms per gradient against 0.78 ms
// with gather_tile_t, whose rows keep eight loads in flight), so a graph
// with a hub is gathered by gather_tile_t.  The choice is made per graph
// on the host: both a hub branch inside the window kernel and a
// tile-walking form of both kernels, either of which would avoid the host
// read, measured slower.
constexpr int kHubThreads = 1024;
__global__ __launch_bounds__(kHubThreads) void graph_has_hub(const int32_t* __restrict__ rp, int n,
                                                            int32_t* __restrict__ hub) {
    int found = 0;
    for (int r = static_cast<int>(threadIdx.x); r < n; r += kHubThreads) found |= (rp[r + 1] - rp[r]) > kLongRow;
    found = __syncthreads_or(found);
    if (threadIdx.x == 0) hub[0] = found ? 1 : 0;
}

template <bool EPI, int FD>
__global__ __launch_bounds__(kSpThreads) void gather_win_t(const int32_t* __restrict__ rp,
                                                          const int32_t* __restrict__ col,
                                                          const int2* __restrict__ split, int n,
                                                          const float* __restrict__ X, int64_t ldx, float scale,
                                                          Epilogue ep, float* __restrict__ out, int64_t ldo,
                                                          int tiles_r, int64_t tiles) {
    constexpr int TC = kWinTC;
    constexpr int WINR = kWinR;
    constexpr int HALO = kHalo;
    constexpr int CHUNK = kWinIdxChunk;
    constexpr int FW = kFarWindow;
    constexpr int LPT = TC / 4;           // 8 lanes per row segment
    constexpr int RG = kSpThreads / LPT;  // 32 row groups
    constexpr int RPT = kTileR / RG;      // 2 rows per thread
    // The window, later reused for the transposed tile (64 x 33 floats).
    __shared__ __attribute__((aligned(16))) float win[(WINR * TC > kTileR * (TC + 1)) ? WINR * TC : kTileR * (TC + 1)];
    __shared__ int sidx[CHUNK];
    __shared__ int srp[kTileR + 1];
    __shared__ int slb[kTileR], sle[kTileR];
    __shared__ int s_local;
    const int64_t t = xcd_tile(blockIdx.x, tiles);
    if (t >= tiles) return;
    const int r0 = static_cast<int>(t % tiles_r) * kTileR;
    const int c0 = static_cast<int>(t / tiles_r) * TC;
    const int wbase = r0 - HALO;
    const int tid = static_cast<int>(threadIdx.x);
    if (tid <= kTileR) srp[tid] = rp[min(r0 + tid, n)];
    if (tid < kTileR) {
        const int r = r0 + tid;
        const int2 sp = r < n ? split[r] : make_int2(rp[n], rp[n]);
        slb[tid] = sp.x;
        sle[tid] = sp.y;
        const int local = wave_sum(sp.y - sp.x);  // wave 0 holds the tile's rows
        if (tid == 0) s_local = local;
    }
    __syncthreads();
    // Staging costs the window's kWinR row segments; it pays when the tile
    // has more local entries than that.
    const bool use_win = s_local > WINR;
    if (use_win) {
        // Window row i is X row clamp(wbase + i, 0, n - 1); rows outside
        // [0, n) are never listed, so their contents do not matter.  Load ls
        // of wave w fills window rows RPW (w + 4 ls) .. + RPW - 1 (1 KiB);
        // the descriptor's range ends at the last window row's column n, so
        // column quads past n read zeros.
        const int base_row = max(wbase, 0);
        const int last_row = min(wbase + WINR - 1, n - 1);
        const uint32_t nrec = static_cast<uint32_t>(((last_row - base_row) * ldx + n - c0) * 4);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float*>(X + static_cast<int64_t>(base_row) * ldx + c0), 0, nrec, kRsrcWord3);
        const int wave = tid >> 6, lane = tid & 63;
        constexpr int RPW = 1024 / (TC * 4);  // window rows per 1 KiB wave load
        static_assert(WINR % (4 * RPW) == 0, "the window must split into whole wave loads");
#pragma unroll
        for (int ls = 0; ls < WINR / (4 * RPW); ++ls) {
            const int wr = RPW * (wave + 4 * ls) + lane / LPT;
            const int grow = min(max(wbase + wr, 0), n - 1);
            const uint32_t voff = static_cast<uint32_t>((static_cast<int64_t>(grow - base_row) * ldx + 4 * (lane % LPT)) * 4);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rs, (__attribute__((address_space(3))) void*)(reinterpret_cast<char*>(win) + 1024 * (wave + 4 * ls)),
                16, voff, 0, 0, 0);
        }
    }
    const int g = tid / LPT;  // row group
    const int q = tid % LPT;  // float4 slot within the segment
    const int cq = c0 + 4 * q;
    float4 acc[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) acc[r] = make_float4(0.f, 0.f, 0.f, 0.f);
    const int e_lo = srp[0];
    const int e_hi = srp[kTileR];
    bool window_ready = !use_win;
    for (int cs = e_lo; cs < e_hi; cs += CHUNK) {
        const int ce = min(e_hi, cs + CHUNK);
        for (int k = cs + tid; k < ce; k += kSpThreads) sidx[k - cs] = col[k];
        __syncthreads();
        // Far entries: [row start, lb) and [le, row end) of each of the
        // thread's rows (the whole row when the window is not used), walked
        // as one list (row 0's runs, then row 1's, ...) FW entries per round,
        // all loaded before any is added; each row's entries are added in
        // list order.  A hub's long list thus keeps FW loads in flight while
        // short lists finish in one round.
        {
            constexpr int NSEG = 2 * RPT;
            int seg_k[NSEG], seg_end[NSEG];  // list position -> entry index per segment
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int rr = g + r * RG;
                const int b0 = srp[rr], e1 = srp[rr + 1];
                const int e0 = use_win ? slb[rr] : e1;
                const int b1 = use_win ? sle[rr] : e1;
                seg_k[2 * r] = max(b0, cs);
                seg_end[2 * r] = max(min(e0, ce), seg_k[2 * r]);
                seg_k[2 * r + 1] = max(b1, cs);
                seg_end[2 * r + 1] = max(min(e1, ce), seg_k[2 * r + 1]);
            }
            int start[NSEG + 1];
            start[0] = 0;
#pragma unroll
            for (int q2 = 0; q2 < NSEG; ++q2) start[q2 + 1] = start[q2] + (seg_end[q2] - seg_k[q2]);
            const int total = start[NSEG];
            for (int off = 0; off < total; off += FW) {
                float4 x[FW];
                int rowof[FW];
#pragma unroll
                for (int u = 0; u < FW; ++u) {
                    const int e = off + u;
                    x[u] = make_float4(0.f, 0.f, 0.f, 0.f);
                    rowof[u] = -1;
                    int sg = 0;
#pragma unroll
                    for (int q2 = 1; q2 < NSEG; ++q2) sg += e >= start[q2] ? 1 : 0;
                    if (e < total) {
                        int k = 0;
#pragma unroll
                        for (int q2 = 0; q2 < NSEG; ++q2)
                            if (q2 == sg) k = seg_k[q2] + (e - start[q2]);
                        rowof[u] = sg >> 1;
                        if (cq < n) x[u] = *reinterpret_cast<const float4*>(X + sidx[k - cs] * ldx + cq);
                    }
                }
#pragma unroll
                for (int u = 0; u < FW; ++u) {
#pragma unroll
                    for (int r = 0; r < RPT; ++r)
                        if (rowof[u] == r) {
                            acc[r].x += x[u].x;
                            acc[r].y += x[u].y;
                            acc[r].z += x[u].z;
                            acc[r].w += x[u].w;
                        }
                }
            }
        }
        if (!window_ready) {
            // The window's buffer loads have landed in every wave.
            __builtin_amdgcn_s_waitcnt(0x0070);
            __syncthreads();
            window_ready = true;
        }
        if (use_win) {
            const float4* __restrict__ w4 = reinterpret_cast<const float4*>(win);
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int rr = g + r * RG;
                const int k0 = max(slb[rr], cs), k1 = min(sle[rr], ce);
                int k = k0;
                for (; k + 8 <= k1; k += 8) {
                    float4 x[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) x[u] = w4[(sidx[k - cs + u] - wbase) * LPT + q];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        acc[r].x += x[u].x;
                        acc[r].y += x[u].y;
                        acc[r].z += x[u].z;
                        acc[r].w += x[u].w;
                    }
                }
                for (; k < k1; ++k) {
                    const float4 x = w4[(sidx[k - cs] - wbase) * LPT + q];
                    acc[r].x += x.x;
                    acc[r].y += x.y;
                    acc[r].z += x.z;
                    acc[r].w += x.w;
                }
            }
        }
        __syncthreads();
    }
    if (!window_ready) {
        // A tile without entries still drains its window loads before the
        // window's memory is reused.
        __builtin_amdgcn_s_waitcnt(0x0070);
        __syncthreads();
    }
    float (*tile)[TC + 1] = reinterpret_cast<float (*)[TC + 1]>(win);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        tile[g + r * RG][4 * q + 0] = acc[r].x;
        tile[g + r * RG][4 * q + 1] = acc[r].y;
        tile[g + r * RG][4 * q + 2] = acc[r].z;
        tile[g + r * RG][4 * q + 3] = acc[r].w;
    }
    __syncthreads();
    store_tile_t<EPI, true, FD, TC>(tile, r0, c0, n, scale, ep, out, ldo);
}

// out (row stride ldo) = X (row stride ldx): row-major streaming copy.
__global__ __launch_bounds__(kSpThreads) void copy_rows_padded(const float* __restrict__ X, int64_t ldx, int n,
                                                              float* __restrict__ out, int64_t ldo) {
    const int64_t r = blockIdx.x;
    const float* src = X + r * ldx;
    float* dst = out + r * ldo;
    for (int j = threadIdx.x * 4; j < n; j += kSpThreads * 4) st_stream4(dst + j, ld_stream4(src + j));
}

// dst[i][j] = src[row_map[i]][col_map[j]] (a null map is the identity): one
// block per output row; the source row is read once through L2 by the
// gathered column loads, the output row is written with 16-byte stores.
template <bool VEC>
__global__ __launch_bounds__(kSpThreads) void permute_rows_cols(const float* __restrict__ src, int64_t lds, int n,
                                                               const int32_t* __restrict__ row_map,
                                                               const int32_t* __restrict__ col_map,
                                                               float* __restrict__ dst, int64_t ldd) {
    const int i = blockIdx.x;
    const int si = row_map ? row_map[i] : i;
    const float* __restrict__ srow = src + static_cast<int64_t>(si) * lds;
    float* __restrict__ drow = dst + static_cast<int64_t>(i) * ldd;
    if (VEC) {
        for (int j = threadIdx.x * 4; j < n; j += kSpThreads * 4) {
            float4 o;
            if (col_map) {
                const int4 c = *reinterpret_cast<const int4*>(col_map + j);
                o = make_float4(srow[c.x], srow[c.y], srow[c.z], srow[c.w]);
            } else {
                o = *reinterpret_cast<const float4*>(srow + j);
            }
            st_stream4(drow + j, o);
        }
    } else {
        for (int j = threadIdx.x; j < n; j += kSpThreads) drow[j] = srow[col_map ? col_map[j] : j];
    }
}

// The same copy with the source row staged in LDS (n floats, n <= the LDS a
// block may hold): the row is read with coalesced 16-byte loads and the
// column gather runs on LDS instead of issuing a cache-line request per
// element.  Blocks stride over the rows.
constexpr int kPermThreads = 1024;

__global__ __launch_bounds__(kPermThreads) void permute_rows_cols_lds(const float* __restrict__ src, int64_t lds,
                                                                     int n, const int32_t* __restrict__ row_map,
                                                                     const int32_t* __restrict__ col_map,
                                                                     float* __restrict__ dst, int64_t ldd) {
    extern __shared__ float srow_lds[];
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int si = row_map ? row_map[i] : i;
        const float* __restrict__ srow = src + static_cast<int64_t>(si) * lds;
        for (int j = threadIdx.x * 4; j < n; j += kPermThreads * 4)
            *reinterpret_cast<float4*>(srow_lds + j) = ld_stream4(srow + j);
        __syncthreads();
        float* __restrict__ drow = dst + static_cast<int64_t>(i) * ldd;
        for (int j = threadIdx.x * 4; j < n; j += kPermThreads * 4) {
            const int4 c = *reinterpret_cast<const int4*>(col_map + j);
            st_stream4(drow + j, make_float4(srow_lds[c.x], srow_lds[c.y], srow_lds[c.z], srow_lds[c.w]));
        }
        __syncthreads();
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

template <bool EPI, int FD, int TC>
int launch_gather_tile_fd(const int32_t* rp, const int32_t* col, int n, const float* X, int64_t ldx, float scale,
                       const Epilogue& ep, float* out, int64_t ldo, hipStream_t s, SecondList second) {
    const bool vec = (n % 4 == 0) && (ldx % 4 == 0) && (ldo % 4 == 0) && aligned16(X) && aligned16(out) &&
                     (!ep.P || aligned16(ep.P)) && (!second.rp || aligned16(second.X)) &&
                     (!ep.fs || (aligned16(ep.fs) && aligned16(ep.ft)));
    const int tiles_r = static_cast<int>(div_up(n, kTileR));
    const int64_t tiles = static_cast<int64_t>(tiles_r) * div_up(n, TC);
    const int64_t blocks = div_up(tiles, kNumXcd) * kNumXcd;
    if (blocks > 0x7fffffffLL) return fail(CG_EUNSUPPORTED, "gradient grid too large (n=%d)", n);
    if (vec)
        hipLaunchKernelGGL((gather_tile_t<EPI, true, FD, TC>), dim3(static_cast<unsigned>(blocks)), dim3(kSpThreads), 0, s, rp,
                           col, n, X, ldx, scale, ep, out, ldo, tiles_r, tiles, second);
    else
        hipLaunchKernelGGL((gather_tile_t<EPI, false, FD, TC>), dim3(static_cast<unsigned>(blocks)), dim3(kSpThreads), 0, s,
                           rp, col, n, X, ldx, scale, ep, out, ldo, tiles_r, tiles, second);
    return check_launch("gather_tile_t");
}

template <bool EPI>
int launch_gather_tile(const int32_t* rp, const int32_t* col, int n, const float* X, int64_t ldx, float scale,
                       const Epilogue& ep, float* out, int64_t ldo, hipStream_t s,
                       SecondList second = SecondList{nullptr, nullptr, nullptr}) {
    // The first product (no epilogue) gathers 64-column strips: 256-byte
    // segments, 4 % faster there; the second gathers 32-column strips, whose
    // 4 MiB footprint at n = 32768 already fills an XCD's L2 (64 columns
    // measured 4 % slower).  Same sums in the same order either way.
    if (!EPI) return launch_gather_tile_fd<EPI, 4, 64>(rp, col, n, X, ldx, scale, ep, out, ldo, s, second);
    if (ep.fs && ep.fdim == 8)
        return launch_gather_tile_fd<EPI, 8, 32>(rp, col, n, X, ldx, scale, ep, out, ldo, s, second);
    return launch_gather_tile_fd<EPI, 4, 32>(rp, col, n, X, ldx, scale, ep, out, ldo, s, second);
}

// The window gather where its 16-byte accesses are aligned (else returns 1
// without launching, and the caller takes the plain gather).
// The window gather where its 16-byte accesses are aligned (else returns 0
// without launching, and the caller takes the plain gather).
template <bool EPI>
int launch_gather_win(const int32_t* rp, const int32_t* col, const int2* split, int n, const float* X, int64_t ldx,
                      float scale, const Epilogue& ep, float* out, int64_t ldo, hipStream_t s, bool& launched) {
    launched = false;
    const bool vec = (n % 4 == 0) && (ldx % 4 == 0) && (ldo % 4 == 0) && aligned16(X) && aligned16(out) &&
                     (!ep.P || aligned16(ep.P)) && (!ep.fs || (aligned16(ep.fs) && aligned16(ep.ft)));
    if (!vec) return 0;
    const int tiles_r = static_cast<int>(div_up(n, kTileR));
    const int64_t tiles = static_cast<int64_t>(tiles_r) * div_up(n, kWinTC);
    const int64_t blocks = div_up(tiles, kNumXcd) * kNumXcd;
    if (blocks > 0x7fffffffLL) return fail(CG_EUNSUPPORTED, "gradient grid too large (n=%d)", n);
    launched = true;
    if constexpr (EPI) {
        if (ep.fs && ep.fdim == 8) {
            hipLaunchKernelGGL((gather_win_t<EPI, 8>), dim3(static_cast<unsigned>(blocks)), dim3(kSpThreads), 0, s, rp,
                               col, split, n, X, ldx, scale, ep, out, ldo, tiles_r, tiles);
            return check_launch("gather_win_t");
        }
    }
    hipLaunchKernelGGL((gather_win_t<EPI, 4>), dim3(static_cast<unsigned>(blocks)), dim3(kSpThreads), 0, s, rp, col,
                       split, n, X, ldx, scale, ep, out, ldo, tiles_r, tiles);
    return check_launch("gather_win_t");
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

// Whether the CSR (rp, col, n) has a hub, measured once per graph (one small
// kernel and a stream synchronisation on the first call) and remembered by
// its device pointers.  Either gather is correct for any graph, so an entry
// left by a freed graph whose memory a new one reuses can cost speed, never
// a result.
int graph_hub_cached(const int32_t* rp, const int32_t* col, int n, int32_t* dev_flag, hipStream_t s, bool& hub) {
    struct Entry {
        const int32_t* rp;
        const int32_t* col;
        int n;
        bool hub;
    };
    static std::mutex mu;
    static Entry cache[16];
    static int next = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        for (const Entry& e : cache)
            if (e.rp == rp && e.col == col && e.n == n && rp) {
                hub = e.hub;
                return 0;
            }
    }
    hipLaunchKernelGGL(graph_has_hub, dim3(1), dim3(kHubThreads), 0, s, rp, n, dev_flag);
    if (int rc = check_launch("graph_has_hub")) return rc;
    int32_t h = 0;
    CG_CHECK_HIP(hipMemcpyAsync(&h, dev_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    CG_CHECK_HIP(hipStreamSynchronize(s));
    hub = h != 0;
    std::lock_guard<std::mutex> lock(mu);
    cache[next] = Entry{rp, col, n, hub};
    next = (next + 1) % 16;
    return 0;
}

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

int cg_adjacency_matmul(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, const float* matrix,
                        float* out, int negate, cg_stream_t stream) {
    if (!col_indices || !row_pointers || !matrix || !out || bad_n(n))
        return fail(CG_EINVAL, "cg_adjacency_matmul: bad arguments");
    Epilogue ep{nullptr, 0, nullptr, nullptr, 0.0f, 4, 0};
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

int cg_permute_copy(const float* src, int64_t ld_src, int64_t n, const int32_t* row_map, const int32_t* col_map,
                    float* dst, int64_t ld_dst, cg_stream_t stream) {
    if (!src || !dst || bad_n(n) || ld_src < n || ld_dst < n) return fail(CG_EINVAL, "cg_permute_copy: bad arguments");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const bool vec = n % 4 == 0 && ld_dst % 4 == 0 && aligned16(dst) && (!col_map || aligned16(col_map)) &&
                     (col_map || (ld_src % 4 == 0 && aligned16(src)));
    const size_t row_bytes = sizeof(float) * static_cast<size_t>(n);
    if (vec && col_map && ld_src % 4 == 0 && aligned16(src) && row_bytes + 1024 <= static_cast<size_t>(max_lds_bytes())) {
        CG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&permute_rows_cols_lds),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(row_bytes)));
        int dev = 0, cus = 0;
        CG_CHECK_HIP(hipGetDevice(&dev));
        CG_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const int64_t per_cu = std::max<int64_t>(1, static_cast<int64_t>(max_lds_bytes()) / (row_bytes + 1024));
        const unsigned blocks = static_cast<unsigned>(std::min<int64_t>(n, std::max(1, cus) * std::min<int64_t>(per_cu, 2)));
        hipLaunchKernelGGL(permute_rows_cols_lds, dim3(blocks), dim3(kPermThreads), row_bytes, s, src, ld_src,
                           static_cast<int>(n), row_map, col_map, dst, ld_dst);
        return check_launch("permute_rows_cols_lds");
    }
    if (vec)
        hipLaunchKernelGGL(permute_rows_cols<true>, dim3(static_cast<unsigned>(n)), dim3(kSpThreads), 0, s, src,
                           ld_src, static_cast<int>(n), row_map, col_map, dst, ld_dst);
    else
        hipLaunchKernelGGL(permute_rows_cols<false>, dim3(static_cast<unsigned>(n)), dim3(kSpThreads), 0, s, src,
                           ld_src, static_cast<int>(n), row_map, col_map, dst, ld_dst);
    return check_launch("permute_rows_cols");
}

int cg_regularize(float* gradient, const float* P, int64_t n, int iteration, cg_stream_t stream) {
    if (!gradient || !P || bad_n(n)) return fail(CG_EINVAL, "cg_regularize: bad arguments");
    const int64_t total = n * n;
    hipLaunchKernelGGL(regularize_kernel, dim3(static_cast<unsigned>(div_up(total, 256))), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), gradient, P, total, static_cast<float>(iteration));
    return check_launch("regularize_kernel");
}

// Row stride of the intermediate: a power-of-two stride would map the lines
// of one column strip onto the same few L2 sets.
static int64_t gradient_ld(int64_t n) {
    const int64_t ld = round_up(n, 4) + 64;
    return (ld * 4) % 4096 == 0 ? ld + 32 : ld;
}

// A row stride whose 128-byte column segments spread over the L2 sets.
static bool gather_friendly_ld(int64_t ld) { return ld % 4 == 0 && (ld * 4) % 4096 != 0; }

// Two padded n x n buffers, a copy of P (gathered by the first product) and
// the intermediate Y, the window runs of both adjacencies and a flag.
size_t cg_gradient_workspace(int64_t n) {
    return n > 0 ? 2 * (sizeof(float) * static_cast<size_t>(n) * static_cast<size_t>(gradient_ld(n)) + 256) +
                       2 * (sizeof(int2) * static_cast<size_t>(n) + 256) + 256 + sizeof(int32_t)
                 : 0;
}

int64_t cg_gradient_padded_ld(int64_t n) { return n > 0 ? gradient_ld(n) : 0; }

namespace {

// -2 A P B (symmetric A, B) or -(A P B^T + A^T P B) (directed: the
// transposes given), + D + factor * lambda (1 - 2P).
int gradient_impl(const int32_t* a_col, const int32_t* a_rowptr, const int32_t* at_col, const int32_t* at_rowptr,
                  const int32_t* b_col, const int32_t* b_rowptr, const int32_t* bt_col, const int32_t* bt_rowptr,
                  int64_t n, const float* P, int64_t ld_p, const float* source_features,
                  const float* target_features, int feature_dim, int lambda_, float regularizer_factor,
                  float* gradient, void* workspace, size_t workspace_bytes, hipStream_t s, const char* who) {
    const bool directed = at_col != nullptr;
    if (!a_col || !a_rowptr || !b_col || !b_rowptr || !P || !gradient || bad_n(n) || ld_p < n ||
        (directed && (!at_rowptr || !bt_col || !bt_rowptr)))
        return fail(CG_EINVAL, "%s: bad arguments", who);
    if ((source_features == nullptr) != (target_features == nullptr))
        return fail(CG_EINVAL, "%s: give both feature matrices or neither", who);
    if (source_features && (!aligned16(source_features) || !aligned16(target_features)))
        return fail(CG_EINVAL, "%s: feature matrices must be 16-byte aligned", who);
    if (feature_dim != 4 && feature_dim != 8) return fail(CG_EINVAL, "%s: feature_dim must be 4 or 8", who);
    const size_t need = directed ? cg_gradient_directed_workspace(n) : cg_gradient_workspace(n);
    if (!workspace || workspace_bytes < need) return fail(CG_EWORKSPACE, "%s: workspace too small", who);
    const int ni = static_cast<int>(n);
    const int64_t ldy = gradient_ld(n);
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    float* Pp = cv.take<float>(static_cast<size_t>(n) * ldy);
    float* Y = cv.take<float>(static_cast<size_t>(n) * ldy);
    float* Y2 = directed ? cv.take<float>(static_cast<size_t>(n) * ldy) : nullptr;
    const Epilogue none{nullptr, 0, nullptr, nullptr, 0.0f, 4, 0};
    const float* X1 = P;
    int64_t ld1 = ld_p;
    if (n % 4 == 0 && aligned16(P) && !gather_friendly_ld(ld_p)) {
        hipLaunchKernelGGL(copy_rows_padded, dim3(static_cast<unsigned>(n)), dim3(kSpThreads), 0, s, P, ld_p, ni,
                           Pp, ldy);
        if (int rc = check_launch("copy_rows_padded")) return rc;
        X1 = Pp;
        ld1 = ldy;
    }
    if (!directed) {
        // Symmetric pairs: each product through the LDS window, unless its
        // adjacency has a hub (then the plain gather).
        int2* a_split = cv.take<int2>(static_cast<size_t>(n));
        int2* b_split = cv.take<int2>(static_cast<size_t>(n));
        int32_t* flag = cv.take<int32_t>(1);
        bool a_hub = false, b_hub = false;
        if (int rc = graph_hub_cached(a_rowptr, a_col, ni, flag, s, a_hub)) return rc;
        if (int rc = graph_hub_cached(b_rowptr, b_col, ni, flag, s, b_hub)) return rc;
        const unsigned sb = static_cast<unsigned>(div_up(n, kSpThreads));
        bool launched = false;
        if (!a_hub) {
            hipLaunchKernelGGL(window_split, dim3(sb), dim3(kSpThreads), 0, s, a_rowptr, a_col, ni, kHalo, a_split);
            if (int rc = check_launch("window_split")) return rc;
            if (int rc = launch_gather_win<false>(a_rowptr, a_col, a_split, ni, X1, ld1, 1.0f, none, Y, ldy, s, launched))
                return rc;
        }
        if (!launched)
            if (int rc = launch_gather_tile<false>(a_rowptr, a_col, ni, X1, ld1, 1.0f, none, Y, ldy, s)) return rc;
        const float rl = regularizer_factor * static_cast<float>(lambda_);
        Epilogue ep{rl != 0.0f ? P : nullptr, ld_p, source_features, target_features, rl, feature_dim,
                    aligned16(P) && ld_p % 4 == 0 ? 1 : 0};
        launched = false;
        if (!b_hub) {
            hipLaunchKernelGGL(window_split, dim3(sb), dim3(kSpThreads), 0, s, b_rowptr, b_col, ni, kHalo, b_split);
            if (int rc = check_launch("window_split")) return rc;
            if (int rc = launch_gather_win<true>(b_rowptr, b_col, b_split, ni, Y, ldy, -2.0f, ep, gradient, n, s,
                                                 launched))
                return rc;
        }
        if (launched) return 0;
        return launch_gather_tile<true>(b_rowptr, b_col, ni, Y, ldy, -2.0f, ep, gradient, n, s);
    }
    // Directed: Y = (A P)^T, Y2 = (A^T P)^T
    if (int rc = launch_gather_tile<false>(a_rowptr, a_col, ni, X1, ld1, 1.0f, none, Y, ldy, s)) return rc;
    if (int rc = launch_gather_tile<false>(at_rowptr, at_col, ni, X1, ld1, 1.0f, none, Y2, ldy, s)) return rc;
    const float rl = regularizer_factor * static_cast<float>(lambda_);
    Epilogue ep{rl != 0.0f ? P : nullptr, ld_p, source_features, target_features, rl, feature_dim,
                aligned16(P) && ld_p % 4 == 0 ? 1 : 0};
    // (B Y + B^T Y2)^T = A P B^T + A^T P B, summed in one tile pass
    return launch_gather_tile<true>(b_rowptr, b_col, ni, Y, ldy, -1.0f, ep, gradient, n, s,
                                    SecondList{bt_rowptr, bt_col, Y2});
}

}  // namespace

size_t cg_gradient_directed_workspace(int64_t n) {
    return n > 0 ? 3 * (sizeof(float) * static_cast<size_t>(n) * static_cast<size_t>(gradient_ld(n)) + 256) : 0;
}

int cg_gradient(const int32_t* a_col, const int32_t* a_rowptr, const int32_t* b_col, const int32_t* b_rowptr,
                int64_t n, const float* P, int64_t ld_p, const float* source_features, const float* target_features,
                int feature_dim, int lambda_, float regularizer_factor, float* gradient, void* workspace,
                size_t workspace_bytes, cg_stream_t stream) {
    return gradient_impl(a_col, a_rowptr, nullptr, nullptr, b_col, b_rowptr, nullptr, nullptr, n, P, ld_p,
                         source_features, target_features, feature_dim, lambda_, regularizer_factor, gradient,
                         workspace, workspace_bytes, reinterpret_cast<hipStream_t>(stream), "cg_gradient");
}

int cg_gradient_directed(const int32_t* a_col, const int32_t* a_rowptr, const int32_t* at_col,
                         const int32_t* at_rowptr, const int32_t* b_col, const int32_t* b_rowptr,
                         const int32_t* bt_col, const int32_t* bt_rowptr, int64_t n, const float* P, int64_t ld_p,
                         const float* source_features, const float* target_features, int feature_dim, int lambda_,
                         float regularizer_factor, float* gradient, void* workspace, size_t workspace_bytes,
                         cg_stream_t stream) {
    if (!at_col) return fail(CG_EINVAL, "cg_gradient_directed: the transposes are required");
    return gradient_impl(a_col, a_rowptr, at_col, at_rowptr, b_col, b_rowptr, bt_col, bt_rowptr, n, P, ld_p,
                         source_features, target_features, feature_dim, lambda_, regularizer_factor, gradient,
                         workspace, workspace_bytes, reinterpret_cast<hipStream_t>(stream), "cg_gradient_directed");
}

}  // extern "C"
