// Frank-Wolfe gradient on the matrix cores (the precision form):
//   G = -2 A P B + D + f * lambda * (1 - 2 P)
// with A, B the 0/1 adjacencies of undirected graphs as exact BF16 operands
// and P split into BF16 hi + lo (hi = bf16(P), lo = bf16(P - hi)), every
// product accumulated in FP32.  The adjacencies are stored as bits (one
// 32-bit word per row and 32-wide k slice, slice-major: the words of 256
// consecutive rows for one slice are 1 KiB contiguous) and expanded to
// BF16 fragments in registers after a 4-byte LDS read.  Same quantity as cugal/pred.py:38-51
// (dense, f = 2) and :54-74 (sparse, f = 1); the reference runs it as torch
// matmuls / cuSPARSE SpMM in FP32.
//
// Three launches over matrices padded with zeros to np = n rounded up to
// 256 (so every tile is whole):
//   1. split_transpose: P^T -> PT_hi, PT_lo (BF16, np x np), through LDS;
//   2. gemm_nt<split R>: Y = A P, i.e. Y[m][c] = sum_k A[m][k] PT[c][k]
//      (hi and lo products into the same accumulators), written back as
//      Y_hi, Y_lo;
//   3. gemm_nt<split L>: G[m][c] = -2 sum_k Y[m][k] B[c][k] (= -2 Y B, B
//      symmetric) with the distance (from the 4-feature rows) and the
//      regulariser fused in the epilogue.
// Both GEMMs are "NT": each operand row holds consecutive k, so an MFMA
// fragment is one 16-byte read (for the bit operand, one byte of a word).
// A block computes a 256 x 256 tile with four waves of 128 x 128
// (v_mfma_f32_32x32x16_bf16, 4 x 4 tiles of 32 x 32), staging 32-wide k
// slices of the split operand's hi and lo tiles (16 KiB each) and the bit
// operand's words (1 KiB) in three LDS buffers filled by direct global ->
// LDS loads (global_load_lds_dwordx4 / _dword: lane i of a wave lands at
// byte 16 i (4 i) of the wave's destination, so each lane fetches the
// chunk whose swizzled slot that is).  A tile row is 64 bytes (four 16-byte chunks); chunk
// c of row r sits at chunk position c ^ f(r), f(r) = ((r >> 1) + (r >> 3))
// & 3, which makes both the stores (4 lanes per row) and the fragment reads
// (consecutive rows, one chunk) hit distinct banks in every 8- or 16-lane
// group.  Tiles are ordered in groups of 4 tile rows and dealt so that the
// 32 blocks one XCD runs together share 4 row panels and 8 column panels
// in its L2.
#include "common.hpp"
#include "cugal_hip.h"

namespace cg {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;

constexpr int kTile = 256;      // block tile (rows and columns)
constexpr int kBK = 32;         // k slice per LDS stage
constexpr int kThreads = 256;   // four waves
constexpr int kRowB = 64;       // LDS bytes per tile row (32 BF16), chunks swizzled
constexpr int kTileB = kTile * kRowB;           // one BF16 operand tile in LDS
constexpr int kWordsB = kTile * 4;              // the bit operand: one word per row
constexpr int kStageB = 2 * kTileB + kWordsB;   // hi + lo + words: 33 KiB
constexpr int kStages = 3;
constexpr int kLdsB = kStages * kStageB;        // three stages: 99 KiB
constexpr int kGroup = 4;       // tile rows per group in the block order
constexpr int kNumXcd = 8;

// LDS chunk swizzle of a tile row (see the file comment).
__device__ __forceinline__ int swz(int row) { return ((row >> 1) + (row >> 3)) & 3; }

__device__ __forceinline__ uint16_t bf16_bits(__bf16 x) { return __builtin_bit_cast(uint16_t, x); }

__device__ __forceinline__ void split_bf16(float x, uint16_t& hi, uint16_t& lo) {
    const __bf16 h = static_cast<__bf16>(x);
    const __bf16 l = static_cast<__bf16>(x - static_cast<float>(h));
    hi = bf16_bits(h);
    lo = bf16_bits(l);
}

// ---- 0/1 adjacency as bits from CSR -------------------------------------
// Word (k / 32) * np + i holds bits k % 32 of row i for k in one 32-wide
// slice; padding rows and columns stay 0.
__global__ __launch_bounds__(256) void pack_rows(const int32_t* __restrict__ col, const int32_t* __restrict__ rp,
                                                int n, int64_t np, uint32_t* __restrict__ out) {
    const int64_t i = blockIdx.x;
    const int b = rp[i];
    const int e = rp[i + 1];
    for (int k = b + static_cast<int>(threadIdx.x); k < e; k += 256) {
        const int j = col[k];
        if (j >= 0 && j < n) atomicOr(&out[static_cast<int64_t>(j >> 5) * np + i], 1u << (j & 31));
    }
}

// Byte `chunk` of a slice word as 8 BF16 values (bit i -> element i: 1.0 or 0).
__device__ __forceinline__ bf16x8 expand_bits(uint32_t word, int chunk) {
    const uint32_t b = (word >> (8 * chunk)) & 0xFFu;
    u32x4 v;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const uint32_t lo = (b >> (2 * p)) & 1u;
        const uint32_t hi = (b >> (2 * p + 1)) & 1u;
        v[p] = (lo * 0x3F80u) | (hi * 0x3F800000u);
    }
    return __builtin_bit_cast(bf16x8, v);
}

// ---- 1. PT_hi / PT_lo = split(P^T), zero padded --------------------------
__global__ __launch_bounds__(256) void split_transpose(const float* __restrict__ P, int64_t ldp, int n, int64_t np,
                                                      uint16_t* __restrict__ hi, uint16_t* __restrict__ lo) {
    __shared__ float tile[64][65];
    const int64_t r0 = static_cast<int64_t>(blockIdx.y) * 64;  // rows of P
    const int64_t c0 = static_cast<int64_t>(blockIdx.x) * 64;  // columns of P
    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;
    for (int i = ty; i < 64; i += 4) {
        const int64_t r = r0 + i, c = c0 + tx;
        tile[i][tx] = (r < n && c < n) ? P[r * ldp + c] : 0.0f;
    }
    __syncthreads();
    for (int j = ty; j < 64; j += 4) {
        uint16_t h, l;
        split_bf16(tile[tx][j], h, l);  // P[r0 + tx][c0 + j]
        const int64_t o = (c0 + j) * np + r0 + tx;
        hi[o] = h;
        lo[o] = l;
    }
}

struct GradEpi {
    const float* P;   // nullptr: no regulariser term
    int64_t ldp;      // row stride of P
    const float* fs;  // source features n x fdim (nullptr: no distance term)
    const float* ft;
    float rl;         // f * lambda
    float* G;         // n x n
    uint16_t* y_hi;   // gemm 1 output (np x np)
    uint16_t* y_lo;
    int fdim;         // 4 or 8
};

// The grid has kNumXcd * ceil(tiles / kNumXcd) blocks; returns false for
// the surplus ones.  Groups run along the first index (ta), which the
// caller binds to the split operand's tiles: 4 of its hi + lo panel pairs
// and 8 panels of the single operand per XCD.
__device__ __forceinline__ bool tile_of_block(int64_t bid, int tiles_1d, int& ta, int& tb) {
    const int64_t tiles = static_cast<int64_t>(tiles_1d) * tiles_1d;
    const int64_t per = div_up(tiles, kNumXcd);
    const int64_t t = (bid % kNumXcd) * per + bid / kNumXcd;  // an XCD's blocks: consecutive t
    if (t >= tiles) return false;
    const int64_t in_group = static_cast<int64_t>(kGroup) * tiles_1d;
    const int64_t g = t / in_group;
    const int64_t first = g * kGroup;
    const int rows = static_cast<int>(min<int64_t>(kGroup, tiles_1d - first));
    const int64_t w = t % in_group;
    ta = static_cast<int>(first + w % rows);
    tb = static_cast<int>(w / rows);
    return true;
}

// C[m][c] = sum_k L[m][k] R[c][k] over k < np, with the split operand on the
// left (SPLIT_L: L = Lh + Ll, R = the bit operand) or on the right (L = the
// bit operand, R = Rh + Rl).  EPI 0: write C as Y_hi / Y_lo; EPI 1: gradient
// epilogue into G.
template <bool SPLIT_L, int EPI, int FD>
__global__ __launch_bounds__(kThreads) void gemm_nt(const uint32_t* __restrict__ Xb,  // bit operand, slice-major
                                                   const uint16_t* __restrict__ Sh,  // split operand, hi
                                                   const uint16_t* __restrict__ Sl,  // split operand, lo
                                                   int n, int64_t np, int tiles_1d, float scale, GradEpi ep) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    int ta, tb;
    if (!tile_of_block(blockIdx.x, tiles_1d, ta, tb)) return;
    const int tm = SPLIT_L ? ta : tb;
    const int tn = SPLIT_L ? tb : ta;
    const int64_t m0 = static_cast<int64_t>(tm) * kTile;
    const int64_t c0 = static_cast<int64_t>(tn) * kTile;
    // Rows of the tiles: the bit operand's and the split operand's (L rows
    // are m, R rows are c).
    const int64_t x_row0 = SPLIT_L ? c0 : m0;
    const int64_t s_row0 = SPLIT_L ? m0 : c0;

    // Global -> LDS per stage.  Split tiles: 256 rows x 64 bytes = 16 KiB
    // each, four 1 KiB wave loads per thread-block instruction; instruction
    // s of wave w fills bytes [1024 (w + 4 s), +1024): lane l lands in row
    // r = 16 (w + 4 s) + l / 4 at chunk slot l % 4, which holds logical
    // chunk (l % 4) ^ swz(r), so the four lanes of a row read its whole
    // 64-byte segment.  Bit words: wave w loads rows 64 w .. 64 w + 63 of
    // the slice (256 contiguous bytes).
    const int lrow = wave * 16 + (lane >> 2);
    const int lchunk = (lane & 3) ^ swz(lrow);  // swz(r + 64 s) == swz(r)
    auto issue_loads = [&](int64_t kt, int stage) {
        char* base = lds + stage * kStageB + wave * 1024;
        const int64_t kx = kt * kBK + lchunk * 8;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int64_t r = lrow + 64 * s;
            __builtin_amdgcn_global_load_lds(Sh + (s_row0 + r) * np + kx,
                                             (__attribute__((address_space(3))) void*)(base + s * 4096), 16, 0, 0);
            __builtin_amdgcn_global_load_lds(Sl + (s_row0 + r) * np + kx,
                                             (__attribute__((address_space(3))) void*)(base + kTileB + s * 4096), 16,
                                             0, 0);
        }
        __builtin_amdgcn_global_load_lds(
            Xb + kt * np + x_row0 + wave * 64 + lane,
            (__attribute__((address_space(3))) void*)(lds + stage * kStageB + 2 * kTileB + wave * 256), 4, 0, 0);
    };
    constexpr int kLoadsPerSlice = 9;

    f32x16 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;

    const int ksteps = static_cast<int>(np / kBK);
    // Fragment rows within the tiles: the wave's 128 rows of L (m) and of R (c).
    const int frag_m = wm * 128 + (lane & 31);
    const int frag_c = wn * 128 + (lane & 31);
    // The lane's 8 k values are chunk 2 ks + lane / 32 of its row; the
    // swizzle is the same for rows 32 apart, so one offset per k step.
    const int frag_sub = lane >> 5;
    const int xr = SPLIT_L ? frag_c : frag_m;
    const int sr = SPLIT_L ? frag_m : frag_c;
    const int swz_s = swz(sr);
    // The split operand's hi (4) and lo (4) fragments of one 16-wide k step.
    auto read_split = [&](const char* base, int ks, bf16x8(&h)[4], bf16x8(&l)[4]) {
        const int ksw = ((2 * ks + frag_sub) ^ swz_s) << 4;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            h[t] = *reinterpret_cast<const bf16x8*>(base + (sr + t * 32) * kRowB + ksw);
            l[t] = *reinterpret_cast<const bf16x8*>(base + kTileB + (sr + t * 32) * kRowB + ksw);
        }
    };
    // The bit operand's words of the slice for the lane's 4 rows.
    auto read_words = [&](const char* base, uint32_t(&w)[4]) {
#pragma unroll
        for (int t = 0; t < 4; ++t) w[t] = *reinterpret_cast<const uint32_t*>(base + 2 * kTileB + (xr + t * 32) * 4);
    };
    // All hi products, then all lo products: a tile's two MFMAs are 16
    // instructions apart, beyond the accumulator latency.  MFMAs
    // [first, first + count) of the step's 32.
    auto mfma_step = [&](const uint32_t(&w)[4], int ks, const bf16x8(&h)[4], const bf16x8(&l)[4], int first,
                         int count) {
        bf16x8 x[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) x[t] = expand_bits(w[t], 2 * ks + frag_sub);
#pragma unroll
        for (int half = 0; half < 2; ++half)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int idx = half * 16 + i * 4 + j;
                    if (idx < first || idx >= first + count) continue;
                    const bf16x8& sv = half ? (SPLIT_L ? l[i] : l[j]) : (SPLIT_L ? h[i] : h[j]);
                    if (SPLIT_L)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sv, x[j], acc[i][j], 0, 0, 0);
                    else
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x[i], sv, acc[i][j], 0, 0, 0);
                }
    };

    // Pipeline: slice kt lives in stage kt % 3.  Iteration kt starts the
    // loads of slice kt + 2, computes slice kt, then waits for slice kt + 1
    // (all but the newest kLoadsPerSlice loads) and synchronises.  The last
    // iterations re-fetch the final slice into a stage no longer read (in
    // bounds), so the load count per iteration, and the wait, stay fixed.
    issue_loads(0, 0);
    issue_loads(min(1, ksteps - 1), 1);
    asm volatile("s_waitcnt vmcnt(9)" ::: "memory");
    __syncthreads();
    uint32_t words[4];
    bf16x8 fh[4], fl[4], gh[4], gl[4];
    read_words(lds, words);
    read_split(lds, 0, fh, fl);
    int stage = 0;
    for (int kt = 0; kt < ksteps; ++kt) {
        const int stage1 = stage == kStages - 1 ? 0 : stage + 1;
        const int stage2 = stage1 == kStages - 1 ? 0 : stage1 + 1;
        issue_loads(min(kt + 2, ksteps - 1), stage2);
        const char* base = lds + stage * kStageB;
        read_split(base, 1, gh, gl);
        mfma_step(words, 0, fh, fl, 0, 32);
        mfma_step(words, 1, gh, gl, 0, 24);
        // Order: the global loads first, then the second step's 8 LDS reads
        // spread over the first step's MFMAs.
        __builtin_amdgcn_sched_group_barrier(0x020, kLoadsPerSlice, 0);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 48, 0);
        // Slice kt + 1 has landed and this wave's reads of the stage the
        // next iteration refills are done; no fence, so the newest loads
        // stay in flight across the barrier.
        asm volatile("s_waitcnt vmcnt(9) lgkmcnt(0)" ::: "memory");
        __builtin_amdgcn_s_barrier();
        uint32_t next_words[4];
        read_words(lds + stage1 * kStageB, next_words);
        read_split(lds + stage1 * kStageB, 0, fh, fl);
        __builtin_amdgcn_sched_barrier(0);
        mfma_step(words, 1, gh, gl, 24, 8);
#pragma unroll
        for (int t = 0; t < 4; ++t) words[t] = next_words[t];
        stage = stage1;
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");

    // Epilogue.  Accumulator r of tile (i, j) is element
    // (m0 + wm*128 + 32 i + 8 (r/4) + 4 (lane/32) + r%4, c0 + wn*128 + 32 j + lane%32).
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t m = m0 + wm * 128 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            FeatureRow fsm{};
            if (EPI == 1 && ep.fs && m < n) fsm = feature_row<FD>(ep.fs, m);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t c = c0 + wn * 128 + j * 32 + (lane & 31);
                const float y = acc[i][j][r] * scale;
                if (EPI == 0) {
                    uint16_t h, l;
                    split_bf16(y, h, l);
                    ep.y_hi[m * np + c] = h;
                    ep.y_lo[m * np + c] = l;
                } else if (m < n && c < n) {
                    float g = y;
                    if (ep.fs) g += feature_distance<FD>(fsm, ep.ft, c);
                    if (ep.P) g += ep.rl - 2.0f * ep.rl * ep.P[m * ep.ldp + c];
                    ep.G[m * n + c] = g;
                }
            }
        }
    }
}

template <bool SPLIT_L, int EPI, int FD>
int launch_gemm(const uint32_t* X, const uint16_t* Sh, const uint16_t* Sl, int n, int64_t np, float scale,
                const GradEpi& ep, hipStream_t s, const char* name) {
    static bool attr = false;
    if (!attr) {
        CG_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&gemm_nt<SPLIT_L, EPI, FD>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, kLdsB));
        attr = true;
    }
    const int tiles_1d = static_cast<int>(np / kTile);
    const int64_t tiles = static_cast<int64_t>(tiles_1d) * tiles_1d;
    const unsigned blocks = static_cast<unsigned>(div_up(tiles, kNumXcd) * kNumXcd);
    hipLaunchKernelGGL((gemm_nt<SPLIT_L, EPI, FD>), dim3(blocks), dim3(kThreads), kLdsB, s, X, Sh, Sl, n, np, tiles_1d,
                       scale, ep);
    return check_launch(name);
}

int64_t padded(int64_t n) { return round_up(n, kTile); }

}  // namespace
}  // namespace cg

using namespace cg;

extern "C" {

size_t cg_adjacency_bits_words(int64_t n) {
    if (n <= 0) return 0;
    const int64_t np = padded(n);
    return static_cast<size_t>(np / 32) * static_cast<size_t>(np);
}

int cg_adjacency_pack_bits(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, uint32_t* out,
                           cg_stream_t stream) {
    if (!col_indices || !row_pointers || !out || n <= 0 || n > (int64_t(1) << 20))
        return fail(CG_EINVAL, "cg_adjacency_pack_bits: bad arguments");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int64_t np = padded(n);
    CG_CHECK_HIP(hipMemsetAsync(out, 0, sizeof(uint32_t) * cg_adjacency_bits_words(n), s));
    hipLaunchKernelGGL(pack_rows, dim3(static_cast<unsigned>(n)), dim3(256), 0, s, col_indices, row_pointers,
                       static_cast<int>(n), np, out);
    return check_launch("pack_rows");
}

size_t cg_gradient_mfma_workspace(int64_t n) {
    if (n <= 0) return 0;
    const size_t e = static_cast<size_t>(padded(n)) * static_cast<size_t>(padded(n));
    Carver cv{nullptr, 0, ~size_t(0)};
    cv.take<uint16_t>(e);
    cv.take<uint16_t>(e);
    cv.take<uint16_t>(e);
    cv.take<uint16_t>(e);
    return cv.used + 256;
}

int cg_gradient_mfma(const uint32_t* a_bits, const uint32_t* b_bits, int64_t n, const float* P, int64_t ld_p,
                     const float* source_features, const float* target_features, int feature_dim, int lambda_,
                     float regularizer_factor, float* gradient, void* workspace, size_t workspace_bytes,
                     cg_stream_t stream) {
    if (!a_bits || !b_bits || !P || !gradient || n <= 0 || n > (int64_t(1) << 20) || ld_p < n)
        return fail(CG_EINVAL, "cg_gradient_mfma: bad arguments");
    if ((source_features == nullptr) != (target_features == nullptr))
        return fail(CG_EINVAL, "cg_gradient_mfma: give both feature matrices or neither");
    if (source_features && (!aligned16(source_features) || !aligned16(target_features)))
        return fail(CG_EINVAL, "cg_gradient_mfma: feature matrices must be 16-byte aligned");
    if (feature_dim != 4 && feature_dim != 8) return fail(CG_EINVAL, "cg_gradient_mfma: feature_dim must be 4 or 8");
    if (!aligned16(a_bits) || !aligned16(b_bits))
        return fail(CG_EINVAL, "cg_gradient_mfma: packed adjacencies must be 16-byte aligned");
    if (!workspace || workspace_bytes < cg_gradient_mfma_workspace(n))
        return fail(CG_EWORKSPACE, "cg_gradient_mfma: workspace too small");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int ni = static_cast<int>(n);
    const int64_t np = padded(n);
    const size_t e = static_cast<size_t>(np * np);
    Carver cv{static_cast<char*>(workspace), 0, workspace_bytes};
    uint16_t* pt_hi = cv.take<uint16_t>(e);
    uint16_t* pt_lo = cv.take<uint16_t>(e);
    uint16_t* y_hi = cv.take<uint16_t>(e);
    uint16_t* y_lo = cv.take<uint16_t>(e);
    const unsigned tb = static_cast<unsigned>(np / 64);
    hipLaunchKernelGGL(split_transpose, dim3(tb, tb), dim3(256), 0, s, P, ld_p, ni, np, pt_hi, pt_lo);
    if (int rc = check_launch("split_transpose")) return rc;
    GradEpi ep1{nullptr, 0, nullptr, nullptr, 0.0f, nullptr, y_hi, y_lo, 4};
    if (int rc = launch_gemm<false, 0, 4>(a_bits, pt_hi, pt_lo, ni, np, 1.0f, ep1, s, "gemm_nt<Y = A P>")) return rc;
    const float rl = regularizer_factor * static_cast<float>(lambda_);
    GradEpi ep2{rl != 0.0f ? P : nullptr, ld_p, source_features, target_features, rl, gradient, nullptr, nullptr,
                feature_dim};
    if (source_features && feature_dim == 8)
        return launch_gemm<true, 1, 8>(b_bits, y_hi, y_lo, ni, np, -2.0f, ep2, s, "gemm_nt<G = -2 Y B>");
    return launch_gemm<true, 1, 4>(b_bits, y_hi, y_lo, ni, np, -2.0f, ep2, s, "gemm_nt<G = -2 Y B>");
}

}  // extern "C"
