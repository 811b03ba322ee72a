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
// 512 (so every tile is whole):
//   1. split_transpose: P^T -> PT_hi, PT_lo (BF16, np x np), through LDS;
//   2. gemm_nt<split R>: Y = A P, i.e. Y[m][c] = sum_k A[m][k] PT[c][k]
//      (hi and lo products into the same accumulators), written back as
//      Y_hi, Y_lo;
//   3. gemm_nt<split L>: G[m][c] = -2 sum_k Y[m][k] B[c][k] (= -2 Y B, B
//      symmetric) with the distance (from the feature rows) and the
//      regulariser fused in the epilogue.
// Both GEMMs are "NT": each operand row holds consecutive k, so an MFMA
// fragment is one 16-byte read (for the bit operand, one byte of a word).
// A block covers 512 rows of the bit operand and 128 of the split operand
// with four waves of 128 x 128 (v_mfma_f32_32x32x16_bf16, 4 x 4 tiles of
// 32 x 32): the split operand, which carries 4 bytes per element (hi +
// lo) against the bit operand's 1/8, is the narrow side, so a 64-wide k
// slice stages 36 KiB (hi and lo tiles of 16 KiB, 4 KiB of words) for as
// many MFMAs as a 256 x 256 tile, whose stage is twice that.  Three stages
// are filled by buffer loads straight into LDS (lane i of a wave lands at
// byte 16 i of the wave's destination, so each lane fetches the chunk
// whose swizzled slot that is); buffer rather than global loads, because
// a pending global load into LDS makes the compiler drain every LDS read
// before the next one is used.  A split row is 128 bytes (eight 16-byte
// chunks); chunk c of row r sits at slot c ^ f(r), f(r) = (r >> 1) & 7, so
// the fragment reads (16 consecutive rows, one logical chunk, per LDS
// cycle) hit 16 distinct 16-byte bank groups.  With one wave per SIMD
// (256 accumulators each), the loop is written in chunks of 4 MFMAs kept
// in order by scheduling barriers, each carrying a share of the next k
// step's LDS reads and bit expansions, so that they issue while the matrix
// pipe is busy.  Tiles are dealt so that the 32 blocks one XCD runs
// together share 2 split panels and 16 bit panels in its L2.
#include <type_traits>

#include "common.hpp"
#include "cugal_hip.h"

namespace cg {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;
template <int V>
using Ic = std::integral_constant<int, V>;  // a compile-time round index

constexpr int kBitTile = 512;   // bit-operand rows per block (4 waves x 128)
constexpr int kSplitTile = 128; // split-operand rows per block
constexpr int kBK = 64;         // k slice per LDS stage (4 MFMA k steps)
constexpr int kThreads = 256;   // four waves
constexpr int kRowB = 128;      // LDS bytes per split row (64 BF16), chunks swizzled
constexpr int kSplitB = kSplitTile * kRowB;     // one BF16 split tile in LDS: 16 KiB
constexpr int kWordsB = kBitTile * 8;           // the bit operand: two words per row
constexpr int kStageB = 2 * kSplitB + kWordsB;  // hi + lo + words: 36 KiB
constexpr int kStages = 3;
constexpr int kLdsB = kStages * kStageB;        // three stages: 108 KiB
constexpr int kSplitGroup = 2;  // split panels per group in the block order
// Word 3 of a raw buffer descriptor on CDNA (32-bit data format, the
// format otherwise unused by raw loads).
constexpr int kRsrcWord3 = 0x00020000;
// s_waitcnt immediates (vmcnt in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt
// 11:8): all but the newest kStages - 2 stages' loads and no LDS
// reads pending; all.
constexpr unsigned kLoadsPerStage = 9;  // per thread: 4 hi + 4 lo + 1 words
constexpr unsigned kLoadsInFlight = kLoadsPerStage * (kStages - 2);
constexpr unsigned kWaitNewestStage = (kLoadsInFlight & 15u) | ((kLoadsInFlight >> 4) << 14) | 0x0070u;
constexpr unsigned kWaitAll = 0x0070;
constexpr int kNumXcd = 8;

// LDS chunk swizzle of a tile row (see the file comment).
__device__ __forceinline__ int swz(int row) { return (row >> 1) & 7; }

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

// Byte `chunk` of a slice word as 8 BF16 values (bit i -> element i: 1.0 or
// 0).  Dword p holds bits 2p (low half) and 2p + 1 (high half).  With
// c = b | (b << 15), bit 2p + 1 of b also sits at bit 16 + 2p of c, so
// c & (0x10001 << 2p) has bit 2p and bit 16 + 2p set as the two elements
// are; times 0x3F80 >> 2p (an integer: 2^7 divides 0x3F80 = BF16 1.0) each
// set bit becomes 0x3F80 in its half.  Both factors are below 2^24, so the
// product is one 24-bit multiply: 10 VALU instructions per fragment.
__device__ __forceinline__ bf16x8 expand_bits(uint32_t word, int chunk) {
    const uint32_t b = (word >> (8 * chunk)) & 0xFFu;
    const uint32_t c = b | (b << 15);
    u32x4 v;
#pragma unroll
    for (int p = 0; p < 4; ++p) v[p] = __umul24(c & (0x10001u << (2 * p)), 0x3F80u >> (2 * p));
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
// the surplus ones.  An XCD's blocks take consecutive tile numbers t, which
// run over pairs of split panels (kSplitGroup) and, within a pair, over all
// bit panels: the 32 blocks one XCD runs together share 2 split panels
// (2 x 32 KiB per stage) and 16 bit panels (16 x 4 KiB) in its L2.
__device__ __forceinline__ bool tile_of_block(int64_t bid, int bit_tiles, int split_tiles, int& tbit, int& tsplit) {
    const int64_t tiles = static_cast<int64_t>(bit_tiles) * split_tiles;
    const int64_t per = div_up(tiles, kNumXcd);
    const int64_t t = (bid % kNumXcd) * per + bid / kNumXcd;
    if (t >= tiles) return false;
    const int64_t in_group = static_cast<int64_t>(kSplitGroup) * bit_tiles;
    const int64_t w = t % in_group;
    tsplit = static_cast<int>((t / in_group) * kSplitGroup + w % kSplitGroup);
    tbit = static_cast<int>(w / kSplitGroup);
    return true;
}

typedef __attribute__((address_space(3))) void* lds_ptr;

// C[m][c] = sum_k L[m][k] R[c][k] over k < np, with the split operand on the
// left (SPLIT_L: L = Lh + Ll, R = the bit operand) or on the right (L = the
// bit operand, R = Rh + Rl).  EPI 0: write C as Y_hi / Y_lo; EPI 1: gradient
// epilogue into G.  A block covers 512 bit-operand rows and 128 split
// rows; wave w owns bit rows 128 w .. 128 w + 127 against all 128 split
// rows (4 x 4 MFMA tiles of 32 x 32).
template <bool SPLIT_L, int EPI, int FD>
__global__ __launch_bounds__(kThreads) void gemm_nt(const uint32_t* __restrict__ Xb,  // bit operand, slice-major
                                                   const uint16_t* __restrict__ Sh,  // split operand, hi
                                                   const uint16_t* __restrict__ Sl,  // split operand, lo
                                                   int n, int64_t np, float scale, GradEpi ep) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    int tbit, tsplit;
    if (!tile_of_block(blockIdx.x, static_cast<int>(np / kBitTile), static_cast<int>(np / kSplitTile), tbit, tsplit))
        return;
    const int64_t x_row0 = static_cast<int64_t>(tbit) * kBitTile;
    const int64_t s_row0 = static_cast<int64_t>(tsplit) * kSplitTile;
    const int64_t m0 = SPLIT_L ? s_row0 : x_row0;
    const int64_t c0 = SPLIT_L ? x_row0 : s_row0;

    // Global -> LDS per stage (buffer loads into LDS: lane i of a wave lands
    // at byte 16 i (4 i) of the wave's destination).  Split tiles: 128 rows
    // x 128 bytes = 16 KiB each; load s of wave w fills bytes
    // [1024 (w + 4 s), +1024): lane l lands in row r = 8 (w + 4 s) + l / 8 at
    // chunk slot l % 8, which holds logical chunk (l % 8) ^ swz(r) (swz(r)
    // does not depend on s).  The buffer descriptors start at the block's
    // split panel; the stage's k offset goes in the scalar offset.  Bit
    // words: plane h (k 32 h .. 32 h + 31 of the stage) of the 512 rows is
    // 2 KiB; wave w loads plane w / 2, rows 256 (w % 2) .. + 255.
    const uint32_t panel_bytes = static_cast<uint32_t>(kSplitTile * np * 2);
    const __amdgpu_buffer_rsrc_t rs_hi =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(Sh + s_row0 * np), 0, panel_bytes, kRsrcWord3);
    const __amdgpu_buffer_rsrc_t rs_lo =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(Sl + s_row0 * np), 0, panel_bytes, kRsrcWord3);
    const int lrow = 8 * wave + (lane >> 3);
    const uint32_t lvoff = static_cast<uint32_t>(lrow * np * 2 + (((lane & 7) ^ swz(lrow)) << 4));
    const uint32_t lstep = static_cast<uint32_t>(32 * np * 2);  // 32 rows further (load s + 1)
    const int plane = wave >> 1;
    const uint32_t wvoff = static_cast<uint32_t>((plane * np + 256 * (wave & 1) + 4 * lane) * 4);
    auto issue_loads = [&](int64_t kt, int stage) {
        char* base = lds + stage * kStageB + wave * 1024;
        const uint32_t koff = static_cast<uint32_t>(kt * kBK * 2);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs_hi, (lds_ptr)(base + s * 4096), 16, lvoff + s * lstep, koff, 0,
                                                     0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs_lo, (lds_ptr)(base + kSplitB + s * 4096), 16,
                                                     lvoff + s * lstep, koff, 0, 0);
        }
        // The stage's two word planes start at slice 2 kt of the bit operand.
        const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint32_t*>(Xb + 2 * kt * np + x_row0), 0, static_cast<uint32_t>(8 * np), kRsrcWord3);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rw, (lds_ptr)(lds + stage * kStageB + 2 * kSplitB + plane * 2048 + (wave & 1) * 1024), 16, wvoff, 0, 0, 0);
    };
    // MFMAs of a stage's last k step issued after the barrier.
    constexpr int kTailMfma = 8;

    f32x16 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;

    const int nst = static_cast<int>(np / kBK);
    // The lane's fragment rows: split rows sr + 32 t, bit rows
    // 128 wave + xr + 32 t.  Its 8 k values of step ks are logical chunk
    // 2 ks + lane / 32 of a split row and byte 2 (ks % 2) + lane / 32 of
    // word plane ks / 2 of a bit row.
    const int frag_sub = lane >> 5;
    const int sr = lane & 31;
    const int swz_s = swz(sr);  // swz(sr + 32 t) == swz(sr)
    const int xr = 128 * wave + (lane & 31);
    // One split fragment (hi or lo, row sr + 32 t) of step ks.
    auto split_frag = [&](const char* stage_base, bool lo, int ks, int t) {
        const int slot = (2 * ks + frag_sub) ^ swz_s;
        return *reinterpret_cast<const bf16x8*>(stage_base + (lo ? kSplitB : 0) + (sr + 32 * t) * kRowB + slot * 16);
    };
    auto read_words = [&](const char* stage_base, uint32_t(&w)[2][4]) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int t = 0; t < 4; ++t)
                w[h][t] = *reinterpret_cast<const uint32_t*>(stage_base + 2 * kSplitB + h * 2048 + (xr + 32 * t) * 4);
    };
    auto bit_frag = [&](const uint32_t(&w)[2][4], int ks, int t) {
        return expand_bits(w[ks >> 1][t], 2 * (ks & 1) + frag_sub);
    };
    // All hi products, then all lo products: a tile's two MFMAs are 16
    // instructions apart.  MFMAs [first, first + count) of the step's 32.
    auto mfma_step = [&](const bf16x8(&x)[4], const bf16x8(&h)[4], const bf16x8(&l)[4], int first, int count) {
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
    // Step ks of the stage at `base` with fragments (x, h, l), in chunks of
    // 4 MFMAs kept in order by scheduling barriers; the first four chunks
    // also fetch step ks + 1's fragments (hi, then lo, one bit fragment per
    // chunk) into (nx, nh, nl).  With one wave per SIMD, the reads and the
    // expansions only overlap the matrix pipe when issued between MFMAs.
    auto step_with_prefetch = [&](const char* base, const uint32_t(&w)[2][4], int ks, const bf16x8(&x)[4],
                                  const bf16x8(&h)[4], const bf16x8(&l)[4], bf16x8(&nx)[4], bf16x8(&nh)[4],
                                  bf16x8(&nl)[4], int chunks) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (c >= chunks) break;
            __builtin_amdgcn_sched_barrier(0);
            if (c < 4) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int t = 2 * (c & 1) + u;
                    if (c < 2)
                        nh[t] = split_frag(base, false, ks + 1, t);
                    else
                        nl[t] = split_frag(base, true, ks + 1, t);
                }
                nx[c] = bit_frag(w, ks + 1, c);
            }
            mfma_step(x, h, l, 4 * c, 4);
        }
    };

    // Pipeline: stage kt % kStages holds k slice kt (64 wide).  Iteration
    // kt starts the loads of slice kt + kStages - 1 into the stage read in
    // iteration kt - 1, runs its four k steps, then waits for slice kt + 1
    // (all but the newest kStages - 2 stages' loads) and synchronises.  The last iterations re-fetch the final slice into a
    // stage no longer read (in bounds), so the load count per iteration, and
    // the wait, stay fixed.
#pragma unroll
    for (int p = 0; p < kStages - 1; ++p) issue_loads(min(p, nst - 1), p);
    __builtin_amdgcn_s_waitcnt(kWaitNewestStage);
    __syncthreads();
    uint32_t words[2][4];
    bf16x8 ax[4], ah[4], al[4], bx[4], bh[4], bl[4];
    read_words(lds, words);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        ah[t] = split_frag(lds, false, 0, t);
        al[t] = split_frag(lds, true, 0, t);
        ax[t] = bit_frag(words, 0, t);
    }
    int stage = 0;
    for (int kt = 0; kt < nst; ++kt) {
        const int stage1 = stage == kStages - 1 ? 0 : stage + 1;
        const int refill = stage == 0 ? kStages - 1 : stage - 1;  // read in iteration kt - 1
        issue_loads(min(kt + kStages - 1, nst - 1), refill);
        const char* base = lds + stage * kStageB;
        step_with_prefetch(base, words, 0, ax, ah, al, bx, bh, bl, 8);
        step_with_prefetch(base, words, 1, bx, bh, bl, ax, ah, al, 8);
        step_with_prefetch(base, words, 2, ax, ah, al, bx, bh, bl, 8);
        // Step 3 up to its last kTailMfma MFMAs.
#pragma unroll
        for (int c = 0; c < (32 - kTailMfma) / 4; ++c) {
            __builtin_amdgcn_sched_barrier(0);
            mfma_step(bx, bh, bl, 4 * c, 4);
        }
        // The MFMAs are register-only and would otherwise be moved past the
        // barrier, leaving the matrix pipe idle while the wave waits there.
        __builtin_amdgcn_sched_barrier(0);
        // Slice kt + 1 has landed and this wave's reads of the stage the
        // next iteration refills are done; no fence, so the newest loads
        // stay in flight across the barrier.
        __builtin_amdgcn_s_waitcnt(kWaitNewestStage);
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
        // Step 3's last MFMAs cover the next slice's LDS reads (words
        // first, so its first bit fragments can be expanded while the
        // split fragments are in flight) and the expansions.
        const char* next = lds + stage1 * kStageB;
        read_words(next, words);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            ah[t] = split_frag(next, false, 0, t);
            al[t] = split_frag(next, true, 0, t);
        }
        constexpr int kPer = kTailMfma / 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            __builtin_amdgcn_sched_barrier(0);
            ax[c] = bit_frag(words, 0, c);
            mfma_step(bx, bh, bl, 32 - kTailMfma + kPer * c, kPer);
        }
        __builtin_amdgcn_sched_barrier(0);
        stage = stage1;
    }
    __builtin_amdgcn_s_waitcnt(kWaitAll);

    // Epilogue.  Accumulator r of tile (i, j) is element (m, c) with
    // m = m0 + [128 wave] + 32 i + 8 (r/4) + 4 (lane/32) + r%4 and
    // c = c0 + [128 wave] + 32 j + lane%32, the wave offset on the bit side.
    const int64_t wm_off = SPLIT_L ? 0 : 128 * wave;
    const int64_t wc_off = SPLIT_L ? 128 * wave : 0;
    auto row_of = [&](int i, int r) { return m0 + wm_off + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); };
    int64_t cols[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) cols[j] = c0 + wc_off + j * 32 + (lane & 31);
    if constexpr (EPI == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t m = row_of(i, r);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    uint16_t h, l;
                    split_bf16(acc[i][j][r] * scale, h, l);
                    ep.y_hi[m * np + cols[j]] = h;
                    ep.y_lo[m * np + cols[j]] = l;
                }
            }
    } else {
        // Eight rounds of 32 outputs (tile row i, accumulators 8 h .. 8 h +
        // 7); round q + 1's loads (P and the source feature rows) are issued
        // before round q's stores, so no load waits behind a store.  The
        // lane's four target feature rows are loaded once.
        FeatureRow ftr[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            ftr[j] = FeatureRow{};
            if (ep.fs && cols[j] < n) ftr[j] = feature_row<FD>(ep.ft, cols[j]);
        }
        struct Round {
            float p[8][4];
            FeatureRow f[8];
        };
        auto load = [&](auto qc, Round& R) {
            constexpr int q = decltype(qc)::value;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int64_t m = row_of(q >> 1, 8 * (q & 1) + u);
                R.f[u] = FeatureRow{};
                if (ep.fs && m < n) R.f[u] = feature_row<FD>(ep.fs, m);
#pragma unroll
                for (int j = 0; j < 4; ++j) R.p[u][j] = (ep.P && m < n && cols[j] < n) ? ep.P[m * ep.ldp + cols[j]] : 0.0f;
            }
        };
        auto store = [&](auto qc, const Round& R) {
            constexpr int q = decltype(qc)::value;
            constexpr int i = q >> 1;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int r = 8 * (q & 1) + u;
                const int64_t m = row_of(i, r);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (m >= n || cols[j] >= n) continue;
                    float g = acc[i][j][r] * scale;
                    if (ep.fs) g += feature_distance_rows<FD>(R.f[u], ftr[j]);
                    if (ep.P) g += ep.rl - 2.0f * ep.rl * R.p[u][j];
                    ep.G[m * n + cols[j]] = g;
                }
            }
        };
        Round ra, rb;
        load(Ic<0>{}, ra);
        load(Ic<1>{}, rb);
        store(Ic<0>{}, ra);
        load(Ic<2>{}, ra);
        store(Ic<1>{}, rb);
        load(Ic<3>{}, rb);
        store(Ic<2>{}, ra);
        load(Ic<4>{}, ra);
        store(Ic<3>{}, rb);
        load(Ic<5>{}, rb);
        store(Ic<4>{}, ra);
        load(Ic<6>{}, ra);
        store(Ic<5>{}, rb);
        load(Ic<7>{}, rb);
        store(Ic<6>{}, ra);
        store(Ic<7>{}, rb);
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
    const int64_t tiles = (np / kBitTile) * (np / kSplitTile);
    const unsigned blocks = static_cast<unsigned>(div_up(tiles, kNumXcd) * kNumXcd);
    hipLaunchKernelGGL((gemm_nt<SPLIT_L, EPI, FD>), dim3(blocks), dim3(kThreads), kLdsB, s, X, Sh, Sl, n, np, scale,
                       ep);
    return check_launch(name);
}

int64_t padded(int64_t n) { return round_up(n, kBitTile); }

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
