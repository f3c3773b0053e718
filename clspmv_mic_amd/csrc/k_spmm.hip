// Mixed-precision block SpMM for gfx950: Y += A X with A a BCSR matrix of
// 4x4 blocks whose values are stored in BF16 or FP16, X (k right-hand
// sides) in the same type, and Y in FP32.  The products run on the matrix
// cores: v_mfma_f32_4x4x4 (16 blocks per wave) multiplies, for each group of
// four lanes, one 4x4 block of A by a 4x4 tile of X (four rows of four
// right-hand sides) and accumulates the 4x4 tile of Y in FP32.
//
// Operand layout of the instruction (16 independent 4x4x4 products per
// wave, lane l = 4*g + t for group g):
//   A[i][k]: lane 4g+i, element k      -> row t of the A block (8 bytes)
//   B[k][j]: lane 4g+j, element k      -> X rows c0..c0+3 of column j
//   D[i][j]: lane 4g+j, register i     -> Y rows r0..r0+3 of column j
// so with X and Y column-major every lane moves whole 8- or 16-byte pieces.
// X may instead come in 4-row tiles (TILED): the four 16-bit values of rows
// 4c'..4c'+3 of right-hand side j at X[(c' * kt + j) * 4], so the four lanes
// of a group read one 32-byte run per quad instead of four columns that lie
// megabytes apart -- a quarter of the cache lines per block for the texture
// path, which bounds the column-major form.
// A group owns one block row and all right-hand sides; each Y element has
// one owner, so Y is updated without atomics and the summation order
// (blocks in storage order, four products per instruction) is fixed.

#include <algorithm>
#include <type_traits>

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 h16x4 __attribute__((ext_vector_type(4)));

template <bool BF16>
__device__ __forceinline__ f32x4 mfma4(s16x4 a, s16x4 b, f32x4 c) {
    if constexpr (BF16) {
        return __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a, b, c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_4x4x4f16(__builtin_bit_cast(h16x4, a),
                                                   __builtin_bit_cast(h16x4, b), c, 0, 0, 0);
    }
}

// Loads 4 consecutive 16-bit values at p (8-byte aligned) or, near the end
// of a column, the ones below `lim` and zeros after them.
__device__ __forceinline__ s16x4 ld4(const unsigned short* __restrict__ p, int avail) {
    if (avail >= 4) return *reinterpret_cast<const s16x4*>(p);
    s16x4 v = {0, 0, 0, 0};
    if (avail > 0) v[0] = (short)p[0];
    if (avail > 1) v[1] = (short)p[1];
    if (avail > 2) v[2] = (short)p[2];
    return v;
}

// A group of four lanes owns one block row and NQ quads of four right-hand
// sides (column-major X: every right-hand side up to 32): each A block it
// loads is multiplied by NQ tiles of X.  Lane t of a group loads row
// t of the block and, per quad q, column 4q + t of X.  A wave holds a slice
// of 16 block rows whose step-j blocks are stored together: its A load of a
// step is one 512-byte run and its column load one 64-byte run.
// LONG (matrices with a block row longer than kLongRowF64 / 4 blocks): the
// FP32 MFMA accumulators are added into FP64 ones every kSpmmChunk blocks,
// so a row's rounding error grows with the chunk, not the row length, and
// y receives one rounding.
// ldx: X's leading dimension, or for TILED the tile stride in right-hand
// sides (the k of the whole tiled X).  With TILED, lane t of a group takes
// right-hand sides t * NQ + q (q < NQ), adjacent tiles, so one 16-byte load
// serves two of them when the stride is even (X16); a right-hand side past
// k reads some in-bounds tile of the chunk -- its column of the product is
// never stored, and an MFMA output column depends only on its own B column.
// NT (TILED, at most 4 * NQ right-hand sides): one wave per slice reads A
// and its block columns once, with non-temporal loads, so they do not evict
// the X tiles that neighbouring slices re-read.  With more right-hand sides
// each slice gets one wave per chunk of 4 * NQ of them, the chunks' waves
// neighbours in one work-group: they run together and read the slice's
// blocks at about the same time, so A comes from memory about once
// (1.03 x the algorithmic bytes at k = 32, against 1.41 x for one wave
// running the chunks one after another, at the same speed); A is then
// loaded normally so that the neighbours find it in cache.
template <bool BF16, int NQ, bool LONG, bool TILED, bool X16, bool NT>
__global__ __launch_bounds__(kBlock) void bspmm4_kernel(
    int brows, int rows, int cols, int kall, const int* __restrict__ rp,
    const int* __restrict__ bcol, const unsigned short* __restrict__ val,
    const unsigned short* __restrict__ Xall, int ldx, float* __restrict__ Yall, int ldy, bool acc_y) {
    static_assert(!X16 || (TILED && NQ >= 2), "16-byte loads take two tiles of the tiled layout");
    static_assert(!NT || TILED, "non-temporal A loads belong to the tiled kernel");
    static_assert(launch::kSpmmSlice * 4 == kWave, "a wave holds one slice, four lanes per block row");
    const int lane = threadIdx.x & (kWave - 1);
    const int t = lane & 3, g = lane >> 2;
    // One wave per slice and chunk of right-hand sides, in XCD-aware order:
    // neighbouring block rows, which share X rows, run on one XCD and share
    // its L2.  (Wave-uniform: in scalar registers, so the loop's exits are
    // scalar branches.)
    const int nch = (TILED && !NT) ? (kall + 4 * NQ - 1) / (4 * NQ) : 1;
    const int gw = __builtin_amdgcn_readfirstlane(
        (int)(((long long)xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x) / kWave));
    const int sl = gw / nch, ch = gw - sl * nch;
    if (sl >= (brows + launch::kSpmmSlice - 1) / launch::kSpmmSlice) return;  // the whole wave
    const int br = sl * launch::kSpmmSlice + g;
    const bool active = br < brows;
    // The instruction is wave-wide: every row of the slice steps through the
    // slice's width, padding slots feeding zeros.
    const int base = __builtin_amdgcn_readfirstlane(rp[sl]);
    const int nmax = __builtin_amdgcn_readfirstlane((rp[sl + 1] - base) / launch::kSpmmSlice);
    // One pass over the slice for right-hand sides [0, k) of X and Y.
    auto run = [&](const unsigned short* __restrict__ X, float* __restrict__ Y, int k) {
        f32x4 acc[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        // Column-major X: operands of step `it` (zeros for padding slots and
        // past the width).
        auto fetch = [&](int it, s16x4& a, s16x4* b, bool& use) {
            a = s16x4{0, 0, 0, 0};
            int c0 = -1;
            if (it < nmax) {
                const size_t j = (size_t)base + (size_t)it * launch::kSpmmSlice + g;
                a = *reinterpret_cast<const s16x4*>(val + j * 16 + t * 4);  // padding holds zeros
                c0 = bcol[j];
            }
            use = c0 >= 0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int xc = q * 4 + t;
                b[q] = s16x4{0, 0, 0, 0};
                if (use && xc < k) b[q] = ld4(X + (size_t)xc * ldx + c0, cols - c0);
            }
        };
        // Software pipeline: block it + 1's loads are in flight while block it
        // is multiplied.
        s16x4 a_cur, a_nxt, b_cur[NQ], b_nxt[NQ];
        bool use_cur, use_nxt;
        double acc64[NQ][4];  // used by LONG only
        if constexpr (LONG) {
#pragma unroll
            for (int q = 0; q < NQ; ++q)
#pragma unroll
                for (int i = 0; i < 4; ++i) acc64[q][i] = 0.0;
        }
        auto carry = [&]() {  // FP32 chunk sums into FP64 (LONG only)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
#pragma unroll
                for (int i = 0; i < 4; ++i) acc64[q][i] += (double)acc[q][i];
                acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            }
        };
        if constexpr (TILED) {
            // Two register sets, P for even steps and Q for odd ones, take
            // turns with nothing copied between them: a step's block column is
            // loaded two steps ahead, its A values and X tiles one step ahead.
            // Every load is unconditional (steps past the end re-read the last
            // slot and are never multiplied), so each wait names an exact count
            // and the next step's loads stay in flight during the MFMAs.
            if (nmax > 0) {
                auto slot = [&](int s) {
                    return (size_t)base + (size_t)min(s, nmax - 1) * launch::kSpmmSlice + g;
                };
                auto ld_col = [&](int s) {
                    if constexpr (NT) return __builtin_nontemporal_load(bcol + slot(s));
                    else return bcol[slot(s)];
                };
                auto ld_a = [&](int s) {
                    const s16x4* p = reinterpret_cast<const s16x4*>(val + slot(s) * 16 + t * 4);
                    if constexpr (NT) return __builtin_nontemporal_load(p);
                    else return *p;
                };
                auto ld_x = [&](int c0, s16x4* b) {
                    // Padding slots (c0 < 0) read tile 0 and are zeroed in mma.
                    const unsigned short* xt = X + (size_t)(c0 >= 0 ? c0 >> 2 : 0) * ldx * 4;
                    if constexpr (X16) {
#pragma unroll
                        for (int p = 0; p < NQ / 2; ++p) {
                            const int r = min(t * NQ + 2 * p, k - 2);
                            const s16x8 v = *reinterpret_cast<const s16x8*>(xt + (size_t)r * 4);
                            b[2 * p] = v.lo;
                            b[2 * p + 1] = v.hi;
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < NQ; ++q)
                            b[q] = *reinterpret_cast<const s16x4*>(xt + (size_t)min(t * NQ + q, k - 1) * 4);
                    }
                };
                auto mma = [&](int s, const s16x4& a, const s16x4* b, bool use) {
                    const s16x4 z = s16x4{0, 0, 0, 0}, au = use ? a : z;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) acc[q] = mfma4<BF16>(au, use ? b[q] : z, acc[q]);
                    if constexpr (LONG) {
                        if (s % launch::kSpmmChunk == launch::kSpmmChunk - 1) carry();  // wave-uniform
                    }
                };
                // Steps go in pairs with no exit inside a pair (a load used only
                // past an exit would be moved down to its use): an odd last step
                // pairs with one past the end, multiplied as zeros (which can
                // turn a zero sum of -0 into +0; nothing else changes).
                int cP = ld_col(0), cQ = ld_col(1);
                s16x4 aP = ld_a(0), aQ = ld_a(1);
                s16x4 xP[NQ], xQ[NQ];
                bool uP = cP >= 0, uQ = false;
                ld_x(cP, xP);
                cP = ld_col(2);
                for (int it = 0; it < nmax; it += 2) {
                    uQ = cQ >= 0 && it + 1 < nmax;
                    ld_x(cQ, xQ);
                    cQ = ld_col(it + 3);
                    mma(it, aP, xP, uP);
                    aP = ld_a(it + 2);
                    uP = cP >= 0 && it + 2 < nmax;
                    ld_x(cP, xP);
                    cP = ld_col(it + 4);
                    mma(it + 1, aQ, xQ, uQ);
                    aQ = ld_a(it + 3);
                }
            }
        } else {
            fetch(0, a_cur, b_cur, use_cur);
            for (int it = 0; it < nmax; ++it) {
                fetch(it + 1, a_nxt, b_nxt, use_nxt);
#pragma unroll
                for (int q = 0; q < NQ; ++q) acc[q] = mfma4<BF16>(a_cur, b_cur[q], acc[q]);
                a_cur = a_nxt;
#pragma unroll
                for (int q = 0; q < NQ; ++q) b_cur[q] = b_nxt[q];
                if constexpr (LONG) {
                    if (it % launch::kSpmmChunk == launch::kSpmmChunk - 1) carry();  // wave-uniform
                }
            }
        }
        if constexpr (LONG) carry();
        if (!active) return;
        const int r0 = br * 4;
        if constexpr (LONG) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int xc = TILED ? t * NQ + q : q * 4 + t;
                if (xc >= k) continue;
                float* yp = Y + (size_t)xc * ldy + r0;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (r0 + i < rows) yp[i] = (float)((acc_y ? (double)yp[i] : 0.0) + acc64[q][i]);
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int xc = TILED ? t * NQ + q : q * 4 + t;
            if (xc >= k) continue;
            float* yp = Y + (size_t)xc * ldy + r0;
            if (r0 + 3 < rows) {
                // Y += A X, or Y = A X with Y not read (then 0 + the sum, the
                // bits a zeroed Y would give).  With several chunks (tiled,
                // more than 4 * NQ right-hand sides) Y goes through
                // non-temporal accesses, which leave the cache to the
                // slice's blocks that the neighbouring chunks' waves re-read
                // (3 % faster at k = 32 than plain accesses).
                f32x4* yq = reinterpret_cast<f32x4*>(yp);
                f32x4 y = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                if (acc_y) {
                    if constexpr (TILED && !NT) y = __builtin_nontemporal_load(yq);
                    else y = *yq;
                }
                y += acc[q];
                if constexpr (TILED && !NT) __builtin_nontemporal_store(y, yq);
                else *yq = y;
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (r0 + i < rows) yp[i] = (acc_y ? yp[i] : 0.0f) + acc[q][i];
            }
        }
    };
    if constexpr (TILED) {
        if constexpr (NT) {
            run(Xall, Yall, kall);  // kall <= 4 * NQ
            return;
        }
        const int c = ch * 4 * NQ;  // this wave's chunk of right-hand sides
        run(Xall + (size_t)c * 4, Yall + (size_t)c * ldy, min(4 * NQ, kall - c));
    } else {
        run(Xall, Yall, kall);
    }
}

// X windows (no FP64 carries; the upload lists each slice's distinct block
// columns): one wave per slice, for every chunk of 4 * NQ right-hand sides
// in turn.  The wave copies the tiles of its slice's window
// (win.tile[sl * win.max ..]) for the chunk into its own win_stride bytes of
// LDS, kWinTile bytes per tile -- then a spare slot that takes the stores of
// idle lanes, a zero tile for padding slots and the window's tile list (4
// bytes per tile, for the chunks after the first) -- and the steps read X
// from there by each slot's window index (win.idx).  A block
// column is fetched once per slice and chunk, with wide contiguous loads,
// instead of once per step through the texture path, where the tiles of 16
// block rows per load cost one cache-line access each and left the steps
// waiting on L2 (k = 32 ran at 193 us; with every tile load hitting L1 it
// would run at 126 us, with every load of one line at 84 us).  The slice's A
// values and window indices are loaded once for all chunks, non-temporally.
// X comes in 4-row tiles (16-byte pieces: two right-hand sides of a tile)
// or, with COLX, column-major (8-byte pieces: the tile's four rows of one
// right-hand side, read from its column; rows past `cols` staged as zeros).
template <int NQ>
constexpr int kWinTile = 8 * 4 * NQ + 16;  // bytes per staged tile: 4 * NQ right-hand sides + 16-byte pad
constexpr int kWinBlock = kBlock;  // (one-wave work-groups measured no faster)

// MULTI: more than 4 * NQ right-hand sides (the chunk loop runs more than
// once).
template <bool BF16, int NQ, bool MULTI, bool COLX>
__global__ __launch_bounds__(kWinBlock) __attribute__((amdgpu_waves_per_eu(3))) void bspmm4_win_kernel(
    int brows, int rows, int cols, int kall, const int* __restrict__ rp, const unsigned short* __restrict__ val,
    launch::SpmmWindows win, int win_stride, const unsigned short* __restrict__ Xall, int ldx,
    float* __restrict__ Yall, int ldy, bool acc_y) {
    static_assert(NQ >= 2, "a lane reads the window in 16-byte pieces of two right-hand sides");
    // Two blocks of kD steps' A values and window indices in registers, each
    // refilled kD * 2 steps ahead once its steps are multiplied; both are in
    // flight while the window is staged, so a slice up to 2 kD blocks wide
    // waits on memory once.  (At 3 waves per SIMD, which the windows' LDS
    // leaves, loads two steps ahead left the steps waiting on memory: 111
    // against 103 us without windows at k = 16.)  The loads of a staging
    // round are all in flight together: a round takes up to 96 tiles of tiled
    // X (kR 16-byte pieces per lane) or 128 of column-major X (two tiles per
    // lane).
    constexpr int kD = 6, PP = 2 * NQ, kR = 3 * NQ, KC = 4 * NQ;
    extern __shared__ __align__(16) unsigned char spmm_lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int t = lane & 3, g = lane >> 2;
    const int sl = __builtin_amdgcn_readfirstlane(
        (int)(((long long)xcd_tile(blockIdx.x, gridDim.x) * kWinBlock + threadIdx.x) / kWave));
    if (sl >= (brows + launch::kSpmmSlice - 1) / launch::kSpmmSlice) return;  // the whole wave
    const int br = sl * launch::kSpmmSlice + g;
    const int base = __builtin_amdgcn_readfirstlane(rp[sl]);
    const int nmax = __builtin_amdgcn_readfirstlane((rp[sl + 1] - base) / launch::kSpmmSlice);
    // Loads from scalar bases with 32-bit lane offsets: the steps' address
    // arithmetic stays in scalar registers.
    const char* const a_base = reinterpret_cast<const char*>(val + (size_t)base * 16);
    const unsigned char* const i_base = win.idx + base;
    const unsigned a_lane = (unsigned)(g * 32 + t * 8);
    auto step_at = [&](int s) { return (unsigned)min(s, nmax - 1) * launch::kSpmmSlice; };
    s16x4 aS[2][kD];
    int cS[2][kD];
    auto fill = [&](int h, int s0) {
#pragma unroll
        for (int d = 0; d < kD; ++d) {
            cS[h][d] = i_base[step_at(s0 + d) + (unsigned)g];
            aS[h][d] = __builtin_nontemporal_load(
                reinterpret_cast<const s16x4*>(a_base + step_at(s0 + d) * 32 + a_lane));
        }
    };
    if (nmax > 0) {
        fill(0, 0);
        fill(1, kD);
    }
    unsigned char* const wbuf = spmm_lds + (threadIdx.x / kWave) * win_stride;
    const int spare = win.max, zero = win.max + 1;
    const int* const w_list = win.tile + (size_t)sl * win.max;
    const unsigned x_tile = (unsigned)ldx * 8;  // bytes per tile of X
    const int npc = win.max * PP;
    int* const wlist_lds = reinterpret_cast<int*>(wbuf + (win.max + 2) * kWinTile<NQ>);
    if (lane < PP)
        *reinterpret_cast<s16x8*>(wbuf + zero * kWinTile<NQ> + lane * 16) = s16x8{0, 0, 0, 0, 0, 0, 0, 0};
    // Staging of a chunk's window from tiled X: 16-byte piece p of tile i
    // (right-hand sides 2p, 2p + 1 of the chunk) to wbuf[i * kWinTile + 16p],
    // kR pieces per lane per round, stored with no branch: pieces past the
    // window go to the spare slot, pieces past the chunk's k (re-reads of its
    // first piece) stay in their tile, where they feed only right-hand sides
    // that are never stored.  The first chunk reads the window's tile list
    // from memory and keeps it in LDS (after the zero tile) for the later
    // chunks.
    auto st_load = [&](int cc, int i0, s16x8* v) {
        const char* const x_base = reinterpret_cast<const char*>(Xall + (size_t)cc * 4);
        const int kc = MULTI ? min(KC, kall - cc) : kall;
        int tl[kR];
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const unsigned j = (unsigned)min(i0 + u * kWave + lane, npc - 1) / PP;
            tl[u] = (MULTI && cc > 0) ? wlist_lds[j] : w_list[j];
        }
        if (MULTI && cc == 0) {
#pragma unroll
            for (int u = 0; u < kR; ++u) {
                const int i = i0 + u * kWave + lane;
                if (i < npc && i % PP == 0) wlist_lds[i / PP] = tl[u];
            }
        }
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const unsigned pc = (unsigned)(i0 + u * kWave + lane) % PP;
            v[u] = *reinterpret_cast<const s16x8*>(x_base + (unsigned)tl[u] * x_tile +
                                                  (2 * pc < (unsigned)kc ? pc * 16 : 0u));
        }
    };
    auto st_store = [&](int i0, const s16x8* v) {
#pragma unroll
        for (int u = 0; u < kR; ++u) {
            const int i = i0 + u * kWave + lane;
            const int tile = i < npc ? i / PP : spare;
            *reinterpret_cast<s16x8*>(wbuf + tile * kWinTile<NQ> + (i % PP) * 16) = v[u];
        }
    };
    // Staging from column-major X: lane L of a round takes window tile j0 + L
    // and loads its four rows from the column of each of the chunk's
    // right-hand sides (in bounds: ldx is a multiple of 4 at least cols), so
    // one load reads one column at the window's tiles -- runs of consecutive
    // tiles, a few cache lines -- and stores them at wbuf[j * kWinTile + 8 r]
    // (the spare slot past the window); rows past `cols` are staged as zeros.
    auto st_colx = [&](int cc) {
        const char* const x_base = reinterpret_cast<const char*>(Xall + (size_t)cc * ldx);
        const int kc = MULTI ? min(KC, kall - cc) : kall;
        constexpr int kT = 2;  // tiles per lane per round: one round for up to 128 tiles
        for (int j0 = 0; j0 < win.max; j0 += kT * kWave) {
            int tl[kT];
#pragma unroll
            for (int h = 0; h < kT; ++h) {
                const int j = j0 + h * kWave + lane, jc = min(j, win.max - 1);
                tl[h] = (MULTI && cc > 0) ? wlist_lds[jc] : w_list[jc];
                if (MULTI && cc == 0 && j < win.max) wlist_lds[j] = tl[h];
            }
            s16x4 v[kT][KC];
#pragma unroll
            for (int h = 0; h < kT; ++h)
#pragma unroll
                for (int r = 0; r < KC; ++r)
                    v[h][r] = *reinterpret_cast<const s16x4*>(
                        x_base + ((r < kc ? (unsigned)r : 0u) * (unsigned)ldx + (unsigned)tl[h] * 4) * 2);
            if (cols & 3) {  // wave-uniform: zeros past the last column
#pragma unroll
                for (int h = 0; h < kT; ++h) {
                    const int past = tl[h] * 4 - cols;  // rows 4 tl + e with e >= -past are past
#pragma unroll
                    for (int r = 0; r < KC; ++r)
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            if (e + past >= 0) v[h][r][e] = 0;
                }
            }
#pragma unroll
            for (int h = 0; h < kT; ++h) {
                const int j = j0 + h * kWave + lane;
                unsigned char* const dst = wbuf + (j < win.max ? j : spare) * kWinTile<NQ>;
#pragma unroll
                for (int r = 0; r < KC; ++r) *reinterpret_cast<s16x4*>(dst + r * 8) = v[h][r];
            }
        }
    };
    // The wave reads only what its own lanes wrote and its LDS accesses
    // complete in order; the fences keep the compiler from moving a chunk's
    // reads across the writes of the window around them.
    auto lds_fence = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    for (int c = 0; c < (MULTI ? kall : 1); c += KC) {
        const int k = MULTI ? min(KC, kall - c) : kall;
        f32x4 acc[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        if (nmax > 0) {
            lds_fence();
            if constexpr (COLX) {
                st_colx(c);
            } else {
                for (int i0 = 0; i0 < npc; i0 += kR * kWave) {
                    s16x8 v[kR];
                    st_load(c, i0, v);
                    st_store(i0, v);
                }
            }
            lds_fence();
            // Lane t of a group takes right-hand sides t * NQ + q.
            auto ld_x = [&](int w, s16x4* b) {
                const unsigned char* wt = wbuf + w * kWinTile<NQ> + t * NQ * 8;
#pragma unroll
                for (int p = 0; p < NQ / 2; ++p) {
                    const s16x8 v = *reinterpret_cast<const s16x8*>(wt + p * 16);
                    b[2 * p] = v.lo;
                    b[2 * p + 1] = v.hi;
                }
            };
            // Padding slots hold zero values and the zero tile's index; steps
            // past the slice's width are skipped, so the products and their
            // order are those of the column-major kernel.
            s16x4 xs[2][NQ];
            for (int b0 = 0; b0 < nmax; b0 += 2 * kD) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int sb = b0 + h * kD;
                    ld_x(cS[h][0], xs[0]);
#pragma unroll
                    for (int d = 0; d < kD; ++d) {
                        if (d + 1 < kD) ld_x(cS[h][d + 1], xs[(d + 1) & 1]);
                        if (sb + d < nmax) {  // wave-uniform
#pragma unroll
                            for (int q = 0; q < NQ; ++q) acc[q] = mfma4<BF16>(aS[h][d], xs[d & 1][q], acc[q]);
                        }
                    }
                    // Refill kD * 2 steps ahead; in a slice wider than 2 kD
                    // steps, the last refills load the first blocks again for
                    // the next chunk.  (Wave-uniform.)
                    if (sb + 2 * kD < nmax || (MULTI && nmax > 2 * kD && c + KC < kall))
                        fill(h, sb + 2 * kD < nmax ? sb + 2 * kD : h * kD);
                }
            }
        }
        // Y += A X, or Y = A X with Y not read (then 0 + the sum, the bits a
        // zeroed Y would give).
        if (br < brows) {
            const int r0 = br * 4;
            float* const Y = Yall + (size_t)c * ldy;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int xc = t * NQ + q;
                if (xc >= k) continue;
                float* yp = Y + (size_t)xc * ldy + r0;
                if (r0 + 3 < rows) {
                    f32x4* yq = reinterpret_cast<f32x4*>(yp);
                    f32x4 y = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                    if (acc_y) y = *yq;
                    y += acc[q];
                    *yq = y;
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (r0 + i < rows) yp[i] = (acc_y ? yp[i] : 0.0f) + acc[q][i];
                }
            }
        }
    }
}

template <bool BF16, int NQ, bool LONG, bool TILED, bool X16>
void launch_kernel(hipStream_t s, unsigned blocks, bool nt, int brows, int rows, int cols, int k, const int* rp,
                   const int* bcol, const unsigned short* val, const unsigned short* X, int ldx, float* Y, int ldy,
                   bool acc_y) {
    if (TILED && nt)
        hipLaunchKernelGGL((bspmm4_kernel<BF16, NQ, LONG, TILED, X16, TILED>), dim3(blocks), dim3(kBlock), 0, s,
                           brows, rows, cols, k, rp, bcol, val, X, ldx, Y, ldy, acc_y);
    else
        hipLaunchKernelGGL((bspmm4_kernel<BF16, NQ, LONG, TILED, X16, false>), dim3(blocks), dim3(kBlock), 0, s,
                           brows, rows, cols, k, rp, bcol, val, X, ldx, Y, ldy, acc_y);
}

// The window kernel, when the matrix has X windows, its rows need no FP64
// carries, a work-group's windows fit kSpmmWinLds and X's byte offsets fit
// 32 bits (the kernel's staging loads take 32-bit offsets from X); sets
// `done` when it launched.  One wave per slice for all k right-hand sides.
template <bool BF16, int NQ, bool COLX>
hipError_t launch_win(hipStream_t s, bool long_rows, int brows, int rows, int cols, int k, const int* rp,
                      const unsigned short* val, const unsigned short* X, int ldx, float* Y, int ldy, bool acc_y,
                      const launch::SpmmWindows& win, bool& done) {
    done = false;
    if (long_rows || !win.tile || !win.idx || win.max <= 0 || win.max > launch::kSpmmWinMax) return hipSuccess;
    const long long x_bytes = COLX ? (long long)k * ldx * 2 : (long long)((cols + 3) / 4) * ldx * 8;
    const int win_stride = (win.max + 2) * kWinTile<NQ> + (win.max * 4 + 15) / 16 * 16;
    if (x_bytes > 0xFFFFFFFFll || (long long)(kWinBlock / kWave) * win_stride > launch::kSpmmWinLds) return hipSuccess;
    const long long slices = (brows + launch::kSpmmSlice - 1) / launch::kSpmmSlice;
    const unsigned wb = (unsigned)((slices * kWave + kWinBlock - 1) / kWinBlock);
    const size_t shm = (size_t)(kWinBlock / kWave) * win_stride;
    if (k > 4 * NQ)
        hipLaunchKernelGGL((bspmm4_win_kernel<BF16, NQ, true, COLX>), dim3(wb), dim3(kWinBlock), shm, s, brows, rows,
                           cols, k, rp, val, win, win_stride, X, ldx, Y, ldy, acc_y);
    else
        hipLaunchKernelGGL((bspmm4_win_kernel<BF16, NQ, false, COLX>), dim3(wb), dim3(kWinBlock), shm, s, brows,
                           rows, cols, k, rp, val, win, win_stride, X, ldx, Y, ldy, acc_y);
    done = true;
    return hipGetLastError();
}

template <bool BF16, int NQ, bool TILED>
hipError_t launch_nq(hipStream_t s, bool long_rows, int brows, int rows, int cols, int k, const int* rp,
                     const int* bcol, const unsigned short* val, const unsigned short* X, int ldx,
                     float* Y, int ldy, bool acc_y, const launch::SpmmWindows& win) {
    const long long slices = (brows + launch::kSpmmSlice - 1) / launch::kSpmmSlice;
    const long long chunks = (TILED && k > 4 * NQ) ? (k + 4 * NQ - 1) / (4 * NQ) : 1;  // waves per slice
    const long long blocks = (slices * chunks * kWave + kBlock - 1) / kBlock;
    // 16-byte tile loads: every tile pair of a chunk 16-byte aligned (an even
    // stride and chunk size -- chunks are 16 right-hand sides but the last,
    // k - 16m, even with k -- and chunks start at multiples of 16) and two
    // of them to load.
    constexpr bool kX16 = TILED && NQ >= 2;
    const bool x16 = kX16 && ldx % 2 == 0 && k % 2 == 0 && ((uintptr_t)X & 15) == 0;
    // One chunk (k <= 4 * NQ): non-temporal A loads.  Measured on the
    // 524,288 x 12 4x4-block matrix: k = 8 / 16 in 76.6 / 122 us against
    // 78.5 / 130 us with normal loads (Y += A X).  Several chunks: a wave per
    // chunk, normal A loads, non-temporal Y accesses (which cost 3-5 us at
    // k = 8 / 16).
    const bool nt = k <= 4 * NQ;
    // X windows in LDS when the matrix has them (16-byte tile pieces).
    if constexpr (kX16) {
        if (x16) {
            bool done = false;
            const hipError_t e = launch_win<BF16, NQ, false>(s, long_rows, brows, rows, cols, k, rp, val, X, ldx, Y,
                                                             ldy, acc_y, win, done);
            if (done || e != hipSuccess) return e;
        }
    }
    const unsigned nb = (unsigned)blocks;
    if (long_rows) {
        if (x16)
            launch_kernel<BF16, NQ, true, TILED, kX16>(s, nb, nt, brows, rows, cols, k, rp, bcol, val, X, ldx, Y, ldy,
                                                        acc_y);
        else
            launch_kernel<BF16, NQ, true, TILED, false>(s, nb, nt, brows, rows, cols, k, rp, bcol, val, X, ldx, Y, ldy,
                                                        acc_y);
    } else {
        if (x16)
            launch_kernel<BF16, NQ, false, TILED, kX16>(s, nb, nt, brows, rows, cols, k, rp, bcol, val, X, ldx, Y, ldy,
                                                        acc_y);
        else
            launch_kernel<BF16, NQ, false, TILED, false>(s, nb, nt, brows, rows, cols, k, rp, bcol, val, X, ldx, Y,
                                                         ldy, acc_y);
    }
    return hipGetLastError();
}

template <bool BF16, bool TILED>
hipError_t launch_k(hipStream_t s, bool long_rows, int brows, int rows, int cols, int k, const int* rp,
                    const int* bcol, const unsigned short* val, const unsigned short* X, int ldx,
                    float* Y, int ldy, bool acc_y, const launch::SpmmWindows& win) {
    // Up to 32 right-hand sides per launch (8 quads of accumulators).  With
    // tiled X one launch takes every right-hand side, 16 per wave (one pass
    // of 8 quads, 86 registers, ran at 5 waves per SIMD and 294 against 247
    // us at k = 32).
    constexpr int kChunk = TILED ? 1 << 30 : 32;
    if constexpr (!TILED) {
        // X windows staged from the columns of X: one launch for every k,
        // 8 or 16 right-hand sides per chunk.
        bool done = false;
        const hipError_t e =
            k <= 8 ? launch_win<BF16, 2, true>(s, long_rows, brows, rows, cols, k, rp, val, X, ldx, Y, ldy, acc_y, win,
                                               done)
                   : launch_win<BF16, 4, true>(s, long_rows, brows, rows, cols, k, rp, val, X, ldx, Y, ldy, acc_y, win,
                                               done);
        if (done || e != hipSuccess) return e;
    }
    for (int c = 0; c < k; c += kChunk) {
        const int kc = std::min(kChunk, k - c);
        // Right-hand side c: column c, or offset c inside every tile.
        const unsigned short* Xc = X + (TILED ? (size_t)c * 4 : (size_t)c * ldx);
        float* Yc = Y + (size_t)c * ldy;
        const int nq = (std::min(kc, TILED ? 16 : 32) + 3) / 4;
        hipError_t e;
        if (nq <= 1)
            e = launch_nq<BF16, 1, TILED>(s, long_rows, brows, rows, cols, kc, rp, bcol, val, Xc, ldx, Yc, ldy, acc_y, win);
        else if (nq <= 2)
            e = launch_nq<BF16, 2, TILED>(s, long_rows, brows, rows, cols, kc, rp, bcol, val, Xc, ldx, Yc, ldy, acc_y, win);
        else if (nq <= 4)
            e = launch_nq<BF16, 4, TILED>(s, long_rows, brows, rows, cols, kc, rp, bcol, val, Xc, ldx, Yc, ldy, acc_y, win);
        else if constexpr (!TILED)
            e = launch_nq<BF16, 8, TILED>(s, long_rows, brows, rows, cols, kc, rp, bcol, val, Xc, ldx, Yc, ldy, acc_y, win);
        else
            e = hipErrorInvalidValue;  // not reached: tiled chunks hold at most 4 quads
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// X (column-major, leading dimension ldx) into 4-row tiles, zeros past the
// last column.  A work-group takes kPackTiles tiles and kPackRhs right-hand
// sides at a time through LDS: it reads each right-hand side's rows of those
// tiles as one contiguous run (kPackTiles * 8 bytes) and writes the tiles'
// right-hand sides as runs of kPackRhs * 8 bytes.  (One thread per tile and
// right-hand side read a different column in every lane: 60 us for 64 MB.)
constexpr int kPackTiles = 64, kPackRhs = 32;
__global__ __launch_bounds__(kBlock) void pack_x4_kernel(int k, int cols, int tiles, const unsigned short* __restrict__ X,
                                                         int ldx, unsigned short* __restrict__ X4) {
    __shared__ s16x4 buf[kPackTiles * kPackRhs];
    constexpr int kPer = kPackTiles * kPackRhs / kBlock;
    const int tile0 = blockIdx.x * kPackTiles;
    const int nt = min(kPackTiles, tiles - tile0);
    for (int j0 = 0; j0 < k; j0 += kPackRhs) {
        const int nj = min(kPackRhs, k - j0);
        __syncthreads();  // the previous group's writes have read buf
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int p = u * kBlock + threadIdx.x, i = p % kPackTiles, j = p / kPackTiles;
            if (i < nt && j < nj) {
                const int c = (tile0 + i) * 4;
                buf[i * kPackRhs + j] = ld4(X + (size_t)(j0 + j) * ldx + c, cols - c);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int p = u * kBlock + threadIdx.x, i = p / kPackRhs, j = p % kPackRhs;
            if (i < nt && j < nj)
                *reinterpret_cast<s16x4*>(X4 + ((size_t)(tile0 + i) * k + j0 + j) * 4) = buf[i * kPackRhs + j];
        }
    }
}

}  // namespace

namespace launch {

hipError_t bspmm4(hipStream_t s, bool bf16, bool long_rows, int brows, int rows, int cols, int k,
                  const int* rp, const int* bcol, const unsigned short* val,
                  const unsigned short* X, int ldx, float* Y, int ldy, bool tiled, bool accumulate,
                  SpmmWindows win) {
    if (brows <= 0 || k <= 0) return hipSuccess;
    if ((long long)brows * 4 / kBlock + 1 > 0x7FFFFFFFll) return hipErrorInvalidValue;
    if (tiled)
        return bf16 ? launch_k<true, true>(s, long_rows, brows, rows, cols, k, rp, bcol, val, X, ldx, Y, ldy,
                                           accumulate, win)
                    : launch_k<false, true>(s, long_rows, brows, rows, cols, k, rp, bcol, val, X, ldx, Y, ldy,
                                            accumulate, win);
    return bf16 ? launch_k<true, false>(s, long_rows, brows, rows, cols, k, rp, bcol, val, X, ldx, Y, ldy, accumulate,
                                        win)
                : launch_k<false, false>(s, long_rows, brows, rows, cols, k, rp, bcol, val, X, ldx, Y, ldy, accumulate,
                                         win);
}

hipError_t pack_x4(hipStream_t s, int k, int cols, const unsigned short* X, int ldx, unsigned short* X4) {
    if (k <= 0 || cols <= 0) return hipSuccess;
    const int tiles = (cols + 3) / 4;
    const int grid = (tiles + kPackTiles - 1) / kPackTiles;
    hipLaunchKernelGGL(pack_x4_kernel, dim3(grid), dim3(kBlock), 0, s, k, cols, tiles, X, ldx, X4);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace clspmv
