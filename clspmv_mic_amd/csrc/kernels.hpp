// Host-callable launchers of the gfx950 SpMV kernels.  Every launcher
// enqueues y += A*x on `s` and returns the launch status; shapes are checked
// by the caller (runtime.cpp) before any launch.
#pragma once

#include <hip/hip_runtime.h>

namespace clspmv {
namespace launch {

// ---- CSR ------------------------------------------------------------------
// Vector kernel: `lanes` (power of two, 1..64) lanes cooperate on one row.
hipError_t csr_vector(hipStream_t s, int rows, int lanes, const int* rp,
                      const int* col, const float* val, const float* x, float* y);
// Adaptive kernel over host-built row blocks (`blocks`, nblocks+1 row starts).
// A block is either a run of rows whose entries fit kAdaptiveNnz, or one
// longer row.
constexpr int kAdaptiveNnz = 1024;
constexpr int kAdaptiveRows = 256;
// Rows longer than this are accumulated in FP64 by the 16-bit-value and
// multi-vector CSR paths, the panel row sums and the block SpMM.
constexpr int kLongRowF64 = 4096;
hipError_t csr_adaptive(hipStream_t s, int nblocks, const int* blocks,
                        const int* rp, const int* col, const float* val,
                        const float* x, float* y);

// ---- COO ------------------------------------------------------------------
// Row-sorted COO in chunks of kCooChunk entries; 2 carry slots per chunk.
constexpr int kCooChunk = 1024;
hipError_t coo_segreduce(hipStream_t s, int nnz, const int* row, const int* col,
                         const float* val, const float* x, float* y,
                         int* carry_row, float* carry_val);

// ---- x-panel encoding (COO and CSR) -------------------------------------------
// Columns are cut into panels of at most kPanelMax floats (a panel's x fits
// in LDS) and rows into blocks.  Entries are grouped by (panel, row block),
// row-sorted inside a group; a group is padded to a multiple of 8 entries
// (one pass-1 lane's share) and cut into sub-chunks of kPanelSub entries
// (the last one shorter;
// sc_start holds each sub-chunk's first entry).  An entry is a 16-bit word
// (bit 15: first entry of a segment; bits 0-14: column inside the panel,
// kPanelPad for padding) and a float.  A segment is one row's run inside
// one sub-chunk; seg_row holds its row, bit 31 when it continues the
// previous segment's row and bit 30 when the next segment continues it.
// Pass 1 (one work-group per run of sub-chunks of one panel, `wg` holds
// (panel, first sub-chunk, end sub-chunk, 0) per group) writes each
// segment's sum to seg_val[sc_seg[sub-chunk] + k].  Segments are numbered
// in (row block, panel, row) order, so pass 2 (one work-group per row block)
// reads its block's segments as one contiguous range, places each in LDS at
// its precomputed slot (seg_dst) of the block's row-major layout (row_slot:
// first slot per row), and adds each row's slots, in panel order, to y.
constexpr int kPanelSub = 512;        // entries per wave-sized sub-chunk (8 per lane)
constexpr int kPanelMax = 32760;      // largest panel width (floats)
constexpr unsigned kPanelPad = 0x7FFF;  // column code of padding entries
constexpr int kPanelThreads = 1024;   // pass-1 work-group size
constexpr int kPanelScPerWg = 512;    // most sub-chunks of one pass-1 work-group
constexpr int kPanelsMax = 2048;      // most panels of one matrix (bounds the encoder's tables)
constexpr int kBlockRowsMax = 4096;   // rows of a pass-2 row block
constexpr int kBlockThreads = 1024;   // pass-2 work-group size
constexpr int kBlockItems = 16;       // segments per pass-2 thread
constexpr int kBlockSegsMax = kBlockThreads * kBlockItems;  // < 65536 (16-bit slots)
// Raises pass 1's dynamic LDS limit on the current device (call before the
// first launch, outside graph capture).
hipError_t panel_prepare();
// Pass 2 spreads the rows with more than kLongSlots slots (blk_long / longs
// list them per block) over its threads and sums them in FP64; a slot holds
// at most one sub-chunk, so every row longer than kLongRowF64 is among them.
constexpr int kLongSlots = 8;
static_assert(kLongSlots <= kLongRowF64 / kPanelSub, "every row longer than kLongRowF64 must be a long row");
// Pass 1 reads x and writes the segment sums seg_val; pass 2 reads seg_val
// and adds the row sums into y.  Pass 1 touches nothing pass 2 of another
// matrix or any other kernel writes, so it may run on a second stream while
// earlier kernels update y, as long as pass 2 waits for it.
// The rows of an FP32 DIA part (y[r] += sum_d val[d * len + r] *
// x[r + off[d]], r < rows) can ride along in a pass-1 launch: extra
// work-groups after the panel ones take kDiaTailRows rows each, with
// dia_kernel method 1's arithmetic per row (its result, bit for bit).
// Pass 1 does not touch y, so the DIA update still lands before pass 2's.
constexpr int kDiaTailRows = 4 * kPanelThreads;
struct DiaTail {
    int rows, cols, len, ndiag;
    const int* off;
    const float* val;
    float* y;
};
// pad_at: entry index of 8 padding entries (column code kPanelPad, value 0);
// tail: a DIA part to sum in the same launch, or null.
hipError_t panel_pass1(hipStream_t s, int nwg, const int* wg, int pw, int cols, const unsigned short* pcol,
                       int dtype, const void* pval, int pad_at, const int* sc_seg, const int* sc_start,
                       const float* x, bool x_aligned, float* seg_val, const DiaTail* tail = nullptr);
hipError_t panel_pass2(hipStream_t s, const float* seg_val, int nblocks, const int* blk_row, const int* blk_seg,
                       const int* row_slot, const unsigned short* seg_dst, const int* blk_long,
                       const int* longs, float* y);

// ---- ELL / SELL --------------------------------------------------------------
hipError_t ell(hipStream_t s, int rows, int hal, int ellnum, int rows_per_lane,
               const int* col, const float* val, const float* x, float* y);
hipError_t sell(hipStream_t s, int rows, int slice_height, const int* ptr,
                const int* col, const float* val, const float* x, float* y);

// ---- DIA / BDIA -------------------------------------------------------------
// rows = number of rows computed (min(height, dia_length_aligned)).
// Methods: 0 one row per lane, 1 two rows per lane (len % 2 == 0),
// 2 / 3 one / four rows per lane with an LDS x window.
// dtype: 0 FP32 values, 1 BF16, 2 FP16 (data holds 16-bit bits).
hipError_t dia(hipStream_t s, int rows, int cols, int len, int ndiag,
               const int* offsets, int min_off, int max_off, int method, int dtype,
               const void* data, const float* x, float* y);
// Largest x window (floats) a work-group stages in LDS.
constexpr int kLdsWindow = 8192;
// LDS-window methods 1, 2, 4 only; methods 0 and 3 are run by dia() over
// the expanded diagonal offsets.
hipError_t bdia(hipStream_t s, int rows, int cols, int len, int nband,
                const int* band_off, const int* bptr, int max_band_width,
                int method, const float* data, const float* x, float* y);

// ---- block formats (bh in {1,2,4,8}, bw in {4,8}) ----------------------------
hipError_t bcsr(hipStream_t s, int brows, int rows, int cols, int bh, int bw,
                int lanes, int aligned, const int* rp, const int* col4,
                const float* data, const float* x, bool x_aligned, float* y);
// BCSR in the slot-major encoding: block rows of at most kBcsrSlotsMax
// slots (the sums are FP32 chains of at most 64 * 8 products per row).
constexpr int kBcsrSlotsMax = 64;
hipError_t bcsr_slots(hipStream_t s, int brows, int rows, int cols, int bh, int bw,
                      int hal, int nslots, const int* col4, const float* data,
                      const float* x, bool x_aligned, float* y);
hipError_t bell(hipStream_t s, int brows, int rows, int cols, int bh, int bw,
                int method, int hal, int f4, int ellnum, const int* col4,
                const float* data, const float* x, bool x_aligned, float* y);
hipError_t sbell(hipStream_t s, int brows, int rows, int cols, int bh, int bw,
                 int slice_height, const int* sptr, const int* col4,
                 const float* data, const float* x, bool x_aligned, float* y);

// ---- multi-right-hand-side SpMV (k <= kMultiMax) --------------------------------
// Y[:, q] += A X[:, q]; X row-major (X[c*k + q]), Y column-major (ldy).
constexpr int kMultiMax = 16;
hipError_t dia_multi(hipStream_t s, int rows, int cols, int len, int ndiag, const int* offsets,
                     int dtype, const void* data, int k, const float* X, float* Y, int ldy);
// CSR over the adaptive row blocks (csr_adaptive's `blocks`).  kb > 0 runs
// the kernel built for kb right-hand sides (k <= kb; 1, 2, 4, 8 or 16)
// whatever k is, so that each right-hand side's summation order stays that
// of kb; kb = 0 picks the smallest width that holds k.
hipError_t csr_multi(hipStream_t s, int nblocks, const int* blocks, const int* rp, const int* col,
                     int dtype, const void* val, int k, const float* X, float* Y, int ldy, int kb = 0);

// ---- mixed-precision 4x4-block SpMM (MFMA) -------------------------------------
// Y[:, 0:k] += A X[:, 0:k]; A in slices of kSpmmSlice block rows: sptr[s] is
// slice s's first slot (its width is (sptr[s+1] - sptr[s]) / kSpmmSlice),
// slot sptr[s] + j * kSpmmSlice + g holds block j of row s * kSpmmSlice + g:
// its element column in scol (a multiple of 4, -1 for padding) and its 16
// BF16 / FP16 values, row-major, in val.  X: 16-bit, column-major (ldx),
// or with `tiled` in 4-row tiles (value r of rows 4c..4c+3 of right-hand
// side j at X[(c * ldx + j) * 4 + r], ldx then the tiles' right-hand-side
// count, zeros past the last column); Y: FP32 column-major (ldy).  ldx, ldy
// multiples of 4 and X, Y 16-byte aligned.  long_rows: some block row holds
// more than kLongRowF64 / 4 blocks; the MFMA sums are then carried in FP64
// every kSpmmChunk blocks.
// With `win` (tiled X only): each slice's distinct block columns as 4-row
// tile indices, ascending, at tile[s * max ..] (max per slice, max <=
// kSpmmWinMax; a slice with fewer repeats its last one), and each slot's
// index into its slice's list in idx (max + 1 for padding: a zero tile).  A wave then
// copies its slice's tiles of X into LDS once and reads them from there;
// used when the windows fit kSpmmWinLds bytes of LDS per work-group.
constexpr int kSpmmChunk = 16;
constexpr int kSpmmSlice = 16;  // block rows per slice: one wave, 4 lanes per block row
constexpr int kSpmmWinMax = 254;
constexpr int kSpmmWinLds = 64 * 1024;
struct SpmmWindows {
    const int* tile = nullptr;
    const unsigned char* idx = nullptr;
    int max = 0;
};
hipError_t bspmm4(hipStream_t s, bool bf16, bool long_rows, int brows, int rows, int cols, int k,
                  const int* sptr, const int* scol, const unsigned short* val,
                  const unsigned short* X, int ldx, float* Y, int ldy, bool tiled = false,
                  bool accumulate = true, SpmmWindows win = {});
// Column-major X (k columns of `cols` values, leading dimension ldx) into
// the tiled layout above (ceil(cols / 4) * k tiles of 4 values).
hipError_t pack_x4(hipStream_t s, int k, int cols, const unsigned short* X, int ldx, unsigned short* X4);

// ---- bandwidth probe ---------------------------------------------------------
// dst[0..4*n4) = src[0..4*n4) with 128-bit loads/stores (16-byte aligned).
hipError_t copy4(hipStream_t s, long long n4, const float* src, float* dst);
// The same copy (unroll = 2, 4 or 8 float4 per thread per step, `grid`
// work-groups of 256), and a read-only stream writing one float per thread
// to out (grid * 256 floats).
hipError_t copy4u(hipStream_t s, long long n4, const float* src, float* dst, int unroll, int grid);
hipError_t read4u(hipStream_t s, long long n4, const float* src, float* out, int unroll, int grid);
// y[i] = x[hash(i) % nx] for i < n: random gathers from an nx-float vector.
hipError_t gather(hipStream_t s, long long n, int nx, const float* x, float* y);

}  // namespace launch
}  // namespace clspmv
