// x-panel SpMV for gfx950: the COO/CSR path for matrices whose columns have
// no locality (power-law rows, random graphs).
//
// Gathering x straight from memory costs a 128-byte L2 line per 4-byte
// element, which caps such kernels near 200 G gathers/s on MI355X whatever
// the HBM bandwidth.  Here the columns are split into panels of at most
// kPanelMax floats; a work-group copies its panel's x into LDS once with
// coalesced loads and then gathers from LDS.  Entries carry a 15-bit
// in-panel column and a segment-start bit in 16 bits, so the matrix stream
// is 6 bytes per nonzero instead of 12 (COO) or 8 (CSR).
//
// Pass 1: each wave takes sub-chunks of 512 entries (8 consecutive ones per
// lane), forms the products, sums each lane's entries in order and reduces
// every segment (one row's run inside the sub-chunk) with a segmented scan
// over the wave's lanes (DPP row shifts and broadcasts); each segment's sum
// goes to seg_val.  Eight entries per lane spread the scan's fixed cost
// over twice the entries of a 4-per-lane layout.  Pass 2: one work-group
// per row block loads the block's segments (one contiguous range), puts
// each at its slot of a row-major LDS layout, and each row adds its slots
// in panel order to y (long rows spread over the work-group).  The
// summation order is fixed by the encoding; no atomics.

#include "device_common.hpp"
#include "kernels.hpp"

namespace clspmv {
namespace {

using namespace dev;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kPer = launch::kPanelSub / kWave;  // entries per lane
static_assert(kPer == 8, "pass 1 is written for 8 entries per lane");
constexpr int kU = 2;  // sub-chunks in flight per wave

// DPP controls (GFX9 encoding): shifts within 16-lane rows, broadcasts of
// lane 15 / 31 to the following rows, and a whole-wave shift by one.
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

// The lane's value moved by DPP; lanes with no source (or in rows outside
// ROW_MASK) receive 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_float(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, true));
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_int(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xF, true);
}

// One step of a segmented inclusive scan: (s, g) absorbs the partner's
// (s, g) unless this lane already starts a segment.  A partner of 0 / 0
// changes nothing, so lanes without a source are left as they are.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void seg_step(float& s, int& g) {
    const float ts = dpp_float<CTRL, ROW_MASK>(s);
    const int tg = dpp_int<CTRL, ROW_MASK>(g);
    if (!g) s += ts;
    g |= tg;
}

// A lane's eight values as loaded (converted to FP32 only when used, so a
// load is never waited for before its sub-chunk's turn).
template <class V>
struct Vals8;
template <>
struct Vals8<ValF32> {
    f32x4 a, b;
    __device__ __forceinline__ void load(const float* __restrict__ p) {
        a = *reinterpret_cast<const f32x4*>(p);
        b = *reinterpret_cast<const f32x4*>(p + 4);
    }
    __device__ __forceinline__ float get(int k) const { return k < 4 ? a[k] : b[k - 4]; }
};
template <class Cvt>
struct Vals8<Val16<Cvt>> {
    u32x4 w;
    __device__ __forceinline__ void load(const unsigned short* __restrict__ p) {
        w = *reinterpret_cast<const u32x4*>(p);
    }
    __device__ __forceinline__ float get(int k) const {
        const unsigned v = w[k >> 1];
        return Cvt::f((unsigned short)(k & 1 ? v >> 16 : v & 0xFFFFu));
    }
};

template <class V>
struct SubChunk {
    u32x4 c;     // eight 16-bit entry words
    Vals8<V> v;  // eight values
};

// One lane's share of a sub-chunk's segment sums, held in registers until
// they are stored.  Its segments are consecutive: `before` - 1 (the one
// open from earlier lanes, closed at the lane's first flag with value
// `first`), then one per further flag (the one closed by flag k has the
// running sum r[k - 1]), and, in the last lane, the sub-chunk's last
// segment (value `last`).  Bit k of `flags` is entry k's segment-start bit.
struct SegSums {
    float first, r[kPer - 1], last;
    int before, flags;
};

// Reduces one sub-chunk (8 entries per lane) to its segment sums.  xs holds
// the panel's x and a 0 at index pad (the panel width): padding entries
// (column word 0x7FFF >= pad) gather it, so every lane reads x with no
// branch and the LDS reads are in flight together.
template <class V>
__device__ __forceinline__ void seg_scan(const SubChunk<V>& q, const float* __restrict__ xs, unsigned pad,
                                         int lane, SegSums& o) {
    unsigned w[kPer];
#pragma unroll
    for (int j = 0; j < kPer / 2; ++j) {
        w[2 * j] = q.c[j] & 0xFFFFu;
        w[2 * j + 1] = q.c[j] >> 16;
    }
    int f[kPer];
    float xv[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        f[k] = (int)(w[k] >> 15);
        xv[k] = xs[min(w[k] & 0x7FFFu, pad)];
    }
    if (lane == 0) f[0] = 1;  // every sub-chunk opens a segment
    // Running sums inside the lane, restarted at each flag (one rounding
    // per entry: the product is fused into the add).
    float r[kPer];
    r[0] = q.v.get(0) * xv[0];
#pragma unroll
    for (int k = 1; k < kPer; ++k) r[k] = fmaf(q.v.get(k), xv[k], f[k] ? 0.0f : r[k - 1]);
    int nf = 0, flags = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        nf += f[k];
        flags |= f[k] << k;
    }
    // Segment numbering: the flags of earlier lanes (an exclusive scan of
    // nf over the wave: DPP row shifts, then the row broadcasts), then this
    // lane's own.
    int incl = nf;
    incl += dpp_int<kRowShr1, 0xF>(incl);
    incl += dpp_int<kRowShr2, 0xF>(incl);
    incl += dpp_int<kRowShr4, 0xF>(incl);
    incl += dpp_int<kRowShr8, 0xF>(incl);
    incl += dpp_int<kRowBcast15, 0xA>(incl);
    incl += dpp_int<kRowBcast31, 0xC>(incl);
    // The lane's head: its entries before its first flag.
    float head = r[kPer - 1];
#pragma unroll
    for (int k = kPer - 1; k >= 1; --k) head = f[k] ? r[k - 1] : head;
    if (f[0]) head = 0.0f;
    // Segmented inclusive scan over lanes of the amount each lane passes on:
    // its tail from the last flag, or its whole sum when it has none -- the
    // last running sum either way.
    float s = r[kPer - 1];
    int g = nf ? 1 : 0;
    seg_step<kRowShr1, 0xF>(s, g);
    seg_step<kRowShr2, 0xF>(s, g);
    seg_step<kRowShr4, 0xF>(s, g);
    seg_step<kRowShr8, 0xF>(s, g);
    seg_step<kRowBcast15, 0xA>(s, g);
    seg_step<kRowBcast31, 0xC>(s, g);
    // The segment open before this lane ends at its first flag.
    o.first = dpp_float<kWaveShr1, 0xF>(s) + head;
#pragma unroll
    for (int k = 0; k < kPer - 1; ++k) o.r[k] = r[k];
    o.last = s;
    o.before = incl - nf;
    o.flags = flags;
}

// out[i] for a small non-negative i: a 32-bit byte offset from the
// wave-uniform base, so the store needs no 64-bit address arithmetic.
__device__ __forceinline__ void store_at(float* __restrict__ out, int i, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(out) + (unsigned)(i << 2)) = v;
}

// Stores one lane's segment sums (predicated, in-lane).  Indices are below
// kPanelSub + 1: a sub-chunk's segments start at out.
__device__ __forceinline__ void seg_store(const SegSums& o, int lane, float* __restrict__ out) {
    const int b = o.before;
    // Segments that open and close inside the lane: flag k closes the one
    // opened by the lane's previous flag, numbered by the flags before k.
#pragma unroll
    for (int k = 1; k < kPer; ++k) {
        const int below = o.flags & ((1 << k) - 1);
        if ((o.flags >> k) & 1 && below) store_at(out, b + __builtin_popcount(below) - 1, o.r[k - 1]);
    }
    const int nf = __builtin_popcount(o.flags);
    if (nf && lane > 0) store_at(out, b - 1, o.first);
    // The last lane closes the sub-chunk's last segment.
    if (lane == kWave - 1) store_at(out, b + nf - 1, o.last);
}

__device__ __forceinline__ float x_at(const float* __restrict__ x, int c, int cols) {
    return (unsigned)c < (unsigned)cols ? x[c] : 0.0f;
}

// The DIA rows of work-group b of the tail: four rows per thread (strided
// by the work-group), per row four partial sums over the diagonals (the
// remainder in the first) added in order -- dia_kernel method 1's
// arithmetic.  Loads of rows past the end are clamped; those rows are not
// written.
template <int NT>
__device__ __forceinline__ void dia_tail_rows(const launch::DiaTail& d, const float* __restrict__ x, int b) {
    constexpr int R = launch::kDiaTailRows / NT, U = 4;
    const int r0 = b * launch::kDiaTailRows + (int)threadIdx.x;
    int rr[R];
#pragma unroll
    for (int k = 0; k < R; ++k) rr[k] = min(r0 + k * NT, d.rows - 1);
    float a[R][U];
#pragma unroll
    for (int k = 0; k < R; ++k)
#pragma unroll
        for (int u = 0; u < U; ++u) a[k][u] = 0.0f;
    int j = 0;
    for (; j + U <= d.ndiag; j += U) {
        float dv[R][U], xv[R][U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int o = d.off[j + u];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                dv[k][u] = d.val[(size_t)(j + u) * d.len + rr[k]];
                xv[k][u] = x_at(x, rr[k] + o, d.cols);
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k)
#pragma unroll
            for (int u = 0; u < U; ++u) a[k][u] = fmaf(dv[k][u], xv[k][u], a[k][u]);
    }
    for (; j < d.ndiag; ++j) {
        const int o = d.off[j];
#pragma unroll
        for (int k = 0; k < R; ++k)
            a[k][0] = fmaf(d.val[(size_t)j * d.len + rr[k]], x_at(x, rr[k] + o, d.cols), a[k][0]);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int r = r0 + k * NT;
        if (r >= d.rows) continue;
        float s = 0.0f;
#pragma unroll
        for (int u = 0; u < U; ++u) s += a[k][u];
        d.y[r] += s;
    }
}

template <int NT, class V>
__global__ __launch_bounds__(NT) void panel_pass1_kernel(
    const int4* __restrict__ wg, int nwg, int pw, int cols,
    const unsigned short* __restrict__ pcol, const typename V::T* __restrict__ pval, int pad_at,
    const int* __restrict__ sc_seg, const int* __restrict__ sc_start,
    const float* __restrict__ x, bool xa, float* __restrict__ seg_val, launch::DiaTail tail) {
    static_assert(launch::kDiaTailRows % NT == 0, "tail rows must split evenly over the threads");
    if ((int)blockIdx.x >= nwg) {  // a work-group of the DIA tail (the whole group)
        dia_tail_rows<NT>(tail, x, (int)blockIdx.x - nwg);
        return;
    }
    extern __shared__ float xs[];
    // The work-group's sub-chunk table, read once (loop indices then come
    // from LDS instead of one dependent memory load per sub-chunk).
    __shared__ int s_start[launch::kPanelScPerWg + 1];
    __shared__ int s_seg[launch::kPanelScPerWg];
    // The table lists a panel's work-groups together; in XCD-aware order
    // they run on one XCD, whose L2 then fetches that panel of x once
    // (block i goes to XCD i mod 8, so the plain order would have every
    // XCD stage every panel).
    const int4 t = wg[xcd_tile(blockIdx.x, nwg)];
    const int panel = t.x, sc0 = t.y, sc1 = t.z;
    const int xb = panel * pw;
    const int w = min(pw, cols - xb);
    const int tid = threadIdx.x;
    if (tid <= sc1 - sc0) s_start[tid] = sc_start[sc0 + tid];
    if (tid < sc1 - sc0) s_seg[tid] = sc_seg[sc0 + tid];
    // Sub-chunk indices are wave-uniform: keep them in scalar registers.
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    const int lane = tid & (kWave - 1);
    constexpr int nw = NT / kWave;
    // Every lane loads on every path: lanes past a short sub-chunk's end,
    // and whole sub-chunks past the work-group's last, read the padding
    // block, which contributes nothing.  With no load under a branch the
    // count of memory operations in flight is the same on every path, so
    // each wait covers just the group it needs (a load that might have
    // been skipped makes the compiler wait for everything).
    auto issue = [&](SubChunk<V>& q, int i0, int i1) {  // the lane's entries from i0 if below i1
        const int i = i0 < i1 ? i0 : pad_at;
        // Plain (cached) loads: measured faster here than non-temporal
        // ones, which the x panel does not need.
        q.c = *reinterpret_cast<const u32x4*>(pcol + i);
        q.v.load(pval + i);
    };
    auto load = [&](SubChunk<V>* q, int sc) {  // sub-chunk bounds from the LDS table
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int c = sc + u * nw;
            int i0 = pad_at, i1 = 0;
            if (c < sc1) {
                i0 = s_start[c - sc0] + lane * kPer;
                i1 = s_start[c - sc0 + 1];
            }
            issue(q[u], i0, i1);
        }
    };
    // The first group's loads go out while the panel is staged, its bounds
    // read from the global table (the LDS copy is not ready before the
    // barrier).
    SubChunk<V> qa[kU], qb[kU];
    auto load_first = [&]() {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int c = sc0 + wave + u * nw;
            int i0 = pad_at, i1 = 0;
            if (c < sc1) {
                i0 = sc_start[c] + lane * kPer;
                i1 = sc_start[c + 1];
            }
            issue(qa[u], i0, i1);
        }
    };
    // Stage the panel: every load is issued before the first LDS store, so
    // the work-group waits for one memory latency, not one per step.
    constexpr int kStage = (launch::kPanelMax + 4 * NT - 1) / (4 * NT);
    const int w4 = w & ~3;  // elements staged with 128-bit loads
    if (xa && w4 >= 4) {
        // Unconditional loads (indices clamped into the panel) keep the
        // loop free of branches, so no wait separates them.
        f32x4 v[kStage];
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int i = min((tid + k * NT) * 4, w4 - 4);
            v[k] = *reinterpret_cast<const f32x4*>(x + xb + i);
        }
        load_first();
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int i = (tid + k * NT) * 4;
            if (i < w4) *reinterpret_cast<f32x4*>(xs + i) = v[k];
        }
        for (int i = w4 + tid; i < w; i += NT) xs[i] = x[xb + i];
        if (tid == 0) xs[pw] = 0.0f;  // the padding entries' x
    } else {
        float v[4 * kStage];
#pragma unroll
        for (int k = 0; k < 4 * kStage; ++k) {
            const int i = min(tid + k * NT, w - 1);
            v[k] = x[xb + i];
        }
        load_first();
#pragma unroll
        for (int k = 0; k < 4 * kStage; ++k) {
            const int i = tid + k * NT;
            if (i < w) xs[i] = v[k];
        }
        if (tid == 0) xs[pw] = 0.0f;  // the padding entries' x
    }
    __syncthreads();

    // The sums of one group of sub-chunks are stored when the next group
    // starts, before its successor's loads are issued.  Memory operations
    // retire in issue order and count together, so the wait for a group's
    // loads then covers only stores issued earlier, which have had a whole
    // iteration to complete; stores issued after the loads would add their
    // latency to every wait.
    SegSums res[kU];
    auto store = [&](int base) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int c = base + u * nw;
            // The segment base is wave-uniform: a scalar base lets every
            // store take a 32-bit lane offset.
            if (c < sc1) seg_store(res[u], lane, seg_val + __builtin_amdgcn_readfirstlane(s_seg[c - sc0]));
        }
    };
    // Two register groups take turns (no copies between them): one is
    // reduced while the other's loads are in flight.  The last group
    // prefetches nothing, so no wave ends waiting for loads of padding.
    constexpr int step = nw * kU;
    int sc = sc0 + wave;
    if (sc >= sc1) return;
    bool pending = false;  // res holds the previous group's sums
    // Reduces the group at sc (in cur), prefetching the next one into nxt;
    // true once the last group is done.
    auto group = [&](SubChunk<V>* cur, SubChunk<V>* nxt) -> bool {
        if (pending) store(sc - step);
        if (sc + step >= sc1) {
#pragma unroll
            for (int u = 0; u < kU; ++u)
                if (sc + u * nw < sc1) seg_scan(cur[u], xs, (unsigned)pw, lane, res[u]);
            store(sc);
            return true;
        }
        load(nxt, sc + step);
#pragma unroll
        for (int u = 0; u < kU; ++u) seg_scan(cur[u], xs, (unsigned)pw, lane, res[u]);
        pending = true;
        sc += step;
        return false;
    };
    while (!group(qa, qb) && !group(qb, qa)) {
    }
}

__global__ __launch_bounds__(launch::kBlockThreads) void panel_pass2_kernel(
    const int* __restrict__ blk_row, const int* __restrict__ blk_seg,
    const int* __restrict__ row_slot, const unsigned short* __restrict__ seg_dst,
    const float* __restrict__ seg_val, const int* __restrict__ blk_long,
    const int4* __restrict__ longs, float* __restrict__ y) {
    constexpr int NT = launch::kBlockThreads, K = launch::kBlockItems;
    __shared__ float slot[launch::kBlockSegsMax];
    constexpr int R = launch::kBlockRowsMax / NT;  // rows (and long rows) per thread at most
    static_assert(R * NT == launch::kBlockRowsMax, "rows of a block must split evenly over the threads");
    const int b = blockIdx.x, tid = threadIdx.x;
    const int row0 = blk_row[b], nrows = blk_row[b + 1] - row0;
    const int j0 = blk_seg[b], j1 = blk_seg[b + 1];
    const int l0 = blk_long[b], nlong = blk_long[b + 1] - l0;
    // Every load of the block -- its segments, the slot range and y of each
    // of this thread's rows, its long rows -- is issued before the first is
    // used, so the work-group waits for about one memory latency.
    unsigned short d[K];
    float v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = j0 + tid + k * NT;
        d[k] = j < j1 ? seg_dst[j] : 0;
        v[k] = j < j1 ? seg_val[j] : 0.0f;
    }
    const int* rs = row_slot + row0 + b;
    int s0[R], s1[R];
    float yv[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * NT;
        const bool live = i < nrows;
        s0[k] = live ? rs[i] : 0;
        s1[k] = live ? rs[i + 1] : 0;
        yv[k] = live ? y[row0 + i] : 0.0f;
    }
    // The first NT long rows (a block rarely has more) are prefetched too.
    int4 lr = tid < nlong ? longs[l0 + tid] : int4{0, 0, 0, 0};
    float ly = tid < nlong ? y[row0 + lr.x] : 0.0f;
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (j0 + tid + k * NT < j1) slot[d[k]] = v[k];
    __syncthreads();
    // Rows of at most kLongSlots slots: each thread adds its rows' slots in
    // order (panel order; a run cut at a sub-chunk boundary keeps its pieces
    // in order), in FP32.
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = tid + k * NT;
        const int s = s0[k], e = s1[k];
        if (i >= nrows || s == e || e - s > launch::kLongSlots) continue;
        float acc = slot[s];
        for (int q = s + 1; q < e; ++q) acc += slot[q];
        y[row0 + i] = yv[k] + acc;
    }
    // Longer rows -- every row longer than kLongRowF64 among them -- spread
    // one per thread over the whole work-group (so the few long rows of a
    // block do not hold up every wave), summed in FP64 in slot order and
    // rounded into y once.
    for (int base = 0; base < nlong; base += NT) {  // uniform
        if (base > 0) {
            const int l = base + tid;
            lr = l < nlong ? longs[l0 + l] : int4{0, 0, 0, 0};
            ly = l < nlong ? y[row0 + lr.x] : 0.0f;
        }
        if (base + tid < nlong) {
            double acc = slot[lr.y];
            for (int q = lr.y + 1; q < lr.z; ++q) acc += slot[q];
            y[row0 + lr.x] = (float)((double)ly + acc);
        }
    }
}

}  // namespace

namespace launch {

hipError_t panel_prepare() {
    // Pass 1 holds up to kPanelMax floats of x in dynamic LDS, beyond the
    // default limit: raise it once per device, outside any graph capture.
    const int lds = (int)(sizeof(float) * (kPanelMax + 1));
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&panel_pass1_kernel<kPanelThreads, ValF32>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&panel_pass1_kernel<kPanelThreads, ValBF16>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&panel_pass1_kernel<kPanelThreads, ValF16>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    return e;
}

hipError_t panel_pass1(hipStream_t s, int nwg, const int* wg, int pw, int cols, const unsigned short* pcol,
                       int dtype, const void* pval, int pad_at, const int* sc_seg, const int* sc_start,
                       const float* x, bool x_aligned, float* seg_val, const DiaTail* tail) {
    if (nwg <= 0) return hipSuccess;
    if (pw <= 0 || pw > kPanelMax) return hipErrorInvalidValue;
    DiaTail dt{0, 0, 0, 0, nullptr, nullptr, nullptr};
    int ntail = 0;
    if (tail && tail->rows > 0 && tail->ndiag > 0) {
        dt = *tail;
        ntail = (dt.rows + kDiaTailRows - 1) / kDiaTailRows;
    }
    const int grid = nwg + ntail;
    const size_t lds = sizeof(float) * ((size_t)pw + 1);  // x panel + the padding entries' 0
    const int4* wg4 = reinterpret_cast<const int4*>(wg);
    const unsigned short* p16 = static_cast<const unsigned short*>(pval);
    switch (dtype) {
        case 0:
            hipLaunchKernelGGL((panel_pass1_kernel<kPanelThreads, ValF32>), dim3(grid), dim3(kPanelThreads), lds,
                               s, wg4, nwg, pw, cols, pcol, static_cast<const float*>(pval), pad_at, sc_seg,
                               sc_start, x, x_aligned, seg_val, dt);
            break;
        case 1:
            hipLaunchKernelGGL((panel_pass1_kernel<kPanelThreads, ValBF16>), dim3(grid), dim3(kPanelThreads),
                               lds, s, wg4, nwg, pw, cols, pcol, p16, pad_at, sc_seg, sc_start, x, x_aligned,
                               seg_val, dt);
            break;
        case 2:
            hipLaunchKernelGGL((panel_pass1_kernel<kPanelThreads, ValF16>), dim3(grid), dim3(kPanelThreads),
                               lds, s, wg4, nwg, pw, cols, pcol, p16, pad_at, sc_seg, sc_start, x, x_aligned,
                               seg_val, dt);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t panel_pass2(hipStream_t s, const float* seg_val, int nblocks, const int* blk_row, const int* blk_seg,
                       const int* row_slot, const unsigned short* seg_dst, const int* blk_long,
                       const int* longs, float* y) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(panel_pass2_kernel, dim3(nblocks), dim3(kBlockThreads), 0, s, blk_row, blk_seg,
                       row_slot, seg_dst, seg_val, blk_long, reinterpret_cast<const int4*>(longs), y);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace clspmv
