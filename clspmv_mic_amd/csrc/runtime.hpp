// Internal definitions of the device runtime shared by runtime.cpp (contexts,
// per-format matrices, drivers) and cocktail.cpp (the Cocktail runtime).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "clspmv_device.h"
#include "common.hpp"

struct clspmv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
};

struct clspmv_dmat {
    clspmv_dmat_info info{};
    int device = 0;
    std::vector<std::pair<void*, size_t>> bufs;  // device allocations owned by the matrix

    // CSR: rp/col/val; adaptive row blocks; vector lanes per row.
    int* rp = nullptr;
    int* col = nullptr;
    float* val = nullptr;
    int* blocks = nullptr;
    int nblocks = 0;
    int lanes = 1;
    int max_row = 0;
    // COO: row/col/val, carry slots (2 per chunk).
    int* row = nullptr;
    int* carry_row = nullptr;
    float* carry_val = nullptr;
    // ELL: col/val, height_aligned, ell_num.  SELL: sptr/col/val.
    int hal = 0, ellnum = 0;
    int* sptr = nullptr;
    // DIA / BDIA: offsets, band pointers, aligned length.
    int* offs = nullptr;
    int* bptr = nullptr;
    int* offs_exp = nullptr;  // BDIA: one offset per diagonal
    int len = 0, ndiag = 0, nband = 0, min_off = 0, max_off = 0, max_band = 0;
    // Block formats: block rows, aligned sizes, group lanes.
    int brows = 0, aligned = 0, f4 = 0;
    // BCSR in the slot-major encoding (bslots > 0, rp == nullptr): block
    // row r's k-th block column at col[k * hal + r] (-1 past the row's end),
    // its (w4, h) float4 at val[((k * bw4 + w4) * bh + h) * hal * 4 + r * 4].
    int bslots = 0;
    bool long_brows = false;  // BSPMM: a block row holds more than kLongRowF64 / 4 blocks
    // BSPMM X windows (wtile == nullptr: absent): slice s's distinct block
    // columns as 4-row tile indices, ascending, at wtile[s * win_max ..],
    // a shorter list repeating its last tile; each slot's index into its
    // slice's list in widx (win_max + 1 for padding); win_max the longest
    // list.
    int* wtile = nullptr;
    unsigned char* widx = nullptr;
    int win_max = 0;
    // x-panel encoding of COO / CSR (pw == 0: absent).  plain == false when
    // the (row, col, val) / CSR arrays were not uploaded.
    int pw = 0, pnwg = 0, pnp = 0, pnb = 0, pnsc = 0;
    int ppad = 0;  // entry index of the padding (one lane's share) after the encoding
    int* pwg = nullptr;  // int4 per pass-1 work-group
    unsigned short* pcol = nullptr;
    float* pval = nullptr;
    unsigned short* plpval = nullptr;  // values in BF16 / FP16 (dtype)
    int* psc_seg = nullptr;
    int* psc_start = nullptr;  // first entry of each sub-chunk (+ end)
    float* pseg = nullptr;
    unsigned short* pseg_dst = nullptr;  // slot of each segment in its block
    int* pblk_row = nullptr;   // first row of each row block (+ end)
    int* pblk_seg = nullptr;   // first segment of each row block (+ end)
    int* prow_slot = nullptr;  // per block: first slot of each row (+ end)
    int* pblk_long = nullptr;  // per block: first of its long rows in plong (+ end)
    int* plong = nullptr;      // long rows: {row in block, first slot, end slot, 0}
    bool use_panel = false;
    // Serving: 1 when batched passes run the multi-vector kernels (they
    // measured faster than separate SpMVs when the matrix was first
    // registered), 0 when not, -1 not measured yet.  Kept with the matrix so
    // that every server runs the same kernels; cleared when the encoding
    // changes.  Written only while a server holds the matrix's claim.
    mutable int serve_multi = -1;
    bool plain = true;
    // Values stored in 16 bits (dtype CLSPMV_DT_BF16 / CLSPMV_DT_F16): DIA
    // data, and BSPMM blocks (16 values per 4x4 block, row-major; rp / col
    // hold the block rows and element columns).  The panel encoding keeps
    // its own 16-bit values in plpval.
    unsigned short* lpval = nullptr;
    int dtype = 0;
};

namespace clspmv {

// HIP call wrapper: records the failing call and returns CLSPMV_ERR_HIP.
#define CLSPMV_HIP(call)                                                        \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            return ::clspmv::fail(CLSPMV_ERR_HIP, "%s failed: %s (%s:%d)", #call,  \
                                  hipGetErrorString(e_), __FILE__, __LINE__);   \
    } while (0)

// Device copy of a host array, owned by `m`.
template <class T>
int upload_array(clspmv_dmat* m, const T* host, size_t n, T** out) {
    void* p = nullptr;
    size_t bytes = sizeof(T) * (n ? n : 1);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
        return fail(CLSPMV_ERR_ALLOC, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    m->bufs.push_back({p, bytes});
    m->info.device_bytes += bytes;
    if (n && host) {
        e = hipMemcpy(p, host, sizeof(T) * n, hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return fail(CLSPMV_ERR_HIP, "hipMemcpy failed: %s", hipGetErrorString(e));
    }
    *out = static_cast<T*>(p);
    return CLSPMV_OK;
}

// Frees one allocation of `m` (no-op for null); the pointer is cleared.
template <class T>
void release_array(clspmv_dmat* m, T** p) {
    if (!*p) return;
    for (size_t i = 0; i < m->bufs.size(); ++i)
        if (m->bufs[i].first == (void*)*p) {
            (void)hipFree(m->bufs[i].first);
            m->info.device_bytes -= m->bufs[i].second;
            m->bufs.erase(m->bufs.begin() + i);
            break;
        }
    *p = nullptr;
}

// x-panel encoding (k_panel.hip) of row-sorted entries, built on the host
// and uploaded into `d`.  pw: panel width (multiple of 8, at most
// launch::kPanelMax); group: sub-chunks per pass-1 work-group, 0 = sized by
// panel_group_size.
int build_panel(clspmv_dmat* d, int rows, int cols, long long nnz, const int* row,
                const int* col, const float* val, int pw, int dtype, int group);
// Sub-chunks per pass-1 work-group for panels holding sc[p] sub-chunks: the
// size whose work-groups finish soonest when `slots` of them run at once,
// each costing its sub-chunks plus the staging of its panel.
int panel_group_size(const std::vector<int>& sc, int slots, int pw);
// Panel width the encoding rule picks for a matrix, or 0 when panels do not
// pay (too few nonzeros per panel to amortize staging x in LDS).
int auto_panel_width(int cols, long long nnz);
// Drops one of two uploaded encodings: the panel one (panel == true) or the
// plain arrays.
void drop_encoding(clspmv_dmat* d, bool panel);

// Looks up, with `device` current, the device address of the scratch array
// of clspmv_verify_spmv_noalloc (k_verify.hip); once per device.
int noalloc_scratch_init(int device);

// Checks that the context's device is current; sets it otherwise.
int use_ctx_device(clspmv_ctx* ctx);

// Launch y += A*x for one device matrix on `stream` (no argument checks
// beyond the method's shape requirements).
int dmat_launch(hipStream_t stream, const clspmv_dmat* A, int method,
                const float* x, float* y);

// The same launch in two phases, for callers that overlap one matrix's
// x-only work with other kernels: dmat_prepass issues the kernels that read
// x and write only the matrix's own scratch (panel pass 1; nothing for
// other encodings), dmat_launch_rest the ones that update y, which must be
// ordered after the prepass.  dmat_launch = prepass then rest.
bool dmat_has_prepass(const clspmv_dmat* A);
// dia (optional, with y): an FP32 DIA part whose rows the prepass launch
// also sums into y (dmat_dia_tail_ok), with DIA method 1's arithmetic.
bool dmat_dia_tail_ok(const clspmv_dmat* D);
int dmat_prepass(hipStream_t stream, const clspmv_dmat* A, const float* x, const clspmv_dmat* dia = nullptr,
                 float* y = nullptr);
int dmat_launch_rest(hipStream_t stream, const clspmv_dmat* A, int method,
                     const float* x, float* y);

// Multi-vector SpMV (DIA, CSR with its own arrays): Y[:, q] += A X[:, q],
// X row-major, Y column-major; no argument checks.  kb > 0: CSR runs the
// kernel built for kb right-hand sides whatever k is (launch::csr_multi);
// each right-hand side's result then does not depend on k.  DIA's never
// does.
bool has_multi(const clspmv_dmat* A);
int multi_launch(hipStream_t s, const clspmv_dmat* A, int k, const float* X, float* Y, int ldy, int kb = 0);

// Method chosen for CLSPMV_METHOD_AUTO.
int auto_method(const clspmv_dmat* A);

}  // namespace clspmv
