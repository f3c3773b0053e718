// Device runtime of the clspmv engine: contexts, device-resident matrices of
// every storage format, launch dispatch, event timing and the per-format
// benchmark drivers.
//
// Uploads replace the reference's cpy_<fmt>_mat (src/spmv_cocktail.cpp:
// 304-962): the host layouts already give coalesced wave accesses, so they
// are copied as they are, plus per-format launch metadata computed here
// (CSR row blocks, DIA offset range, band widths, lane-group sizes).  Column
// indices are range-checked on upload so that no kernel can read outside x.

#include <algorithm>
#include <cstring>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <numeric>

#include "kernels.hpp"
#include "runtime.hpp"

using namespace clspmv;

namespace clspmv {

int use_ctx_device(clspmv_ctx* ctx) {
    int cur = -1;
    CLSPMV_HIP(hipGetDevice(&cur));
    if (cur != ctx->device) CLSPMV_HIP(hipSetDevice(ctx->device));
    return CLSPMV_OK;
}

namespace {

int pow2_at_least(int v, int cap) {
    int p = 1;
    while (p < v && p < cap) p <<= 1;
    return p;
}

int check_cols(const int* c, size_t n, int cols, int scale, const char* what) {
    for (size_t i = 0; i < n; ++i) {
        long long e = (long long)c[i] * scale;
        if (c[i] < 0 || e >= cols)
            return fail(CLSPMV_ERR_INVALID, "%s: column %lld outside width %d",
                        what, e, cols);
    }
    return CLSPMV_OK;
}

clspmv_dmat* new_dmat(int format, const clspmv_matinfo& mi) {
    clspmv_dmat* m = new clspmv_dmat();
    m->info.format = format;
    m->info.rows = mi.height;
    m->info.cols = mi.width;
    m->info.nnz = mi.nnz;
    m->info.bheight = 1;
    m->info.bwidth = 1;
    (void)hipGetDevice(&m->device);
    return m;
}

// Row blocks of the adaptive CSR kernel: a block is a maximal run of at most
// kAdaptiveRows rows whose entries fit kAdaptiveNnz, or a single longer row.
std::vector<int> adaptive_blocks(const int* rp, int rows) {
    std::vector<int> b{0};
    int r = 0;
    while (r < rows) {
        if (rp[r + 1] - rp[r] > launch::kAdaptiveNnz) {
            b.push_back(++r);
            continue;
        }
        int r2 = r;
        while (r2 < rows && r2 - r < launch::kAdaptiveRows &&
               rp[r2 + 1] - rp[r] <= launch::kAdaptiveNnz)
            ++r2;
        b.push_back(r2);
        r = r2;
    }
    return b;
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

int auto_method(const clspmv_dmat* A) {
    switch (A->info.format) {
        case CLSPMV_FMT_CSR: {
            double avg = A->info.rows ? (double)A->info.nnz / A->info.rows : 0.0;
            return A->max_row > 8.0 * std::max(avg, 4.0) ? 1 : 0;
        }
        case CLSPMV_FMT_ELL: return A->hal % 4 == 0 ? 1 : 0;
        case CLSPMV_FMT_DIA: return 0;
        case CLSPMV_FMT_BDIA: return 0;
        default: return 0;
    }
}

bool dmat_has_prepass(const clspmv_dmat* A) {
    const clspmv_dmat_info& I = A->info;
    return (I.format == CLSPMV_FMT_CSR || I.format == CLSPMV_FMT_COO) && A->use_panel && I.rows > 0 &&
           I.cols > 0 && I.nnz > 0;
}

bool dmat_dia_tail_ok(const clspmv_dmat* D) {
    const clspmv_dmat_info& I = D->info;
    return I.format == CLSPMV_FMT_DIA && D->dtype == CLSPMV_DT_F32 && I.rows > 0 && I.cols > 0 && I.nnz > 0 &&
           D->ndiag > 0;
}

int dmat_prepass(hipStream_t s, const clspmv_dmat* A, const float* x, const clspmv_dmat* dia, float* y) {
    if (!dmat_has_prepass(A)) return CLSPMV_OK;
    launch::DiaTail tail{0, 0, 0, 0, nullptr, nullptr, nullptr};
    if (dia) {
        if (!dmat_dia_tail_ok(dia) || !y) return fail(CLSPMV_ERR_INVALID, "DIA part not fit for the pass-1 tail");
        tail = launch::DiaTail{std::min(dia->info.rows, dia->len), dia->info.cols, dia->len, dia->ndiag, dia->offs,
                               dia->val, y};
    }
    const hipError_t e = launch::panel_pass1(s, A->pnwg, A->pwg, A->pw, A->info.cols, A->pcol, A->dtype,
                                             A->dtype ? (const void*)A->plpval : (const void*)A->pval, A->ppad,
                                             A->psc_seg, A->psc_start, x, aligned16(x), A->pseg,
                                             dia ? &tail : nullptr);
    if (e != hipSuccess)
        return fail(CLSPMV_ERR_HIP, "panel pass 1 failed (format %d): %s", A->info.format, hipGetErrorString(e));
    return CLSPMV_OK;
}

namespace {
int resolve_method(const clspmv_dmat* A, int* method) {
    if (*method == CLSPMV_METHOD_AUTO) *method = auto_method(A);
    if (*method < 0 || *method >= clspmv_method_count(A->info.format))
        return fail(CLSPMV_ERR_UNSUPPORTED, "method %d not available for format %d", *method,
                    A->info.format);
    return CLSPMV_OK;
}
}  // namespace

int dmat_launch(hipStream_t s, const clspmv_dmat* A, int method,
                const float* x, float* y) {
    int st = resolve_method(A, &method);
    if (!st) st = dmat_prepass(s, A, x);
    return st ? st : dmat_launch_rest(s, A, method, x, y);
}

int dmat_launch_rest(hipStream_t s, const clspmv_dmat* A, int method,
                     const float* x, float* y) {
    const clspmv_dmat_info& I = A->info;
    const int st = resolve_method(A, &method);
    if (st) return st;
    if (I.rows == 0 || I.cols == 0 || I.nnz == 0) return CLSPMV_OK;
    hipError_t e = hipSuccess;
    if (dmat_has_prepass(A)) {
        // The panel encoding serves every method of the format; pass 1 has
        // filled the segment sums.
        e = launch::panel_pass2(s, A->pseg, A->pnb, A->pblk_row, A->pblk_seg, A->prow_slot, A->pseg_dst,
                                A->pblk_long, A->plong, y);
        if (e != hipSuccess)
            return fail(CLSPMV_ERR_HIP, "panel pass 2 failed (format %d): %s", I.format,
                        hipGetErrorString(e));
        return CLSPMV_OK;
    }
    const bool xa = aligned16(x);
    switch (I.format) {
        case CLSPMV_FMT_CSR:
            if (A->dtype) {  // 16-bit values: the multi-vector kernel with k = 1
                e = launch::csr_multi(s, A->nblocks, A->blocks, A->rp, A->col, A->dtype, A->lpval, 1, x, y,
                                      I.rows);
                break;
            }
            e = method == 0 ? launch::csr_vector(s, I.rows, A->lanes, A->rp, A->col, A->val, x, y)
                            : launch::csr_adaptive(s, A->nblocks, A->blocks, A->rp, A->col, A->val, x, y);
            break;
        case CLSPMV_FMT_COO:
            e = launch::coo_segreduce(s, I.nnz, A->row, A->col, A->val, x, y,
                                      A->carry_row, A->carry_val);
            break;
        case CLSPMV_FMT_ELL:
            if (method == 1 && A->hal % 4 != 0)
                return fail(CLSPMV_ERR_UNSUPPORTED, "ELL method 1 needs height_aligned %% 4 == 0");
            e = launch::ell(s, I.rows, A->hal, A->ellnum, method == 1 ? 4 : 1, A->col, A->val, x, y);
            break;
        case CLSPMV_FMT_SELL:
            e = launch::sell(s, I.rows, I.slice_height, A->sptr, A->col, A->val, x, y);
            break;
        case CLSPMV_FMT_DIA: {
            const int span = A->max_off - A->min_off;
            if (method == 1 && A->len % 2)
                return fail(CLSPMV_ERR_UNSUPPORTED, "DIA method 1 needs an even aligned length");
            if (method >= 2 && (method == 3 ? 4 : 1) * 256 + span > launch::kLdsWindow)
                return fail(CLSPMV_ERR_UNSUPPORTED, "DIA method %d: offset span %d too wide for LDS", method, span);
            e = launch::dia(s, std::min(I.rows, A->len), I.cols, A->len, A->ndiag, A->offs,
                            A->min_off, A->max_off, method, A->dtype,
                            A->dtype ? (const void*)A->lpval : (const void*)A->val, x, y);
            break;
        }
        case CLSPMV_FMT_BDIA: {
            static const int rpt[5] = {1, 1, 2, 2, 4};
            static const bool lds[5] = {false, true, true, false, true};
            if (lds[method] && rpt[method] * 256 + A->max_band > launch::kLdsWindow / 2)
                return fail(CLSPMV_ERR_UNSUPPORTED, "BDIA method %d: band width %d too wide for LDS", method, A->max_band);
            if (method == 0 || method == 3) {
                if (method == 3 && A->len % 2)
                    return fail(CLSPMV_ERR_UNSUPPORTED, "BDIA method 3 needs an even aligned length");
                e = launch::dia(s, std::min(I.rows, A->len), I.cols, A->len, A->ndiag, A->offs_exp,
                                A->min_off, A->max_off, method == 0 ? 0 : 1, 0, A->val, x, y);
                break;
            }
            e = launch::bdia(s, std::min(I.rows, A->len), I.cols, A->len, A->nband, A->offs,
                             A->bptr, A->max_band, method, A->val, x, y);
            break;
        }
        case CLSPMV_FMT_BCSR:
            if (A->bslots > 0)
                e = launch::bcsr_slots(s, A->brows, I.rows, I.cols, I.bheight, I.bwidth, A->hal, A->bslots,
                                       A->col, A->val, x, xa, y);
            else
                e = launch::bcsr(s, A->brows, I.rows, I.cols, I.bheight, I.bwidth, A->lanes,
                                 A->aligned, A->rp, A->col, A->val, x, xa, y);
            break;
        case CLSPMV_FMT_BELL:
            e = launch::bell(s, A->brows, I.rows, I.cols, I.bheight, I.bwidth, method,
                             A->hal, A->f4, A->ellnum, A->col, A->val, x, xa, y);
            break;
        case CLSPMV_FMT_SBELL:
            e = launch::sbell(s, A->brows, I.rows, I.cols, I.bheight, I.bwidth,
                              I.slice_height, A->sptr, A->col, A->val, x, xa, y);
            break;
        default:
            return fail(CLSPMV_ERR_INVALID, "unknown format %d", I.format);
    }
    if (e != hipSuccess)
        return fail(CLSPMV_ERR_HIP, "launch failed (format %d method %d): %s", I.format,
                    method, hipGetErrorString(e));
    return CLSPMV_OK;
}

bool has_multi(const clspmv_dmat* A) {
    return A->info.format == CLSPMV_FMT_DIA || (A->info.format == CLSPMV_FMT_CSR && A->plain);
}

int multi_launch(hipStream_t s, const clspmv_dmat* A, int k, const float* X, float* Y, int ldy, int kb) {
    const clspmv_dmat_info& I = A->info;
    hipError_t e;
    if (I.format == CLSPMV_FMT_DIA) {
        e = launch::dia_multi(s, std::min(I.rows, A->len), I.cols, A->len, A->ndiag, A->offs, A->dtype,
                              A->dtype ? (const void*)A->lpval : (const void*)A->val, k, X, Y, ldy);
    } else if (I.format == CLSPMV_FMT_CSR && A->plain) {
        e = launch::csr_multi(s, A->nblocks, A->blocks, A->rp, A->col, A->dtype,
                              A->dtype ? (const void*)A->lpval : (const void*)A->val, k, X, Y, ldy, kb);
    } else {
        return fail(CLSPMV_ERR_UNSUPPORTED, "multi-vector SpMV exists for DIA and CSR (own arrays) only");
    }
    if (e != hipSuccess) return fail(CLSPMV_ERR_HIP, "multi-vector launch failed: %s", hipGetErrorString(e));
    return CLSPMV_OK;
}

}  // namespace clspmv

extern "C" {

// ---- contexts ----------------------------------------------------------------

int clspmv_ctx_create(int device_ordinal, clspmv_ctx** out) {
    if (!out) return fail(CLSPMV_ERR_INVALID, "null output");
    int n = 0;
    CLSPMV_HIP(hipGetDeviceCount(&n));
    if (device_ordinal < 0 || device_ordinal >= n)
        return fail(CLSPMV_ERR_INVALID, "device %d not present (%d devices)", device_ordinal, n);
    CLSPMV_HIP(hipSetDevice(device_ordinal));
    int st = noalloc_scratch_init(device_ordinal);
    if (st) return st;
    clspmv_ctx* c = new clspmv_ctx();
    c->device = device_ordinal;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_start);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_stop);
    if (e != hipSuccess) {
        delete c;
        return fail(CLSPMV_ERR_HIP, "stream/event creation failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return CLSPMV_OK;
}

int clspmv_ctx_destroy(clspmv_ctx* ctx) {
    if (!ctx) return CLSPMV_OK;
    use_ctx_device(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(ctx->ev_start);
    (void)hipEventDestroy(ctx->ev_stop);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CLSPMV_OK;
}

int clspmv_ctx_synchronize(clspmv_ctx* ctx) {
    if (!ctx) return fail(CLSPMV_ERR_INVALID, "null context");
    CLSPMV_HIP(hipStreamSynchronize(ctx->stream));
    return CLSPMV_OK;
}

void* clspmv_ctx_stream(clspmv_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }
int clspmv_ctx_device(clspmv_ctx* ctx) { return ctx ? ctx->device : -1; }

int clspmv_device_info(clspmv_ctx* ctx, char* name, int name_len, int* cu_count) {
    if (!ctx) return fail(CLSPMV_ERR_INVALID, "null context");
    hipDeviceProp_t p;
    CLSPMV_HIP(hipGetDeviceProperties(&p, ctx->device));
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    return CLSPMV_OK;
}

int clspmv_device_malloc(clspmv_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(CLSPMV_ERR_INVALID, "null argument");
    int st = use_ctx_device(ctx);
    if (st) return st;
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(CLSPMV_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return CLSPMV_OK;
}

int clspmv_device_free(clspmv_ctx* ctx, void* ptr) {
    if (!ctx) return fail(CLSPMV_ERR_INVALID, "null context");
    if (ptr) CLSPMV_HIP(hipFree(ptr));
    return CLSPMV_OK;
}

int clspmv_memcpy_h2d(clspmv_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return fail(CLSPMV_ERR_INVALID, "null context");
    if (!bytes) return CLSPMV_OK;
    CLSPMV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    CLSPMV_HIP(hipStreamSynchronize(ctx->stream));
    return CLSPMV_OK;
}

int clspmv_memcpy_d2h(clspmv_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return fail(CLSPMV_ERR_INVALID, "null context");
    if (!bytes) return CLSPMV_OK;
    CLSPMV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    CLSPMV_HIP(hipStreamSynchronize(ctx->stream));
    return CLSPMV_OK;
}

int clspmv_memcpy_d2d(clspmv_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return fail(CLSPMV_ERR_INVALID, "null context");
    if (!bytes) return CLSPMV_OK;
    CLSPMV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return CLSPMV_OK;
}

int clspmv_memset_zero(clspmv_ctx* ctx, void* dst, size_t bytes) {
    if (!ctx) return fail(CLSPMV_ERR_INVALID, "null context");
    if (!bytes) return CLSPMV_OK;
    CLSPMV_HIP(hipMemsetAsync(dst, 0, bytes, ctx->stream));
    return CLSPMV_OK;
}

// ---- device matrices -----------------------------------------------------------

int clspmv_method_count(int format) {
    switch (format) {
        case CLSPMV_FMT_BSPMM: return 0;  // clspmv_spmm only
        case CLSPMV_FMT_CSR: return CLSPMV_CSR_IMP_NUM;
        case CLSPMV_FMT_COO: return CLSPMV_COO_IMP_NUM;
        case CLSPMV_FMT_ELL: return CLSPMV_ELL_IMP_NUM;
        case CLSPMV_FMT_SELL: return CLSPMV_SELL_IMP_NUM;
        case CLSPMV_FMT_DIA: return CLSPMV_DIA_IMP_NUM;
        case CLSPMV_FMT_BDIA: return CLSPMV_BDIA_IMP_NUM;
        case CLSPMV_FMT_BCSR: return CLSPMV_BCSR_IMP_NUM;
        case CLSPMV_FMT_BELL: return CLSPMV_BELL_IMP_NUM;
        case CLSPMV_FMT_SBELL: return CLSPMV_SBELL_IMP_NUM;
        default: return 0;
    }
}

#define UPLOAD_PROLOGUE(m, fmt)                                               \
    if (!ctx || !m || !out) return fail(CLSPMV_ERR_INVALID, "null argument"); \
    {                                                                         \
        int st0 = use_ctx_device(ctx);                                        \
        if (st0) return st0;                                                  \
    }                                                                         \
    clspmv_dmat* d = new_dmat(fmt, m->matinfo);                               \
    int st = CLSPMV_OK;

#define UPLOAD_CHECK(expr)            \
    do {                              \
        st = (expr);                  \
        if (st) {                     \
            clspmv_dmat_free(d);      \
            return st;                \
        }                             \
    } while (0)

// Builds the panel encoding asked for by opts (a width, or
// CLSPMV_PANEL_AUTO / CLSPMV_PANEL_BOTH) from row-sorted entries; with AUTO
// it replaces the plain arrays, with BOTH it is added beside them.
static int add_panel(clspmv_dmat* d, const clspmv_upload_opts& o, const int* row, const int* col,
                     const float* val) {
    const clspmv_dmat_info& I = d->info;
    int pw = o.panel_width;
    // The plain arrays can serve alone: FP32 values, or 16-bit CSR values
    // kept beside the panels (BOTH).
    const bool plain_ok = o.dtype == CLSPMV_DT_F32 ||
                          (o.panel_width == CLSPMV_PANEL_BOTH && I.format == CLSPMV_FMT_CSR);
    if (pw == CLSPMV_PANEL_AUTO || pw == CLSPMV_PANEL_BOTH) {
        pw = auto_panel_width(I.cols, I.nnz);
        if (pw == 0 && !plain_ok) pw = std::min(launch::kPanelMax, round_up(I.cols, 8));
    }
    if (pw == 0) return CLSPMV_OK;
    int st = build_panel(d, I.rows, I.cols, I.nnz, row, col, val, pw, o.dtype, o.panel_group);
    if (st == CLSPMV_ERR_UNSUPPORTED && o.panel_width < 0 && plain_ok)
        return CLSPMV_OK;  // stay plain
    if (st) return st;
    d->dtype = o.dtype;
    if (o.panel_width != CLSPMV_PANEL_BOTH) drop_encoding(d, false);
    return CLSPMV_OK;
}

static int check_opts(const clspmv_upload_opts& o, bool csr) {
    if (o.dtype != CLSPMV_DT_F32 && o.dtype != CLSPMV_DT_BF16 && o.dtype != CLSPMV_DT_F16)
        return fail(CLSPMV_ERR_INVALID, "dtype %d", o.dtype);
    if (o.panel_group != 0 && (o.panel_group < 8 || o.panel_group > launch::kPanelScPerWg))
        return fail(CLSPMV_ERR_INVALID, "panel_group %d: need 0 or [8, %d]", o.panel_group,
                    launch::kPanelScPerWg);
    const int pw = o.panel_width;
    if (o.dtype != CLSPMV_DT_F32 && !csr && (pw == CLSPMV_PANEL_BOTH || pw == 0))
        return fail(CLSPMV_ERR_UNSUPPORTED, "16-bit COO values need the panel encoding only");
    if (pw == 0 || pw == CLSPMV_PANEL_AUTO || pw == CLSPMV_PANEL_BOTH) return CLSPMV_OK;
    if (pw < 8 || pw % 8 || pw > launch::kPanelMax)
        return fail(CLSPMV_ERR_INVALID, "panel width %d: need 0, -1, -2 or a multiple of 8 in [8, %d]", pw,
                    launch::kPanelMax);
    return CLSPMV_OK;
}

int clspmv_upload_csr(clspmv_ctx* ctx, const clspmv_csr_matrix* m, clspmv_dmat** out) {
    return clspmv_upload_csr_ex(ctx, m, nullptr, out);
}

int clspmv_upload_coo(clspmv_ctx* ctx, const clspmv_coo_matrix* m, clspmv_dmat** out) {
    return clspmv_upload_coo_ex(ctx, m, nullptr, out);
}

int clspmv_upload_csr_ex(clspmv_ctx* ctx, const clspmv_csr_matrix* m,
                         const clspmv_upload_opts* opts, clspmv_dmat** out) {
    const clspmv_upload_opts o = opts ? *opts : clspmv_upload_opts{0, CLSPMV_DT_F32};
    {
        int st0 = check_opts(o, true);
        if (st0) return st0;
    }
    UPLOAD_PROLOGUE(m, CLSPMV_FMT_CSR)
    const int rows = m->matinfo.height, nnz = m->matinfo.nnz;
    if (rows < 0 || nnz < 0 || (rows > 0 && !m->csr_row_ptr))
        UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "malformed CSR matrix"));
    if (rows > 0 && (m->csr_row_ptr[0] != 0 || m->csr_row_ptr[rows] != nnz))
        UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "CSR row pointer does not span nnz"));
    int max_row = 0;
    for (int r = 0; r < rows; ++r) {
        int len = m->csr_row_ptr[r + 1] - m->csr_row_ptr[r];
        if (len < 0) UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "CSR row pointer decreases at %d", r));
        max_row = std::max(max_row, len);
    }
    UPLOAD_CHECK(check_cols(m->csr_col_id, nnz, m->matinfo.width, 1, "csr"));
    d->info.stored = nnz;
    d->max_row = max_row;
    double avg = rows ? (double)nnz / rows : 0.0;
    d->lanes = pow2_at_least((int)std::ceil(avg), 64);
    UPLOAD_CHECK(upload_array(d, m->csr_row_ptr, (size_t)rows + 1, &d->rp));
    UPLOAD_CHECK(upload_array(d, m->csr_col_id, nnz, &d->col));
    if (o.dtype != CLSPMV_DT_F32) {
        // Own arrays with 16-bit values (run by the multi-vector kernel).
        std::vector<unsigned short> lp(nnz);
        UPLOAD_CHECK(clspmv_convert_from_f32(m->csr_data, lp.data(), (size_t)nnz, o.dtype));
        UPLOAD_CHECK(upload_array(d, lp.data(), lp.size(), &d->lpval));
        d->dtype = o.dtype;
    } else {
        UPLOAD_CHECK(upload_array(d, m->csr_data, nnz, &d->val));
    }
    std::vector<int> blocks = adaptive_blocks(m->csr_row_ptr, rows);
    d->nblocks = (int)blocks.size() - 1;
    UPLOAD_CHECK(upload_array(d, blocks.data(), blocks.size(), &d->blocks));

    if (o.panel_width != 0 && nnz > 0) {
        std::vector<int> row_of(nnz);
        for (int r = 0; r < rows; ++r)
            std::fill(row_of.begin() + m->csr_row_ptr[r], row_of.begin() + m->csr_row_ptr[r + 1], r);
        UPLOAD_CHECK(add_panel(d, o, row_of.data(), m->csr_col_id, m->csr_data));
    }
    *out = d;
    return CLSPMV_OK;
}

int clspmv_upload_coo_ex(clspmv_ctx* ctx, const clspmv_coo_matrix* m,
                         const clspmv_upload_opts* opts, clspmv_dmat** out) {
    const clspmv_upload_opts o = opts ? *opts : clspmv_upload_opts{CLSPMV_PANEL_AUTO, CLSPMV_DT_F32};
    {
        int st0 = check_opts(o, false);
        if (st0) return st0;
    }
    UPLOAD_PROLOGUE(m, CLSPMV_FMT_COO)
    UPLOAD_CHECK(clspmv_validate_coo(m));
    const int nnz = m->matinfo.nnz;
    d->info.stored = nnz;
    // The segmented reduction needs row-sorted entries: sort a copy if needed.
    clspmv_coo_matrix sorted = *m;
    std::vector<int> r, c;
    std::vector<float> v;
    if (!coo_sorted(*m)) {
        r.assign(m->coo_row_id, m->coo_row_id + nnz);
        c.assign(m->coo_col_id, m->coo_col_id + nnz);
        v.assign(m->coo_data, m->coo_data + nnz);
        sorted.coo_row_id = r.data();
        sorted.coo_col_id = c.data();
        sorted.coo_data = v.data();
        coo_sort(sorted);
    }
    UPLOAD_CHECK(upload_array(d, sorted.coo_row_id, nnz, &d->row));
    UPLOAD_CHECK(upload_array(d, sorted.coo_col_id, nnz, &d->col));
    UPLOAD_CHECK(upload_array(d, sorted.coo_data, nnz, &d->val));
    const int chunks = (nnz + launch::kCooChunk - 1) / launch::kCooChunk;
    UPLOAD_CHECK(upload_array<int>(d, nullptr, 2 * (size_t)chunks, &d->carry_row));
    UPLOAD_CHECK(upload_array<float>(d, nullptr, 2 * (size_t)chunks, &d->carry_val));
    if (o.panel_width != 0 && nnz > 0)
        UPLOAD_CHECK(add_panel(d, o, sorted.coo_row_id, sorted.coo_col_id, sorted.coo_data));
    *out = d;
    return CLSPMV_OK;
}

int clspmv_upload_ell(clspmv_ctx* ctx, const clspmv_ell_matrix* m, clspmv_dmat** out) {
    UPLOAD_PROLOGUE(m, CLSPMV_FMT_ELL)
    const size_t total = (size_t)m->ell_height_aligned * m->ell_num;
    if (m->ell_height_aligned < m->matinfo.height || m->ell_num < 0)
        UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "malformed ELL matrix"));
    UPLOAD_CHECK(check_cols(m->ell_col_id, total, std::max(m->matinfo.width, 1), 1, "ell"));
    d->hal = m->ell_height_aligned;
    d->ellnum = m->ell_num;
    d->info.stored = (long long)total;
    UPLOAD_CHECK(upload_array(d, m->ell_col_id, total, &d->col));
    UPLOAD_CHECK(upload_array(d, m->ell_data, total, &d->val));
    *out = d;
    return CLSPMV_OK;
}

int clspmv_upload_sell(clspmv_ctx* ctx, const clspmv_sell_matrix* m, clspmv_dmat** out) {
    UPLOAD_PROLOGUE(m, CLSPMV_FMT_SELL)
    const int h = (int)m->sell_slice_height, ns = m->sell_slice_num;
    if (h <= 0 || ns < 0 || (long long)ns * h < m->matinfo.height || !m->sell_slice_ptr)
        UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "malformed SELL matrix"));
    for (int s = 0; s < ns; ++s) {
        int w = m->sell_slice_ptr[s + 1] - m->sell_slice_ptr[s];
        if (w < 0 || w % h) UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "SELL slice %d size %d", s, w));
    }
    const size_t total = (size_t)m->sell_slice_ptr[ns];
    UPLOAD_CHECK(check_cols(m->sell_col_id, total, std::max(m->matinfo.width, 1), 1, "sell"));
    d->info.slice_height = h;
    d->info.stored = (long long)total;
    UPLOAD_CHECK(upload_array(d, m->sell_slice_ptr, (size_t)ns + 1, &d->sptr));
    UPLOAD_CHECK(upload_array(d, m->sell_col_id, total, &d->col));
    UPLOAD_CHECK(upload_array(d, m->sell_data, total, &d->val));
    *out = d;
    return CLSPMV_OK;
}

int clspmv_upload_dia(clspmv_ctx* ctx, const clspmv_dia_matrix* m, clspmv_dmat** out) {
    return clspmv_upload_dia_ex(ctx, m, nullptr, out);
}

int clspmv_upload_dia_ex(clspmv_ctx* ctx, const clspmv_dia_matrix* m, const clspmv_upload_opts* opts,
                         clspmv_dmat** out) {
    const int dtype = opts ? opts->dtype : CLSPMV_DT_F32;
    if (dtype != CLSPMV_DT_F32 && dtype != CLSPMV_DT_BF16 && dtype != CLSPMV_DT_F16)
        return fail(CLSPMV_ERR_INVALID, "dtype %d", dtype);
    UPLOAD_PROLOGUE(m, CLSPMV_FMT_DIA)
    if (m->dia_num < 0 || m->dia_length_aligned < 0 || (m->dia_num > 0 && !m->dia_offsets))
        UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "malformed DIA matrix"));
    d->len = m->dia_length_aligned;
    d->ndiag = m->dia_num;
    d->min_off = d->max_off = 0;
    for (int j = 0; j < m->dia_num; ++j) {
        if (j == 0 || m->dia_offsets[j] < d->min_off) d->min_off = m->dia_offsets[j];
        if (j == 0 || m->dia_offsets[j] > d->max_off) d->max_off = m->dia_offsets[j];
    }
    d->info.stored = (long long)m->dia_num * m->dia_length_aligned;
    UPLOAD_CHECK(upload_array(d, m->dia_offsets, (size_t)m->dia_num, &d->offs));
    const size_t total = (size_t)m->dia_num * m->dia_length_aligned;
    if (dtype == CLSPMV_DT_F32) {
        UPLOAD_CHECK(upload_array(d, m->dia_data, total, &d->val));
    } else {
        // A DIA row holds at most dia_num entries; the DIA kernels sum in FP32,
        // so 16-bit values are refused where rows could exceed the FP32 limit.
        if (m->dia_num > launch::kLongRowF64)
            UPLOAD_CHECK(fail(CLSPMV_ERR_UNSUPPORTED,
                              "16-bit DIA with %d diagonals: rows longer than %d need FP64 sums (use CSR)",
                              m->dia_num, launch::kLongRowF64));
        std::vector<unsigned short> lp(total);
        UPLOAD_CHECK(clspmv_convert_from_f32(m->dia_data, lp.data(), total, dtype));
        UPLOAD_CHECK(upload_array(d, lp.data(), total, &d->lpval));
        d->dtype = dtype;
    }
    *out = d;
    return CLSPMV_OK;
}

int clspmv_upload_bdia(clspmv_ctx* ctx, const clspmv_bdia_matrix* m, clspmv_dmat** out) {
    UPLOAD_PROLOGUE(m, CLSPMV_FMT_BDIA)
    const int nb = m->bdia_band_num;
    if (nb < 0 || m->bdia_length_aligned < 0 || (nb > 0 && (!m->bdia_offsets || !m->bdia_bptr)))
        UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "malformed BDIA matrix"));
    d->len = m->bdia_length_aligned;
    d->nband = nb;
    d->max_band = 0;
    for (int b = 0; b < nb; ++b) {
        int w = m->bdia_bptr[b + 1] - m->bdia_bptr[b];
        if (w <= 0) UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "BDIA band %d is empty", b));
        d->max_band = std::max(d->max_band, w);
    }
    const int ndiag = nb > 0 ? m->bdia_bptr[nb] : 0;
    d->ndiag = ndiag;
    std::vector<int> exp_off;
    for (int b = 0; b < nb; ++b)
        for (int j = m->bdia_bptr[b]; j < m->bdia_bptr[b + 1]; ++j)
            exp_off.push_back(m->bdia_offsets[b] + (j - m->bdia_bptr[b]));
    if (!exp_off.empty()) {
        d->min_off = *std::min_element(exp_off.begin(), exp_off.end());
        d->max_off = *std::max_element(exp_off.begin(), exp_off.end());
    }
    UPLOAD_CHECK(upload_array(d, exp_off.data(), exp_off.size(), &d->offs_exp));
    d->info.stored = (long long)ndiag * m->bdia_length_aligned;
    UPLOAD_CHECK(upload_array(d, m->bdia_offsets, (size_t)nb, &d->offs));
    UPLOAD_CHECK(upload_array(d, m->bdia_bptr, (size_t)nb + 1, &d->bptr));
    UPLOAD_CHECK(upload_array(d, m->bdia_data, (size_t)ndiag * m->bdia_length_aligned, &d->val));
    *out = d;
    return CLSPMV_OK;
}

static int block_shape_ok(unsigned bh, unsigned bw) {
    return (bh == 1 || bh == 2 || bh == 4 || bh == 8) && (bw == 4 || bw == 8);
}

int clspmv_upload_b4csr(clspmv_ctx* ctx, const clspmv_b4csr_matrix* m, clspmv_dmat** out) {
    UPLOAD_PROLOGUE(m, CLSPMV_FMT_BCSR)
    const unsigned bh = m->b4csr_bheight, bw = m->b4csr_bwidth;
    if (!block_shape_ok(bh, bw))
        UPLOAD_CHECK(fail(CLSPMV_ERR_UNSUPPORTED, "b4csr block %ux%u not supported", bh, bw));
    const int nbr = m->b4csr_row_num, nb = m->b4csr_block_num;
    if (nbr < 0 || nb < 0 || (long long)nbr * bh < m->matinfo.height ||
        m->b4csr_aligned_size < 4LL * nb || !m->b4csr_row_ptr)
        UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "malformed b4csr matrix"));
    // col4 + bw/4 - 1 float4s must start inside the padded width.
    UPLOAD_CHECK(check_cols(m->b4csr_col_id, nb, std::max(m->matinfo.width, 1), 4, "b4csr"));
    d->info.bheight = (int)bh;
    d->info.bwidth = (int)bw;
    d->brows = nbr;
    d->aligned = m->b4csr_aligned_size;
    d->info.stored = (long long)nb * bh * bw;
    double avg = nbr ? (double)nb / nbr : 0.0;
    d->lanes = pow2_at_least((int)std::ceil(avg), 64);
    const int* rp = m->b4csr_row_ptr;
    int w = 0;
    for (int r = 0; r < nbr; ++r) {
        if (rp[r + 1] < rp[r]) UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "b4csr row pointers decrease at %d", r));
        w = std::max(w, rp[r + 1] - rp[r]);
    }
    if (rp[0] != 0 || rp[nbr] != nb) UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "b4csr row pointers do not span the blocks"));
    // Slot-major encoding when every block row fits kBcsrSlotsMax slots and
    // padding the rows to the longest adds at most 10 % to the blocks: a
    // lane per block row then reads each slot's columns and values as
    // contiguous runs across the wave, as BELL does.  The row-pointer
    // arrays are not uploaded.
    if (nb > 0 && w <= launch::kBcsrSlotsMax && (long long)w * nbr * 10 <= (long long)nb * 11) {
        const int bw4 = (int)bw / 4;
        const long long hal = ((long long)nbr + 63) / 64 * 64;
        std::vector<int> scol((size_t)(hal * w), -1);
        std::vector<float> sval((size_t)(hal * w * bw4 * bh * 4), 0.0f);
        const size_t plane = (size_t)m->b4csr_aligned_size * bh;
        for (int r = 0; r < nbr; ++r)
            for (int k = 0; k < rp[r + 1] - rp[r]; ++k) {
                const int j = rp[r] + k;
                scol[(size_t)k * hal + r] = m->b4csr_col_id[j];
                for (int w4 = 0; w4 < bw4; ++w4)
                    for (unsigned h = 0; h < bh; ++h) {
                        const float* src = m->b4csr_data + w4 * plane + (size_t)h * m->b4csr_aligned_size + (size_t)j * 4;
                        float* dst = sval.data() + (((size_t)k * bw4 + w4) * bh + h) * hal * 4 + (size_t)r * 4;
                        std::memcpy(dst, src, sizeof(float) * 4);
                    }
            }
        d->hal = (int)hal;
        d->bslots = w;
        UPLOAD_CHECK(upload_array(d, scol.data(), scol.size(), &d->col));
        UPLOAD_CHECK(upload_array(d, sval.data(), sval.size(), &d->val));
        *out = d;
        return CLSPMV_OK;
    }
    UPLOAD_CHECK(upload_array(d, m->b4csr_row_ptr, (size_t)nbr + 1, &d->rp));
    UPLOAD_CHECK(upload_array(d, m->b4csr_col_id, (size_t)nb, &d->col));
    UPLOAD_CHECK(upload_array(d, m->b4csr_data, (size_t)m->b4csr_aligned_size * bh * (bw / 4), &d->val));
    *out = d;
    return CLSPMV_OK;
}

int clspmv_upload_b4ell(clspmv_ctx* ctx, const clspmv_b4ell_matrix* m, clspmv_dmat** out) {
    UPLOAD_PROLOGUE(m, CLSPMV_FMT_BELL)
    const unsigned bh = m->b4ell_bheight, bw = m->b4ell_bwidth;
    if (!block_shape_ok(bh, bw))
        UPLOAD_CHECK(fail(CLSPMV_ERR_UNSUPPORTED, "b4ell block %ux%u not supported", bh, bw));
    const int nbr = m->b4ell_row_num, ell = m->b4ell_block_num;
    if (nbr < 0 || ell < 0 || (long long)nbr * bh < m->matinfo.height ||
        m->b4ell_height_aligned < nbr || m->b4ell_float4_aligned < 4LL * nbr)
        UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "malformed b4ell matrix"));
    const size_t ncol = (size_t)m->b4ell_height_aligned * ell;
    // Check the columns of real block rows only; padding rows are never read.
    for (int s = 0; s < ell; ++s)
        UPLOAD_CHECK(check_cols(m->b4ell_col_id + (size_t)s * m->b4ell_height_aligned, nbr,
                                std::max(m->matinfo.width, 1), 4, "b4ell"));
    d->info.bheight = (int)bh;
    d->info.bwidth = (int)bw;
    d->brows = nbr;
    d->hal = m->b4ell_height_aligned;
    d->f4 = m->b4ell_float4_aligned;
    d->ellnum = ell;
    d->info.stored = (long long)nbr * ell * bh * bw;
    UPLOAD_CHECK(upload_array(d, m->b4ell_col_id, ncol, &d->col));
    UPLOAD_CHECK(upload_array(d, m->b4ell_data, (size_t)m->b4ell_float4_aligned * bh * (bw / 4) * ell, &d->val));
    *out = d;
    return CLSPMV_OK;
}

int clspmv_upload_sbell(clspmv_ctx* ctx, const clspmv_sbell_matrix* m, clspmv_dmat** out) {
    UPLOAD_PROLOGUE(m, CLSPMV_FMT_SBELL)
    const unsigned bh = m->sbell_bheight, bw = m->sbell_bwidth;
    if (!block_shape_ok(bh, bw))
        UPLOAD_CHECK(fail(CLSPMV_ERR_UNSUPPORTED, "sbell block %ux%u not supported", bh, bw));
    const int nbr = m->sbell_row_num, ns = m->sbell_slice_num, sh = (int)m->sbell_slice_height;
    if (nbr < 0 || ns < 0 || sh <= 0 || (long long)ns * sh < nbr ||
        (long long)nbr * bh < m->matinfo.height || !m->sbell_slice_ptr)
        UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "malformed sbell matrix"));
    for (int s = 0; s < ns; ++s) {
        int w = m->sbell_slice_ptr[s + 1] - m->sbell_slice_ptr[s];
        if (w < 0 || w % sh) UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "sbell slice %d size %d", s, w));
    }
    const size_t total = (size_t)m->sbell_slice_ptr[ns];
    UPLOAD_CHECK(check_cols(m->sbell_col_id, total, std::max(m->matinfo.width, 1), 4, "sbell"));
    d->info.bheight = (int)bh;
    d->info.bwidth = (int)bw;
    d->info.slice_height = sh;
    d->brows = nbr;
    d->info.stored = (long long)total * bh * bw;
    UPLOAD_CHECK(upload_array(d, m->sbell_slice_ptr, (size_t)ns + 1, &d->sptr));
    UPLOAD_CHECK(upload_array(d, m->sbell_col_id, total, &d->col));
    UPLOAD_CHECK(upload_array(d, m->sbell_data, total * bh * bw, &d->val));
    *out = d;
    return CLSPMV_OK;
}


int clspmv_dmat_free(clspmv_dmat* m) {
    if (!m) return CLSPMV_OK;
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != m->device) (void)hipSetDevice(m->device);
    for (auto& b : m->bufs) (void)hipFree(b.first);
    if (cur != m->device && cur >= 0) (void)hipSetDevice(cur);
    delete m;
    return CLSPMV_OK;
}

// ---- mixed-precision block SpMM ------------------------------------------------

namespace {

uint16_t f32_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu))
        return (uint16_t)((u >> 16) | 0x0040u);  // quiet NaN
    u += 0x7FFFu + ((u >> 16) & 1u);            // round to nearest even
    return (uint16_t)(u >> 16);
}

float bf16_to_f32(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

uint16_t f32_to_f16(float f) {
    const _Float16 h = (_Float16)f;  // round to nearest even
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

float f16_to_f32(uint16_t u) {
    _Float16 h;
    std::memcpy(&h, &u, 2);
    return (float)h;
}

}  // namespace

int clspmv_convert_from_f32(const float* src, void* dst, size_t n, int dtype) {
    if ((!src || !dst) && n) return fail(CLSPMV_ERR_INVALID, "null argument");
    uint16_t* d = static_cast<uint16_t*>(dst);
    if (dtype == CLSPMV_DT_BF16) {
        for (size_t i = 0; i < n; ++i) d[i] = f32_to_bf16(src[i]);
    } else if (dtype == CLSPMV_DT_F16) {
        for (size_t i = 0; i < n; ++i) d[i] = f32_to_f16(src[i]);
    } else {
        return fail(CLSPMV_ERR_INVALID, "dtype %d is not a 16-bit type", dtype);
    }
    return CLSPMV_OK;
}

int clspmv_convert_to_f32(const void* src, float* dst, size_t n, int dtype) {
    if ((!src || !dst) && n) return fail(CLSPMV_ERR_INVALID, "null argument");
    const uint16_t* s = static_cast<const uint16_t*>(src);
    if (dtype == CLSPMV_DT_BF16) {
        for (size_t i = 0; i < n; ++i) dst[i] = bf16_to_f32(s[i]);
    } else if (dtype == CLSPMV_DT_F16) {
        for (size_t i = 0; i < n; ++i) dst[i] = f16_to_f32(s[i]);
    } else {
        return fail(CLSPMV_ERR_INVALID, "dtype %d is not a 16-bit type", dtype);
    }
    return CLSPMV_OK;
}

int clspmv_upload_bcsr_lowp(clspmv_ctx* ctx, const clspmv_bcsr_matrix* m, int dtype,
                            clspmv_dmat** out) {
    if (dtype != CLSPMV_DT_BF16 && dtype != CLSPMV_DT_F16)
        return fail(CLSPMV_ERR_INVALID, "dtype %d: need CLSPMV_DT_BF16 or CLSPMV_DT_F16", dtype);
    UPLOAD_PROLOGUE(m, CLSPMV_FMT_BSPMM)
    if (m->bcsr_bwidth != 4 || m->bcsr_bheight != 4)
        UPLOAD_CHECK(fail(CLSPMV_ERR_UNSUPPORTED, "block SpMM needs 4x4 blocks (got %ux%u)",
                          m->bcsr_bheight, m->bcsr_bwidth));
    const int nbr = m->bcsr_row_num, nb = m->bcsr_block_num, rows = m->matinfo.height,
              cols = m->matinfo.width;
    if (nbr < 0 || nb < 0 || (long long)nbr * 4 < rows || m->bcsr_aligned_size < nb ||
        (nbr > 0 && (!m->bcsr_row_ptr || m->bcsr_row_ptr[0] != 0 || m->bcsr_row_ptr[nbr] != nb)))
        UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "malformed BCSR matrix"));
    for (int r = 0; r < nbr; ++r)
        if (m->bcsr_row_ptr[r + 1] < m->bcsr_row_ptr[r])
            UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "BCSR row pointer decreases at %d", r));
    for (int j = 0; j < nb; ++j) {
        const int c = m->bcsr_col_id[j];
        if (c < 0 || c >= cols || c % 4)
            UPLOAD_CHECK(fail(CLSPMV_ERR_INVALID, "BCSR block %d: column %d", j, c));
    }
    // Device layout: slices of kSpmmSlice block rows (one wave's), each
    // padded to its longest block row; block j of a slice's rows are stored
    // together, so a wave's loads of one step are contiguous.  Slot
    // (slice s, step j, row g) holds the block's column (-1 for padding) and
    // its 16 values row by row; element (h, w) of BCSR block b sits at
    // data[b + (4h + w) * aligned] (include/matrix_storage.h:1719-1735).
    constexpr int S = launch::kSpmmSlice;
    const int nsl = (nbr + S - 1) / S;
    std::vector<int> sptr((size_t)nsl + 1, 0);
    for (int sl = 0; sl < nsl; ++sl) {
        int w = 0;
        for (int r = sl * S; r < std::min(nbr, (sl + 1) * S); ++r)
            w = std::max(w, m->bcsr_row_ptr[r + 1] - m->bcsr_row_ptr[r]);
        if ((long long)sptr[sl] + (long long)w * S > INT_MAX / 16)
            UPLOAD_CHECK(fail(CLSPMV_ERR_TOO_LARGE, "block SpMM layout exceeds 32-bit offsets"));
        sptr[sl + 1] = sptr[sl] + w * S;
    }
    const size_t slots = (size_t)sptr[nsl];
    std::vector<int> scol(slots, -1);
    std::vector<float> blocks(slots * 16, 0.0f);
    for (int r = 0; r < nbr; ++r) {
        const int sl = r / S, g = r % S;
        for (int j = 0; j < m->bcsr_row_ptr[r + 1] - m->bcsr_row_ptr[r]; ++j) {
            const int b = m->bcsr_row_ptr[r] + j;
            const size_t at = (size_t)sptr[sl] + (size_t)j * S + g;
            scol[at] = m->bcsr_col_id[b];
            for (int e = 0; e < 16; ++e) blocks[at * 16 + e] = m->bcsr_data[b + (size_t)e * m->bcsr_aligned_size];
        }
    }
    std::vector<uint16_t> lp(blocks.size());
    UPLOAD_CHECK(clspmv_convert_from_f32(blocks.data(), lp.data(), lp.size(), dtype));
    d->dtype = dtype;
    d->brows = nbr;
    for (int r = 0; r < nbr && !d->long_brows; ++r)
        d->long_brows = (long long)(m->bcsr_row_ptr[r + 1] - m->bcsr_row_ptr[r]) * 4 > launch::kLongRowF64;
    d->info.bheight = d->info.bwidth = 4;
    d->info.stored = (long long)slots * 16;
    UPLOAD_CHECK(upload_array(d, sptr.data(), sptr.size(), &d->rp));
    UPLOAD_CHECK(upload_array(d, scol.data(), scol.size(), &d->col));
    UPLOAD_CHECK(upload_array(d, lp.data(), lp.size(), &d->lpval));
    // X windows for the tiled kernel: every slice's distinct block columns
    // (as tile indices) and each slot's index among them, the lists at a
    // fixed stride (the longest list's length) so that a wave finds its
    // slice's list without a pointer load; shorter lists repeat their last
    // tile.  Not built when a slice has more than kSpmmWinMax of them, when
    // the padding would more than double the tiles staged, nor for
    // FP64-carried long rows.
    if (!d->long_brows && nsl > 0) {
        std::vector<std::vector<int>> lists((size_t)nsl);
        std::vector<unsigned char> widx(slots, 0xFF);
        size_t wmax = 0, total = 0;
        bool ok = true;
        for (int sl = 0; sl < nsl && ok; ++sl) {
            std::vector<int>& cur = lists[sl];
            for (int at = sptr[sl]; at < sptr[sl + 1]; ++at)
                if (scol[at] >= 0) cur.push_back(scol[at] / 4);
            std::sort(cur.begin(), cur.end());
            cur.erase(std::unique(cur.begin(), cur.end()), cur.end());
            if (cur.size() > (size_t)launch::kSpmmWinMax) ok = false;
            for (int at = sptr[sl]; at < sptr[sl + 1] && ok; ++at)
                if (scol[at] >= 0)
                    widx[at] = (unsigned char)(std::lower_bound(cur.begin(), cur.end(), scol[at] / 4) - cur.begin());
            wmax = std::max(wmax, cur.size());
            total += cur.size();
        }
        if (ok && wmax > 0 && (size_t)nsl * wmax <= 2 * total + 64) {
            std::vector<int> wtile((size_t)nsl * wmax);
            for (int sl = 0; sl < nsl; ++sl) {
                const std::vector<int>& cur = lists[sl];
                for (size_t i = 0; i < wmax; ++i)
                    wtile[(size_t)sl * wmax + i] = cur.empty() ? 0 : cur[std::min(i, cur.size() - 1)];
            }
            for (unsigned char& w : widx)
                if (w == 0xFF) w = (unsigned char)(wmax + 1);  // the zero tile after the spare slot
            UPLOAD_CHECK(upload_array(d, wtile.data(), wtile.size(), &d->wtile));
            UPLOAD_CHECK(upload_array(d, widx.data(), widx.size(), &d->widx));
            d->win_max = (int)wmax;
        }
    }
    *out = d;
    return CLSPMV_OK;
}

#undef UPLOAD_PROLOGUE
#undef UPLOAD_CHECK

int clspmv_spmm(clspmv_ctx* ctx, const clspmv_dmat* A, int k, const void* d_X, int ldx,
                float* d_Y, int ldy) {
    if (!ctx || !A || (!d_X && k) || (!d_Y && k)) return fail(CLSPMV_ERR_INVALID, "null argument");
    const clspmv_dmat_info& I = A->info;
    if (I.format != CLSPMV_FMT_BSPMM)
        return fail(CLSPMV_ERR_UNSUPPORTED, "clspmv_spmm needs a matrix from clspmv_upload_bcsr_lowp");
    if (k < 0) return fail(CLSPMV_ERR_INVALID, "k = %d", k);
    if (ldx < I.cols || ldx % 4 || ldy < I.rows || ldy % 4)
        return fail(CLSPMV_ERR_INVALID, "ldx %d / ldy %d: need >= cols / rows and multiples of 4", ldx, ldy);
    if (((uintptr_t)d_X & 7) || ((uintptr_t)d_Y & 15))
        return fail(CLSPMV_ERR_INVALID, "X must be 8-byte and Y 16-byte aligned");
    int st = use_ctx_device(ctx);
    if (st) return st;
    if (k == 0 || I.rows == 0) return CLSPMV_OK;
    hipError_t e = launch::bspmm4(ctx->stream, A->dtype == CLSPMV_DT_BF16, A->long_brows, A->brows, I.rows, I.cols, k,
                                  A->rp, A->col, A->lpval, static_cast<const unsigned short*>(d_X), ldx,
                                  d_Y, ldy, /*tiled=*/false, /*accumulate=*/true,
                                  launch::SpmmWindows{A->wtile, A->widx, A->win_max});
    if (e != hipSuccess) return fail(CLSPMV_ERR_HIP, "spmm launch failed: %s", hipGetErrorString(e));
    return CLSPMV_OK;
}

int clspmv_spmm_tiled(clspmv_ctx* ctx, const clspmv_dmat* A, int k, const void* d_X4, float* d_Y, int ldy) {
    return clspmv_spmm_tiled_ex(ctx, A, k, d_X4, d_Y, ldy, 1);
}

int clspmv_spmm_tiled_ex(clspmv_ctx* ctx, const clspmv_dmat* A, int k, const void* d_X4, float* d_Y, int ldy,
                         int accumulate) {
    if (!ctx || !A || (!d_X4 && k) || (!d_Y && k)) return fail(CLSPMV_ERR_INVALID, "null argument");
    const clspmv_dmat_info& I = A->info;
    if (I.format != CLSPMV_FMT_BSPMM)
        return fail(CLSPMV_ERR_UNSUPPORTED, "clspmv_spmm_tiled needs a matrix from clspmv_upload_bcsr_lowp");
    if (k < 0) return fail(CLSPMV_ERR_INVALID, "k = %d", k);
    if (ldy < I.rows || ldy % 4) return fail(CLSPMV_ERR_INVALID, "ldy %d: need >= rows and a multiple of 4", ldy);
    if (((uintptr_t)d_X4 & 7) || ((uintptr_t)d_Y & 15))
        return fail(CLSPMV_ERR_INVALID, "X must be 8-byte and Y 16-byte aligned");
    int st = use_ctx_device(ctx);
    if (st) return st;
    if (k == 0 || I.rows == 0) return CLSPMV_OK;
    hipError_t e = launch::bspmm4(ctx->stream, A->dtype == CLSPMV_DT_BF16, A->long_brows, A->brows, I.rows, I.cols, k,
                                  A->rp, A->col, A->lpval, static_cast<const unsigned short*>(d_X4), k, d_Y, ldy,
                                  true, accumulate != 0,
                                  launch::SpmmWindows{A->wtile, A->widx, A->win_max});
    if (e != hipSuccess) return fail(CLSPMV_ERR_HIP, "spmm launch failed: %s", hipGetErrorString(e));
    return CLSPMV_OK;
}

int clspmv_spmm_pack_x(clspmv_ctx* ctx, int k, int cols, const void* d_X, int ldx, void* d_X4) {
    if (!ctx || (k > 0 && cols > 0 && (!d_X || !d_X4))) return fail(CLSPMV_ERR_INVALID, "null argument");
    if (k < 0 || cols < 0) return fail(CLSPMV_ERR_INVALID, "k = %d, cols = %d", k, cols);
    if (ldx < cols || ldx % 4) return fail(CLSPMV_ERR_INVALID, "ldx %d: need >= cols and a multiple of 4", ldx);
    if (((uintptr_t)d_X & 7) || ((uintptr_t)d_X4 & 7)) return fail(CLSPMV_ERR_INVALID, "X and X4 must be 8-byte aligned");
    int st = use_ctx_device(ctx);
    if (st) return st;
    const hipError_t e = launch::pack_x4(ctx->stream, k, cols, static_cast<const unsigned short*>(d_X), ldx,
                                         static_cast<unsigned short*>(d_X4));
    if (e != hipSuccess) return fail(CLSPMV_ERR_HIP, "pack launch failed: %s", hipGetErrorString(e));
    return CLSPMV_OK;
}

int clspmv_spmv_multi(clspmv_ctx* ctx, const clspmv_dmat* A, int k, const float* d_X, float* d_Y,
                      int ldy) {
    if (!ctx || !A || (k > 0 && (!d_X || !d_Y))) return fail(CLSPMV_ERR_INVALID, "null argument");
    const clspmv_dmat_info& I = A->info;
    if (k < 0 || k > launch::kMultiMax) return fail(CLSPMV_ERR_INVALID, "k = %d outside 0..%d", k, launch::kMultiMax);
    if (ldy < I.rows) return fail(CLSPMV_ERR_INVALID, "ldy %d < rows %d", ldy, I.rows);
    int st = use_ctx_device(ctx);
    if (st) return st;
    if (k == 0 || I.rows == 0 || I.cols == 0 || I.nnz == 0) return CLSPMV_OK;
    return multi_launch(ctx->stream, A, k, d_X, d_Y, ldy);
}

int clspmv_dmat_get_info(const clspmv_dmat* m, clspmv_dmat_info* out) {
    if (!m || !out) return fail(CLSPMV_ERR_INVALID, "null argument");
    *out = m->info;
    out->panel_width = m->pw;
    out->panel_active = m->use_panel ? 1 : 0;
    out->panel_groups = m->pw ? m->pnwg : 0;
    out->panel_subchunks = m->pw ? m->pnsc : 0;
    out->panel_blocks = m->pw ? m->pnb : 0;
    out->plain = m->plain ? 1 : 0;
    out->dtype = m->dtype;
    out->block_slots = m->bslots;
    out->spmm_window = m->wtile ? m->win_max : 0;
    return CLSPMV_OK;
}

int clspmv_dmat_set_panel(clspmv_dmat* m, int use_panel) {
    if (!m) return fail(CLSPMV_ERR_INVALID, "null argument");
    if (use_panel && !m->pw) return fail(CLSPMV_ERR_UNSUPPORTED, "matrix has no panel encoding");
    if (!use_panel && !m->plain) return fail(CLSPMV_ERR_UNSUPPORTED, "matrix has only the panel encoding");
    m->use_panel = use_panel != 0;
    m->serve_multi = -1;
    return CLSPMV_OK;
}

int clspmv_spmv(clspmv_ctx* ctx, const clspmv_dmat* A, int method,
                const float* d_x, float* d_y) {
    if (!ctx || !A) return fail(CLSPMV_ERR_INVALID, "null argument");
    if (A->info.nnz > 0 && A->info.rows > 0 && (!d_x || !d_y))
        return fail(CLSPMV_ERR_INVALID, "null vector");
    if (A->device != ctx->device)
        return fail(CLSPMV_ERR_INVALID, "matrix lives on device %d, context on %d",
                    A->device, ctx->device);
    int st = use_ctx_device(ctx);
    if (st) return st;
    return dmat_launch(ctx->stream, A, method, d_x, d_y);
}

int clspmv_time_spmv(clspmv_ctx* ctx, const clspmv_dmat* A, int method,
                     const float* d_x, float* d_y, int warmup, int ntimes,
                     double* ms_per_call) {
    if (!ctx || !A || ntimes <= 0) return fail(CLSPMV_ERR_INVALID, "bad argument");
    for (int i = 0; i < warmup; ++i) {
        int st = clspmv_spmv(ctx, A, method, d_x, d_y);
        if (st) return st;
    }
    CLSPMV_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    for (int i = 0; i < ntimes; ++i) {
        int st = clspmv_spmv(ctx, A, method, d_x, d_y);
        if (st) return st;
    }
    CLSPMV_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    CLSPMV_HIP(hipEventSynchronize(ctx->ev_stop));
    float ms = 0.0f;
    CLSPMV_HIP(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (ms_per_call) *ms_per_call = (double)ms / ntimes;
    return CLSPMV_OK;
}

int clspmv_time_spmv_multi(clspmv_ctx* ctx, const clspmv_dmat* A, int k, const float* d_X, float* d_Y,
                           int ldy, int warmup, int ntimes, double* ms_per_call) {
    if (!ctx || !A || ntimes <= 0 || warmup < 0) return fail(CLSPMV_ERR_INVALID, "bad argument");
    for (int i = 0; i < warmup; ++i) {
        int st = clspmv_spmv_multi(ctx, A, k, d_X, d_Y, ldy);
        if (st) return st;
    }
    CLSPMV_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    for (int i = 0; i < ntimes; ++i) {
        int st = clspmv_spmv_multi(ctx, A, k, d_X, d_Y, ldy);
        if (st) return st;
    }
    CLSPMV_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    CLSPMV_HIP(hipEventSynchronize(ctx->ev_stop));
    float ms = 0.0f;
    CLSPMV_HIP(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    if (ms_per_call) *ms_per_call = (double)ms / ntimes;
    return CLSPMV_OK;
}

namespace {
// (float4 per thread per step, work-groups) of the probes' launch shapes.
constexpr int kProbeShapes = 5;
constexpr int kProbeShape[kProbeShapes][2] = {{2, 4096}, {4, 2048}, {4, 4096}, {8, 1024}, {8, 2048}};
}  // namespace

int clspmv_read_probe(clspmv_ctx* ctx, size_t bytes, int ntimes, double* ms_per_read, double* gbps) {
    if (!ctx || ntimes <= 0 || bytes < 16) return fail(CLSPMV_ERR_INVALID, "bad argument");
    int st = use_ctx_device(ctx);
    if (st) return st;
    const long long n4 = (long long)(bytes / 16);
    float *a = nullptr, *out = nullptr;
    CLSPMV_HIP(hipMalloc(&a, n4 * 16));
    hipError_t e = hipMalloc(&out, sizeof(float) * 4096 * 256);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return fail(CLSPMV_ERR_ALLOC, "probe buffer: %s", hipGetErrorString(e));
    }
    e = hipMemsetAsync(a, 0, n4 * 16, ctx->stream);
    double per = 1e30;
    for (int v = 0; v < kProbeShapes && e == hipSuccess; ++v) {
        auto run = [&]() { return launch::read4u(ctx->stream, n4, a, out, kProbeShape[v][0], kProbeShape[v][1]); };
        float ms = 0.0f;
        e = run();
        if (e == hipSuccess) e = hipEventRecord(ctx->ev_start, ctx->stream);
        for (int i = 0; i < ntimes && e == hipSuccess; ++i) e = run();
        if (e == hipSuccess) e = hipEventRecord(ctx->ev_stop, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(ctx->ev_stop);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
        if (e == hipSuccess) per = std::min(per, (double)ms / ntimes);
    }
    (void)hipFree(a);
    (void)hipFree(out);
    if (e != hipSuccess) return fail(CLSPMV_ERR_HIP, "probe failed: %s", hipGetErrorString(e));
    if (ms_per_read) *ms_per_read = per;
    if (gbps) *gbps = per > 0 ? (double)n4 * 16 / (per * 1e-3) / 1e9 : 0.0;
    return CLSPMV_OK;
}

int clspmv_bandwidth_probe(clspmv_ctx* ctx, size_t bytes, int ntimes,
                           double* ms_per_copy, double* gbps) {
    if (!ctx || ntimes <= 0 || bytes < 16) return fail(CLSPMV_ERR_INVALID, "bad argument");
    int st = use_ctx_device(ctx);
    if (st) return st;
    const long long n4 = (long long)(bytes / 16);
    float *a = nullptr, *b = nullptr;
    CLSPMV_HIP(hipMalloc(&a, n4 * 16));
    hipError_t e = hipMalloc(&b, n4 * 16);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return fail(CLSPMV_ERR_ALLOC, "probe buffer: %s", hipGetErrorString(e));
    }
    e = hipMemsetAsync(a, 0, n4 * 16, ctx->stream);
    // The plain one-float4-per-thread copy, then unrolled copies with
    // non-temporal stores over a few launch shapes; the fastest counts.
    double per = 1e30;
    for (int v = 0; v < kProbeShapes + 1 && e == hipSuccess; ++v) {
        auto run = [&]() {
            return v == 0 ? launch::copy4(ctx->stream, n4, a, b)
                          : launch::copy4u(ctx->stream, n4, a, b, kProbeShape[v - 1][0], kProbeShape[v - 1][1]);
        };
        float ms = 0.0f;
        e = run();
        if (e == hipSuccess) e = hipEventRecord(ctx->ev_start, ctx->stream);
        for (int i = 0; i < ntimes && e == hipSuccess; ++i) e = run();
        if (e == hipSuccess) e = hipEventRecord(ctx->ev_stop, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(ctx->ev_stop);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
        if (e == hipSuccess) per = std::min(per, (double)ms / ntimes);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(CLSPMV_ERR_HIP, "probe failed: %s", hipGetErrorString(e));
    if (ms_per_copy) *ms_per_copy = per;
    if (gbps) *gbps = per > 0 ? 2.0 * n4 * 16 / (per * 1e-3) / 1e9 : 0.0;
    return CLSPMV_OK;
}

int clspmv_gather_probe(clspmv_ctx* ctx, int nx, long long n, int ntimes,
                        double* ms_per_call, double* ggathers_per_s) {
    if (!ctx || nx <= 0 || n <= 0 || ntimes <= 0) return fail(CLSPMV_ERR_INVALID, "bad argument");
    int st = use_ctx_device(ctx);
    if (st) return st;
    float *x = nullptr, *y = nullptr;
    CLSPMV_HIP(hipMalloc(&x, sizeof(float) * nx));
    hipError_t e = hipMalloc(&y, sizeof(float) * n);
    if (e != hipSuccess) {
        (void)hipFree(x);
        return fail(CLSPMV_ERR_ALLOC, "probe buffer: %s", hipGetErrorString(e));
    }
    float ms = 0.0f;
    e = hipMemsetAsync(x, 0, sizeof(float) * nx, ctx->stream);
    if (e == hipSuccess) e = launch::gather(ctx->stream, n, nx, x, y);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev_start, ctx->stream);
    for (int i = 0; i < ntimes && e == hipSuccess; ++i) e = launch::gather(ctx->stream, n, nx, x, y);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev_stop, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev_stop);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    (void)hipFree(x);
    (void)hipFree(y);
    if (e != hipSuccess) return fail(CLSPMV_ERR_HIP, "gather probe failed: %s", hipGetErrorString(e));
    const double per = (double)ms / ntimes;
    if (ms_per_call) *ms_per_call = per;
    if (ggathers_per_s) *ggathers_per_s = per > 0 ? (double)n / (per * 1e-3) / 1e9 : 0.0;
    return CLSPMV_OK;
}

}  // extern "C"
