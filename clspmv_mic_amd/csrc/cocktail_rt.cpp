// Cocktail device runtime (src/spmv_cocktail.cpp): uploads every part of a
// partitioned matrix, launches them into one y in the reference's order
// (BDIA, DIA, SBELL, BELL, BCSR, SELL, ELL, CSR, COO; do_spmv :1088-1132),
// and exposes the reference's copy helpers.
//
// The pass over all parts is captured once into a HIP graph, so one SpMV is
// a single graph launch instead of up to ten kernel launches (the reference
// pays a full launch per part, 0.27 ms each on its device, runlog.txt:65).
// Parts accumulate into y one after another on one stream, which fixes the
// summation order of every row: results do not depend on timing.  The panel
// encoding's first pass (CSR/COO parts) reads only x and writes the part's
// own segment sums, so it runs on a second stream beside the earlier parts'
// kernels (the DIA band of a power-law matrix, say); the second pass, which
// adds into y, waits for it and keeps its place in the order.  Or (fuse) an
// FP32 DIA part directly ahead of such a part rides along in that part's
// first-pass launch as extra work-groups, which fill the first pass's
// ragged end; its update still reaches y before the second pass's.

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>

#include "clspmv_cocktail.h"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace clspmv;

struct clspmv_cocktail_gpu {
    int device = 0;
    int width = 0, height = 0, nnz = 0;
    float* vec = nullptr;  // x, padded to a multiple of 8 floats with zeros
    float* res = nullptr;  // y
    int vec_len = 0;
    std::vector<clspmv_dmat*> parts;  // launch order
    std::vector<int> methods;
};

struct clspmv_cocktail_kernels {
    clspmv_ctx* ctx = nullptr;
    clspmv_cocktail_gpu* gpu = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // kGraphPasses passes in one graph: do_spmv(ntimes) launches it while
    // ntimes allows, so a pass of a small matrix is not held to the host's
    // launch rate.
    hipGraph_t graph_many = nullptr;
    hipGraphExec_t exec_many = nullptr;
    bool use_graph = true;  // false: the pass is issued as plain launches
    // Panel first passes on `side`, forked from and joined back into the
    // launching stream by the two events; false: one stream throughout.
    bool overlap = true;
    bool fuse = false;  // DIA part summed in the next part's first-pass launch
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

namespace {

// Whether `method` can run on `A` (LDS windows and 128-bit paths have shape
// requirements); the partitioner's choice comes from the tables and may
// name a variant the part's shape does not allow.
bool method_runs(const clspmv_dmat* A, int m) {
    const int span = A->max_off - A->min_off;
    switch (A->info.format) {
        case CLSPMV_FMT_ELL: return m == 0 || A->hal % 4 == 0;
        case CLSPMV_FMT_DIA:
            if (m == 1 && A->len % 2) return false;
            return m < 2 || (m == 3 ? 1024 : 256) + span <= 8192;
        case CLSPMV_FMT_BDIA: {
            static const int rpt[5] = {1, 1, 2, 2, 4};
            static const bool lds[5] = {false, true, true, false, true};
            if (m == 3 && A->len % 2) return false;
            return !lds[m] || rpt[m] * 256 + A->max_band <= 4096;
        }
        default: return true;
    }
}

// Index of the first part with a first pass, if an earlier part can run
// beside it; -1 when overlapping gains nothing.
int overlap_from(const clspmv_cocktail_gpu* g) {
    for (size_t i = 0; i < g->parts.size(); ++i)
        if (dmat_has_prepass(g->parts[i])) return i > 0 ? (int)i : -1;
    return -1;
}

// Index of an FP32 DIA part directly followed by a part with a first pass,
// or -1.
int fuse_at(const clspmv_cocktail_gpu* g) {
    for (size_t i = 0; i + 1 < g->parts.size(); ++i)
        if (dmat_dia_tail_ok(g->parts[i]) && dmat_has_prepass(g->parts[i + 1])) return (int)i;
    return -1;
}

bool fuses(const clspmv_cocktail_kernels* k) { return k->fuse && fuse_at(k->gpu) >= 0; }
bool overlaps(const clspmv_cocktail_kernels* k) {
    return !fuses(k) && k->overlap && k->side && overlap_from(k->gpu) >= 0;
}

// One pass over every part into y, on stream s (and k->side).
int launch_all(hipStream_t s, const clspmv_cocktail_kernels* k) {
    const clspmv_cocktail_gpu* g = k->gpu;
    if (fuses(k)) {
        const int d = fuse_at(g);
        for (size_t i = 0; i < g->parts.size(); ++i) {
            if ((int)i == d) continue;  // summed by the next part's first-pass launch
            int st = (int)i == d + 1 ? dmat_prepass(s, g->parts[i], g->vec, g->parts[d], g->res)
                                     : dmat_launch(s, g->parts[i], g->methods[i], g->vec, g->res);
            if (!st && (int)i == d + 1) st = dmat_launch_rest(s, g->parts[i], g->methods[i], g->vec, g->res);
            if (st) return st;
        }
        return CLSPMV_OK;
    }
    const int first = overlaps(k) ? overlap_from(g) : -1;
    if (first < 0) {
        for (size_t i = 0; i < g->parts.size(); ++i) {
            int st = dmat_launch(s, g->parts[i], g->methods[i], g->vec, g->res);
            if (st) return st;
        }
        return CLSPMV_OK;
    }
    // Fork: the side stream starts after everything already on s (x is in
    // place, and the previous pass's second passes have read the segment
    // sums these first passes overwrite).
    CLSPMV_HIP(hipEventRecord(k->fork, s));
    CLSPMV_HIP(hipStreamWaitEvent(k->side, k->fork, 0));
    for (size_t i = first; i < g->parts.size(); ++i) {
        int st = dmat_prepass(k->side, g->parts[i], g->vec);
        if (st) return st;
    }
    CLSPMV_HIP(hipEventRecord(k->join, k->side));
    for (size_t i = 0; i < g->parts.size(); ++i) {
        const clspmv_dmat* A = g->parts[i];
        if ((int)i == first) CLSPMV_HIP(hipStreamWaitEvent(s, k->join, 0));  // join
        int st = (int)i >= first ? dmat_launch_rest(s, A, g->methods[i], g->vec, g->res)
                                 : dmat_launch(s, A, g->methods[i], g->vec, g->res);
        if (st) return st;
    }
    return CLSPMV_OK;
}

constexpr int kGraphPasses = 16;

// Captures `passes` consecutive passes into *graph / *exec.
int capture_passes(clspmv_cocktail_kernels* k, int passes, hipGraph_t* graph, hipGraphExec_t* exec) {
    hipStream_t s = k->ctx->stream;
    CLSPMV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = CLSPMV_OK;
    for (int i = 0; i < passes && !st; ++i) st = launch_all(s, k);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(CLSPMV_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    *graph = g;
    CLSPMV_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return CLSPMV_OK;
}

void drop_graphs(clspmv_cocktail_kernels* k) {
    if (k->exec) (void)hipGraphExecDestroy(k->exec);
    if (k->graph) (void)hipGraphDestroy(k->graph);
    if (k->exec_many) (void)hipGraphExecDestroy(k->exec_many);
    if (k->graph_many) (void)hipGraphDestroy(k->graph_many);
    k->exec = k->exec_many = nullptr;
    k->graph = k->graph_many = nullptr;
}

// One pass (exec) and kGraphPasses passes (exec_many) as graphs.  On an
// error neither is left, and do_spmv refuses to run.
int capture(clspmv_cocktail_kernels* k) {
    drop_graphs(k);
    int st = capture_passes(k, 1, &k->graph, &k->exec);
    if (!st) st = capture_passes(k, kGraphPasses, &k->graph_many, &k->exec_many);
    if (st) drop_graphs(k);
    return st;
}

template <class F>
int add_part(clspmv_cocktail_gpu* g, clspmv_ctx* ctx, bool use, int meth, F&& upload) {
    if (!use) return CLSPMV_OK;
    clspmv_dmat* A = nullptr;
    int st = upload(ctx, &A);
    if (st) return st;
    if (A->info.nnz == 0) {  // nothing to launch
        clspmv_dmat_free(A);
        return CLSPMV_OK;
    }
    int m = meth;
    if (m < 0 || m >= clspmv_method_count(A->info.format) || !method_runs(A, m)) m = auto_method(A);
    g->parts.push_back(A);
    g->methods.push_back(m);
    return CLSPMV_OK;
}

void destroy_gpu(clspmv_cocktail_gpu* g) {
    if (!g) return;
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != g->device) (void)hipSetDevice(g->device);
    for (clspmv_dmat* p : g->parts) clspmv_dmat_free(p);
    if (g->vec) (void)hipFree(g->vec);
    if (g->res) (void)hipFree(g->res);
    delete g;
}

void destroy_kernels(clspmv_cocktail_kernels* k) {
    if (!k) return;
    drop_graphs(k);
    if (k->fork) (void)hipEventDestroy(k->fork);
    if (k->join) (void)hipEventDestroy(k->join);
    if (k->side) (void)hipStreamDestroy(k->side);
    clspmv_ctx_destroy(k->ctx);
    delete k;
}

int sync(clspmv_cocktail_kernels* k) {
    CLSPMV_HIP(hipStreamSynchronize(k->ctx->stream));
    return CLSPMV_OK;
}

}  // namespace

int clspmv_init_mat_kernels_on(int device, const clspmv_cocktail* m, const float* vec,
                               const float* res, clspmv_cocktail_gpu** gpumat,
                               clspmv_cocktail_kernels** kernels) {
    if (!m || !gpumat || !kernels) return fail(CLSPMV_ERR_INVALID, "null argument");
    *gpumat = nullptr;
    *kernels = nullptr;
    clspmv_cocktail_kernels* k = new clspmv_cocktail_kernels();
    int st = clspmv_ctx_create(device, &k->ctx);
    if (st) {
        delete k;
        return st;
    }
    hipError_t e = hipStreamCreateWithFlags(&k->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&k->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&k->join, hipEventDisableTiming);
    if (e != hipSuccess) {
        destroy_kernels(k);
        return fail(CLSPMV_ERR_HIP, "cocktail streams: %s", hipGetErrorString(e));
    }
    clspmv_cocktail_gpu* g = new clspmv_cocktail_gpu();
    g->device = device;
    g->width = m->mat_width;
    g->height = m->mat_height;
    g->nnz = m->mat_nnz;
    k->gpu = g;
    auto cleanup = [&](int code) {
        destroy_gpu(g);
        k->gpu = nullptr;
        destroy_kernels(k);
        return code;
    };
    // x padded to a multiple of 8 (cpy_shared, src/spmv_cocktail.cpp:230).
    g->vec_len = round_up(std::max(g->width, 1), 8);
    e = hipMalloc(&g->vec, sizeof(float) * g->vec_len);
    if (e == hipSuccess) e = hipMalloc(&g->res, sizeof(float) * std::max(g->height, 1));
    if (e == hipSuccess) e = hipMemsetAsync(g->vec, 0, sizeof(float) * g->vec_len, k->ctx->stream);
    if (e == hipSuccess && vec && g->width)
        e = hipMemcpyAsync(g->vec, vec, sizeof(float) * g->width, hipMemcpyHostToDevice, k->ctx->stream);
    if (e == hipSuccess && res && g->height)
        e = hipMemcpyAsync(g->res, res, sizeof(float) * g->height, hipMemcpyHostToDevice, k->ctx->stream);
    else if (e == hipSuccess && g->height)
        e = hipMemsetAsync(g->res, 0, sizeof(float) * g->height, k->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(k->ctx->stream);
    if (e != hipSuccess)
        return cleanup(fail(CLSPMV_ERR_ALLOC, "cocktail vectors: %s", hipGetErrorString(e)));

    clspmv_ctx* c = k->ctx;
    if ((st = add_part(g, c, m->ifusebdia, m->bdia_meth_num,
                       [&](clspmv_ctx* x, clspmv_dmat** A) { return clspmv_upload_bdia(x, &m->bdia, A); })) ||
        (st = add_part(g, c, m->ifusedia, m->dia_meth_num,
                       [&](clspmv_ctx* x, clspmv_dmat** A) { return clspmv_upload_dia(x, &m->dia, A); })) ||
        (st = add_part(g, c, m->ifusesbell, m->sbell_meth_num,
                       [&](clspmv_ctx* x, clspmv_dmat** A) { return clspmv_upload_sbell(x, &m->sbell, A); })) ||
        (st = add_part(g, c, m->ifusebell, m->bell_meth_num,
                       [&](clspmv_ctx* x, clspmv_dmat** A) { return clspmv_upload_b4ell(x, &m->bell, A); })) ||
        (st = add_part(g, c, m->ifusebcsr, m->bcsr_meth_num,
                       [&](clspmv_ctx* x, clspmv_dmat** A) { return clspmv_upload_b4csr(x, &m->bcsr, A); })) ||
        (st = add_part(g, c, m->ifusesell, m->sell_meth_num,
                       [&](clspmv_ctx* x, clspmv_dmat** A) { return clspmv_upload_sell(x, &m->sell, A); })) ||
        (st = add_part(g, c, m->ifuseell, m->ell_meth_num,
                       [&](clspmv_ctx* x, clspmv_dmat** A) { return clspmv_upload_ell(x, &m->ell, A); })) ||
        (st = add_part(g, c, m->ifusecsr, m->csr_meth_num,
                       [&](clspmv_ctx* x, clspmv_dmat** A) {
                           const clspmv_upload_opts o{CLSPMV_PANEL_BOTH, CLSPMV_DT_F32};
                           return clspmv_upload_csr_ex(x, &m->csr, &o, A);
                       })) ||
        (st = add_part(g, c, m->ifusecoo, m->coo_meth_num,
                       [&](clspmv_ctx* x, clspmv_dmat** A) {
                           const clspmv_upload_opts o{CLSPMV_PANEL_BOTH, CLSPMV_DT_F32};
                           return clspmv_upload_coo_ex(x, &m->coo, &o, A);
                       })))
        return cleanup(st);
    if ((st = capture(k))) return cleanup(st);
    *gpumat = g;
    *kernels = k;
    return CLSPMV_OK;
}

int clspmv_init_mat_kernels(const clspmv_cocktail* cpumat, const float* vec, const float* res,
                            clspmv_cocktail_gpu** gpumat, clspmv_cocktail_kernels** kernels,
                            int use_tex) {
    (void)use_tex;  // no texture path on CDNA
    return clspmv_init_mat_kernels_on(0, cpumat, vec, res, gpumat, kernels);
}

int clspmv_do_spmv_async(clspmv_cocktail_kernels* k, int ntimes) {
    if (!k || !k->exec) return fail(CLSPMV_ERR_INVALID, "kernels not initialised");
    int st = use_ctx_device(k->ctx);
    if (st) return st;
    if (!k->use_graph) {
        for (int i = 0; i < ntimes; ++i)
            if ((st = launch_all(k->ctx->stream, k))) return st;
        return CLSPMV_OK;
    }
    int i = 0;
    for (; i + kGraphPasses <= ntimes; i += kGraphPasses) CLSPMV_HIP(hipGraphLaunch(k->exec_many, k->ctx->stream));
    for (; i < ntimes; ++i) CLSPMV_HIP(hipGraphLaunch(k->exec, k->ctx->stream));
    return CLSPMV_OK;
}

int clspmv_do_spmv(clspmv_cocktail_kernels* k, int ntimes) {
    int st = clspmv_do_spmv_async(k, ntimes);
    if (st) return st;
    return sync(k);
}

namespace {
int check_pair(clspmv_cocktail_gpu* g, clspmv_cocktail_kernels* k) {
    if (!g || !k || k->gpu != g) return fail(CLSPMV_ERR_INVALID, "gpumat/kernels mismatch");
    return use_ctx_device(k->ctx);
}

int copy(clspmv_cocktail_kernels* k, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    if (bytes) CLSPMV_HIP(hipMemcpyAsync(dst, src, bytes, kind, k->ctx->stream));
    return sync(k);
}
}  // namespace

int clspmv_cpy_vector_from_cpu(clspmv_cocktail_gpu* g, clspmv_cocktail_kernels* k, const float* vec, int n) {
    int st = check_pair(g, k);
    if (st) return st;
    if (n < 0 || n > g->width) return fail(CLSPMV_ERR_INVALID, "vector size %d > width %d", n, g->width);
    return copy(k, g->vec, vec, sizeof(float) * n, hipMemcpyHostToDevice);
}

int clspmv_cpy_vector_from_gpu(clspmv_cocktail_gpu* g, clspmv_cocktail_kernels* k, const float* d_vec, int n) {
    int st = check_pair(g, k);
    if (st) return st;
    if (n < 0 || n > g->width) return fail(CLSPMV_ERR_INVALID, "vector size %d > width %d", n, g->width);
    return copy(k, g->vec, d_vec, sizeof(float) * n, hipMemcpyDeviceToDevice);
}

int clspmv_cpy_result_from_cpu(clspmv_cocktail_gpu* g, clspmv_cocktail_kernels* k, const float* res, int n) {
    int st = check_pair(g, k);
    if (st) return st;
    if (n < 0 || n > g->height) return fail(CLSPMV_ERR_INVALID, "result size %d > height %d", n, g->height);
    return copy(k, g->res, res, sizeof(float) * n, hipMemcpyHostToDevice);
}

int clspmv_cpy_result_from_gpu(clspmv_cocktail_gpu* g, clspmv_cocktail_kernels* k, const float* d_res, int n) {
    int st = check_pair(g, k);
    if (st) return st;
    if (n < 0 || n > g->height) return fail(CLSPMV_ERR_INVALID, "result size %d > height %d", n, g->height);
    return copy(k, g->res, d_res, sizeof(float) * n, hipMemcpyDeviceToDevice);
}

int clspmv_cpy_result_to_cpu(clspmv_cocktail_gpu* g, clspmv_cocktail_kernels* k, float* res, int n) {
    int st = check_pair(g, k);
    if (st) return st;
    if (n < 0 || n > g->height) return fail(CLSPMV_ERR_INVALID, "result size %d > height %d", n, g->height);
    return copy(k, res, g->res, sizeof(float) * n, hipMemcpyDeviceToHost);
}

int clspmv_cpy_result_to_gpu(clspmv_cocktail_gpu* g, clspmv_cocktail_kernels* k, float* d_res, int n) {
    int st = check_pair(g, k);
    if (st) return st;
    if (n < 0 || n > g->height) return fail(CLSPMV_ERR_INVALID, "result size %d > height %d", n, g->height);
    return copy(k, d_res, g->res, sizeof(float) * n, hipMemcpyDeviceToDevice);
}

int clspmv_free_mat_kernels(clspmv_cocktail_gpu* g, clspmv_cocktail_kernels* k) {
    if (k && k->ctx) {
        use_ctx_device(k->ctx);
        (void)hipStreamSynchronize(k->ctx->stream);
        if (k->side) (void)hipStreamSynchronize(k->side);
    }
    destroy_gpu(g);
    if (k) k->gpu = nullptr;
    destroy_kernels(k);
    return CLSPMV_OK;
}

float* clspmv_cocktail_gpu_vec(clspmv_cocktail_gpu* g) { return g ? g->vec : nullptr; }
float* clspmv_cocktail_gpu_res(clspmv_cocktail_gpu* g) { return g ? g->res : nullptr; }
clspmv_ctx* clspmv_cocktail_kernels_ctx(clspmv_cocktail_kernels* k) { return k ? k->ctx : nullptr; }

int clspmv_cocktail_profile(clspmv_cocktail_kernels* k, int ntimes, clspmv_part_timing* parts,
                            int max_parts, int* nparts, double* graph_ms) {
    if (!k || !k->gpu || ntimes <= 0) return fail(CLSPMV_ERR_INVALID, "bad argument");
    int st = use_ctx_device(k->ctx);
    if (st) return st;
    clspmv_cocktail_gpu* g = k->gpu;
    hipStream_t s = k->ctx->stream;
    // Profiling accumulates into y: keep a copy and restore it afterwards.
    float* saved = nullptr;
    CLSPMV_HIP(hipMalloc(&saved, sizeof(float) * std::max(g->height, 1)));
    hipError_t e = hipMemcpyAsync(saved, g->res, sizeof(float) * g->height, hipMemcpyDeviceToDevice, s);
    int n = 0;
    for (size_t i = 0; i < g->parts.size() && e == hipSuccess; ++i) {
        // one warm-up launch, then ntimes timed
        st = dmat_launch(s, g->parts[i], g->methods[i], g->vec, g->res);
        if (!st) e = hipEventRecord(k->ctx->ev_start, s);
        for (int r = 0; r < ntimes && !st && e == hipSuccess; ++r)
            st = dmat_launch(s, g->parts[i], g->methods[i], g->vec, g->res);
        if (st) break;
        e = hipEventRecord(k->ctx->ev_stop, s);
        if (e == hipSuccess) e = hipEventSynchronize(k->ctx->ev_stop);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, k->ctx->ev_start, k->ctx->ev_stop);
        if (parts && n < max_parts) {
            parts[n].format = g->parts[i]->info.format;
            parts[n].method = g->methods[i];
            parts[n].nnz = g->parts[i]->info.nnz;
            parts[n].stored = g->parts[i]->info.stored;
            parts[n].ms = (double)ms / ntimes;
        }
        ++n;
    }
    if (!st && e == hipSuccess && graph_ms) {
        // The whole pass in the kernels' launch mode (graph or plain).
        st = clspmv_do_spmv_async(k, 1);
        if (!st) e = hipEventRecord(k->ctx->ev_start, s);
        if (!st && e == hipSuccess) st = clspmv_do_spmv_async(k, ntimes);
        if (!st && e == hipSuccess) e = hipEventRecord(k->ctx->ev_stop, s);
        if (e == hipSuccess) e = hipEventSynchronize(k->ctx->ev_stop);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, k->ctx->ev_start, k->ctx->ev_stop);
        *graph_ms = (double)ms / ntimes;
    }
    hipError_t e2 = hipMemcpyAsync(g->res, saved, sizeof(float) * g->height, hipMemcpyDeviceToDevice, s);
    (void)hipStreamSynchronize(s);
    (void)hipFree(saved);
    if (st) return st;
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(CLSPMV_ERR_HIP, "profile failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    if (nparts) *nparts = n;
    return CLSPMV_OK;
}

int clspmv_cocktail_time_cold(clspmv_cocktail_kernels* k, int ntimes, size_t scrub_bytes, double* ms_per_pass) {
    if (!k || !k->gpu || !k->exec || ntimes <= 0 || !ms_per_pass) return fail(CLSPMV_ERR_INVALID, "bad argument");
    int st = use_ctx_device(k->ctx);
    if (st) return st;
    clspmv_cocktail_gpu* g = k->gpu;
    hipStream_t s = k->ctx->stream;
    // The scrub only reads (a read-only stream over its own buffer), so
    // the timed pass starts with clean caches full of unrelated lines and no
    // write-backs pending.
    const long long n4 = (long long)std::max<size_t>(scrub_bytes / 16, 16);
    constexpr int kScrubUnroll = 4, kScrubGrid = 2048;
    float* saved = nullptr;
    float* scrub = nullptr;
    float* sink = nullptr;
    CLSPMV_HIP(hipMalloc(&saved, sizeof(float) * std::max(g->height, 1)));
    hipError_t e = hipMalloc(&scrub, (size_t)n4 * 16);
    if (e == hipSuccess) e = hipMalloc(&sink, sizeof(float) * kScrubGrid * 256);
    if (e != hipSuccess) {
        (void)hipFree(saved);
        if (scrub) (void)hipFree(scrub);
        return fail(CLSPMV_ERR_ALLOC, "scrub buffer of %zu bytes: %s", (size_t)n4 * 16, hipGetErrorString(e));
    }
    e = hipMemsetAsync(scrub, 0, (size_t)n4 * 16, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(saved, g->res, sizeof(float) * g->height, hipMemcpyDeviceToDevice, s);
    double total = 0.0;
    for (int r = 0; r < ntimes && e == hipSuccess && !st; ++r) {
        e = launch::read4u(s, n4, scrub, sink, kScrubUnroll, kScrubGrid);
        if (e == hipSuccess) e = hipEventRecord(k->ctx->ev_start, s);
        if (e != hipSuccess) break;
        if (k->use_graph)
            e = hipGraphLaunch(k->exec, s);
        else
            st = launch_all(s, k);
        if (e == hipSuccess && !st) e = hipEventRecord(k->ctx->ev_stop, s);
        if (e == hipSuccess && !st) e = hipEventSynchronize(k->ctx->ev_stop);
        float ms = 0.0f;
        if (e == hipSuccess && !st) e = hipEventElapsedTime(&ms, k->ctx->ev_start, k->ctx->ev_stop);
        total += ms;
    }
    hipError_t e2 = hipMemcpyAsync(g->res, saved, sizeof(float) * g->height, hipMemcpyDeviceToDevice, s);
    (void)hipStreamSynchronize(s);
    (void)hipFree(saved);
    (void)hipFree(scrub);
    (void)hipFree(sink);
    if (st) return st;
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(CLSPMV_ERR_HIP, "cold timing failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    *ms_per_pass = total / ntimes;
    return CLSPMV_OK;
}

namespace {

// Average ms of `ntimes` runs of `fn` on the kernels' stream (one untimed
// run first).
template <class F>
int time_runs(clspmv_cocktail_kernels* k, int ntimes, F&& fn, double* ms_out) {
    hipStream_t s = k->ctx->stream;
    int st = fn(s);
    if (st) return st;
    CLSPMV_HIP(hipEventRecord(k->ctx->ev_start, s));
    for (int r = 0; r < ntimes; ++r)
        if ((st = fn(s))) return st;
    CLSPMV_HIP(hipEventRecord(k->ctx->ev_stop, s));
    CLSPMV_HIP(hipEventSynchronize(k->ctx->ev_stop));
    float ms = 0.0f;
    CLSPMV_HIP(hipEventElapsedTime(&ms, k->ctx->ev_start, k->ctx->ev_stop));
    *ms_out = (double)ms / ntimes;
    return CLSPMV_OK;
}

// Lane counts worth timing around the upload's choice (powers of two).
std::vector<int> lane_candidates(int lanes) {
    std::vector<int> c;
    for (int l = std::max(1, lanes / 4); l <= std::min(64, lanes * 4); l <<= 1) c.push_back(l);
    return c;
}

// Trial number (counted from 1 over one tuning call) at which tune_parts
// fails on purpose, from CLSPMV_TUNE_FAIL_TRIAL; 0 = never.  A host-side
// error for tests of the recovery path: no kernel is involved.
int injected_fail_trial() {
    const char* v = std::getenv("CLSPMV_TUNE_FAIL_TRIAL");
    return v ? std::atoi(v) : 0;
}

int tune_parts(clspmv_cocktail_kernels* k, int ntimes) {
    clspmv_cocktail_gpu* g = k->gpu;
    const int fail_at = injected_fail_trial();
    int trials = 0;
    for (size_t i = 0; i < g->parts.size(); ++i) {
        clspmv_dmat* A = g->parts[i];
        const int fmt = A->info.format;
        double best = 1e30;
        int best_m = g->methods[i], best_l = A->lanes;
        bool best_panel = A->use_panel;
        // Lanes per row: CSR method 0, BCSR in its row-pointer encoding.
        const bool lane_fmt = fmt == CLSPMV_FMT_CSR || (fmt == CLSPMV_FMT_BCSR && A->bslots == 0);
        const std::vector<int> lanes = lane_fmt ? lane_candidates(A->lanes) : std::vector<int>{A->lanes};
        // The part's state before its trials, restored if one fails.
        const int upload_lanes = A->lanes, upload_method = g->methods[i];
        const bool upload_panel = A->use_panel;
        auto trial = [&](int m, int l, bool panel) -> int {
            if (++trials == fail_at) return fail(CLSPMV_ERR_HIP, "tune: injected failure at trial %d", trials);
            A->lanes = l;
            A->use_panel = panel;
            g->methods[i] = m;
            // The whole pass is timed, not the part alone: a part run by
            // itself keeps x and y in cache between launches, which the
            // other parts' streams evict in a real pass.  Best of three
            // batches: one batch's average can carry a stall that flips the
            // choice between kernels a few microseconds apart.
            double ms = 1e30;
            for (int rep = 0; rep < 3; ++rep) {
                double t = 0.0;
                int st = time_runs(k, ntimes, [&](hipStream_t s) { return launch_all(s, k); }, &t);
                if (st) return st;
                ms = std::min(ms, t);
            }
            if (ms < best) {
                best = ms;
                best_m = m;
                best_l = l;
                best_panel = panel;
            }
            return CLSPMV_OK;
        };
        int st = CLSPMV_OK;
        if (A->pw)  // the panel encoding runs the same kernels for every method
            st = trial(0, upload_lanes, true);
        for (int m = 0; !st && A->plain && m < clspmv_method_count(fmt); ++m) {
            if (!method_runs(A, m)) continue;
            const bool uses_lanes = lane_fmt && (fmt == CLSPMV_FMT_BCSR || m == 0);
            for (int l : uses_lanes ? lanes : std::vector<int>{upload_lanes})
                if ((st = trial(m, l, false))) break;
        }
        if (st) {
            A->lanes = upload_lanes;
            A->use_panel = upload_panel;
            g->methods[i] = upload_method;
            return st;
        }
        g->methods[i] = best_m;
        A->lanes = best_l;
        A->use_panel = best_panel;
        // Keep only the encoding that won.
        if (A->pw && A->plain) drop_encoding(A, !best_panel);
    }
    return CLSPMV_OK;
}

}  // namespace

int clspmv_cocktail_tune(clspmv_cocktail_kernels* k, int ntimes, int* graph) {
    if (!k || !k->gpu || ntimes <= 0) return fail(CLSPMV_ERR_INVALID, "bad argument");
    int st = use_ctx_device(k->ctx);
    if (st) return st;
    clspmv_cocktail_gpu* g = k->gpu;
    hipStream_t s = k->ctx->stream;
    // Tuning launches accumulate into y: keep a copy and restore it after.
    float* saved = nullptr;
    const size_t ybytes = sizeof(float) * g->height;
    CLSPMV_HIP(hipMalloc(&saved, sizeof(float) * std::max(g->height, 1)));
    hipError_t e = hipMemcpyAsync(saved, g->res, ybytes, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) st = fail(CLSPMV_ERR_HIP, "tune: %s", hipGetErrorString(e));
    if (!st) st = tune_parts(k, ntimes);
    // Launch mode (graph or plain launches) and whether the panel first
    // passes run beside the earlier parts: every combination captured and
    // timed (best of three batches each), the fastest kept.  Neither wins
    // everywhere: graph nodes of one stream are spaced further apart than
    // plain launches, and a second stream costs plain launches an event
    // round trip per pass.
    // runs calls of fn per batch, each `per` passes: ms per pass.
    auto best_of3 = [&](auto&& fn, int runs, int per, double* out) -> int {
        double best = 1e30;
        for (int rep = 0; rep < 3; ++rep) {
            double t = 0.0;
            int r = time_runs(k, runs, fn, &t);
            if (r) return r;
            best = std::min(best, t / per);
        }
        *out = best;
        return CLSPMV_OK;
    };
    double best_ms = 1e30;
    bool best_graph = true, best_overlap = true, best_fuse = false;
    // (overlap, fuse) candidates: the first is the default.
    std::vector<std::pair<bool, bool>> modes{{true, false}};
    if (overlap_from(g) >= 0) modes.push_back({false, false});
    if (fuse_at(g) >= 0) modes.push_back({false, true});
    for (size_t m = 0; m < modes.size() && !st; ++m) {
        k->overlap = modes[m].first;
        k->fuse = modes[m].second;
        double t_graph = 0.0, t_direct = 0.0;
        // Graph launches as do_spmv issues them (kGraphPasses passes each).
        if (!(st = capture(k)))
            st = best_of3([&](hipStream_t q) -> int {
                CLSPMV_HIP(hipGraphLaunch(k->exec_many, q));
                return CLSPMV_OK;
            }, std::max(2, ntimes / kGraphPasses), kGraphPasses, &t_graph);
        if (!st) st = best_of3([&](hipStream_t q) { return launch_all(q, k); }, ntimes, 1, &t_direct);
        if (st) break;
        if (t_graph < best_ms) best_ms = t_graph, best_graph = true, best_overlap = k->overlap, best_fuse = k->fuse;
        if (t_direct < best_ms) best_ms = t_direct, best_graph = false, best_overlap = k->overlap, best_fuse = k->fuse;
    }
    if (!st && (k->overlap != best_overlap || k->fuse != best_fuse)) {
        k->overlap = best_overlap;
        k->fuse = best_fuse;
        st = capture(k);
    }
    e = hipMemcpyAsync(g->res, saved, ybytes, hipMemcpyDeviceToDevice, s);
    hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(saved);
    if (st) {
        // Parts tuned before the failure may have dropped an encoding that
        // the graph captured earlier still points at: capture the pass again
        // from the parts' current state.  If that fails too, capture() has
        // left no graph, and do_spmv refuses to run.
        const int err = st;
        (void)capture(k);
        return err;
    }
    if (e != hipSuccess || e2 != hipSuccess)
        return fail(CLSPMV_ERR_HIP, "tune: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    k->use_graph = best_graph;
    if (graph) *graph = k->use_graph ? 1 : 0;
    return CLSPMV_OK;
}

int clspmv_cocktail_set_graph(clspmv_cocktail_kernels* k, int graph) {
    if (!k || !k->exec) return fail(CLSPMV_ERR_INVALID, "kernels not initialised");
    k->use_graph = graph != 0;
    return CLSPMV_OK;
}

int clspmv_cocktail_set_overlap(clspmv_cocktail_kernels* k, int overlap) {
    if (!k || !k->exec) return fail(CLSPMV_ERR_INVALID, "kernels not initialised");
    int st = use_ctx_device(k->ctx);
    if (st) return st;
    if (k->overlap == (overlap != 0)) return CLSPMV_OK;
    k->overlap = overlap != 0;
    return capture(k);
}

int clspmv_cocktail_set_fuse(clspmv_cocktail_kernels* k, int fuse) {
    if (!k || !k->exec) return fail(CLSPMV_ERR_INVALID, "kernels not initialised");
    int st = use_ctx_device(k->ctx);
    if (st) return st;
    if (k->fuse == (fuse != 0)) return CLSPMV_OK;
    k->fuse = fuse != 0;
    return capture(k);
}

int clspmv_cocktail_launch_modes(const clspmv_cocktail_kernels* k, int* graph, int* overlap, int* fuse) {
    if (!k || !k->gpu) return fail(CLSPMV_ERR_INVALID, "kernels not initialised");
    if (graph) *graph = k->use_graph ? 1 : 0;
    if (overlap) *overlap = overlaps(k) ? 1 : 0;
    if (fuse) *fuse = fuses(k) ? 1 : 0;
    return CLSPMV_OK;
}

// evaluate (src/spmv_cocktail.cpp:1431) with check_and_time (:1134).
int clspmv_evaluate(const clspmv_cocktail* cpumat, const float* vec, float* res,
                    const float* coores, clspmv_eval_report* report) {
    if (!cpumat || !vec || !res) return fail(CLSPMV_ERR_INVALID, "null argument");
    clspmv_cocktail_gpu* g = nullptr;
    clspmv_cocktail_kernels* k = nullptr;
    int st = clspmv_init_mat_kernels(cpumat, vec, res, &g, &k, 1);
    if (st) return st;
    const int h = cpumat->mat_height;
    std::vector<float> y(std::max(h, 1));
    st = clspmv_do_spmv(k, 1);
    if (!st) st = clspmv_cpy_result_to_cpu(g, k, y.data(), h);
    clspmv_eval_report rep{};
    if (!st && coores) {
        // The reference's report (src/util.cpp:93-131).
        const VecCompare c = two_vec_compare(coores, y.data(), h);
        rep.l2_distance = c.distance;
        rep.max_diff = c.max_diff;
        rep.max_ratio = c.max_ratio;
        std::printf("\nCorrectness Check: Distance %e max diff %e max diff ratio %e vec size %d\n",
                    rep.l2_distance, rep.max_diff, rep.max_ratio, h);
    }
    if (!st) {
        const int ntimes = 1000;
        st = use_ctx_device(k->ctx);
        hipError_t e = hipSuccess;
        if (!st) e = hipEventRecord(k->ctx->ev_start, k->ctx->stream);
        if (!st && e == hipSuccess) st = clspmv_do_spmv_async(k, ntimes);
        if (!st && e == hipSuccess) e = hipEventRecord(k->ctx->ev_stop, k->ctx->stream);
        if (!st && e == hipSuccess) e = hipEventSynchronize(k->ctx->ev_stop);
        float ms = 0.0f;
        if (!st && e == hipSuccess) e = hipEventElapsedTime(&ms, k->ctx->ev_start, k->ctx->ev_stop);
        if (!st && e != hipSuccess) st = fail(CLSPMV_ERR_HIP, "evaluate timing: %s", hipGetErrorString(e));
        if (!st) {
            rep.ms_per_spmv = (double)ms / ntimes;
            rep.gflops = 2.0 * cpumat->mat_nnz / (rep.ms_per_spmv * 1e-3) / 1e9;
            std::printf("\nCocktail time %lf ms GFLOPS %lf\n\n", rep.ms_per_spmv, rep.gflops);
        }
    }
    std::copy(y.begin(), y.begin() + h, res);
    clspmv_free_mat_kernels(g, k);
    if (report) *report = rep;
    return st;
}
