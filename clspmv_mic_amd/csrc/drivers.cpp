// Per-format benchmark drivers: the spmv_<fmt>() entry points of the
// reference (include/spmv_<fmt>.h:7, e.g. src/spmv_bdia.cpp:413-440).
//
// Each driver converts the COO matrix to its format with the reference's
// parameters (alignment GPU_ALIGNMENT, slice heights 32/128, the eight block
// shapes), uploads it, and for every kernel variant: zeroes y, runs one SpMV
// with x = 1, keeps the result, then times `ntimes` launches after three
// warm-up launches (src/spmv_bdia.cpp:111-137).  It prints one line per
// variant and the best one, and fills a report.  Like the reference, every
// variant's first result is checked against the serial COO SpMV of the same
// matrix on the host (spmv_only + two_vec_compare, src/spmv_bdia.cpp:118,
// :422-423), and also scored against an FP64 host product by the row-scaled
// tolerance |y - y64| <= 1e-5 (|A||x|); variants over it are counted.

#include <algorithm>
#include <cmath>
#include <functional>

#include "runtime.hpp"

using namespace clspmv;

namespace {

struct Variant {
    clspmv_dmat* A;
    int method;
    int bh, bw;
};

constexpr double kDriverTol = 1e-5;

struct DriverRun {
    clspmv_ctx* ctx = nullptr;
    float* x = nullptr;
    float* y = nullptr;
    int rows = 0, cols = 0;
    std::vector<float> first;
    std::vector<float> coores;         // serial COO SpMV, x = 1 (spmv_only, src/util.cpp:85)
    std::vector<double> y64, bound;    // FP64 product and sum |a x| per row
    ~DriverRun() {
        if (ctx) {
            if (x) (void)hipFree(x);
            if (y) (void)hipFree(y);
            clspmv_ctx_destroy(ctx);
        }
    }
};

// The host references of a driver run for x = 1: the reference's serial
// check (entries in storage order, FP32) and an FP64 product with its
// per-row bound.  The matrix is sorted first, as every conversion of the
// reference sorts it in place (include/matrix_storage.h:1271).
void host_references(DriverRun& R, clspmv_coo_matrix* mat) {
    if (!coo_sorted(*mat)) coo_sort(*mat);
    const int nnz = mat->matinfo.nnz;
    R.coores.assign(std::max(R.rows, 1), 0.0f);
    R.y64.assign(std::max(R.rows, 1), 0.0);
    R.bound.assign(std::max(R.rows, 1), 0.0);
    for (int i = 0; i < nnz; ++i) {
        const int r = mat->coo_row_id[i];
        const float a = mat->coo_data[i];
        R.coores[r] += a * 1.0f;
        R.y64[r] += (double)a;
        R.bound[r] += std::fabs((double)a);
    }
}

int driver_setup(DriverRun& R, int device, clspmv_coo_matrix* mat) {
    int st = clspmv_ctx_create(device, &R.ctx);
    if (st) return st;
    R.rows = mat->matinfo.height;
    R.cols = mat->matinfo.width;
    host_references(R, mat);
    CLSPMV_HIP(hipMalloc(&R.x, sizeof(float) * std::max(R.cols, 1)));
    CLSPMV_HIP(hipMalloc(&R.y, sizeof(float) * std::max(R.rows, 1)));
    std::vector<float> ones(std::max(R.cols, 1), 1.0f);
    return clspmv_memcpy_h2d(R.ctx, R.x, ones.data(), sizeof(float) * R.cols);
}

// Runs one variant; returns its status.  Records timing in `rep`.
int run_variant(DriverRun& R, const Variant& v, int ntimes, const char* name,
                clspmv_driver_report* rep, float* y_out, double nnz) {
    CLSPMV_HIP(hipMemsetAsync(R.y, 0, sizeof(float) * R.rows, R.ctx->stream));
    int st = clspmv_spmv(R.ctx, v.A, v.method, R.x, R.y);
    if (st) return st;
    std::vector<float> y(R.rows);
    st = clspmv_memcpy_d2h(R.ctx, y.data(), R.y, sizeof(float) * R.rows);
    if (st) return st;
    if (rep->variants == 0) {
        R.first = y;
    } else {
        for (int i = 0; i < R.rows; ++i) {
            double a = std::fabs((double)y[i]), b = std::fabs((double)R.first[i]);
            double m = std::max(a, b);
            if (m > 0) rep->max_rel_diff = std::max(rep->max_rel_diff, std::fabs((double)y[i] - R.first[i]) / m);
        }
    }
    // The reference's check against the serial result, and the tolerance.
    const VecCompare c = two_vec_compare(R.coores.data(), y.data(), R.rows);
    std::printf("\nCorrectness Check: Distance %e max diff %e max diff ratio %e vec size %d\n", c.distance,
                c.max_diff, c.max_ratio, R.rows);
    rep->check_distance = std::max(rep->check_distance, c.distance);
    rep->check_max_diff = std::max(rep->check_max_diff, c.max_diff);
    rep->check_max_ratio = std::max(rep->check_max_ratio, c.max_ratio);
    double worst = 0.0;
    for (int i = 0; i < R.rows; ++i) {
        const double d = std::fabs((double)y[i] - R.y64[i]);
        double e = R.bound[i] > 0 ? d / R.bound[i] : (d > 0 ? INFINITY : 0.0);
        if (e != e) e = INFINITY;
        worst = std::max(worst, e);
    }
    rep->max_scaled_error = std::max(rep->max_scaled_error, worst);
    if (worst > kDriverTol) rep->variants_over_tol++;
    double ms = 0.0;
    st = clspmv_time_spmv(R.ctx, v.A, v.method, R.x, R.y, 3, std::max(ntimes, 1), &ms);
    if (st) return st;
    double gflops = ms > 0 ? 2.0 * nnz / (ms * 1e-3) / 1e9 : 0.0;
    std::printf("%s method %d block %dx%d: %.6f ms, %.3f GFLOPS\n", name, v.method, v.bh,
                v.bw, ms, gflops);
    int k = rep->variants;
    if (k < 64) {
        rep->ms[k] = ms;
        rep->method[k] = v.method;
        rep->bheight[k] = v.bh;
        rep->bwidth[k] = v.bw;
    }
    if (rep->variants == 0 || ms < rep->best_ms) {
        rep->best_ms = ms;
        rep->best_gflops = gflops;
        rep->best_method = v.method;
        rep->best_bheight = v.bh;
        rep->best_bwidth = v.bw;
        if (y_out) std::copy(y.begin(), y.end(), y_out);
    }
    rep->variants++;
    return CLSPMV_OK;
}

// Shared driver body: `make(shape_index, &dmat)` converts and uploads for one
// block shape (or the single layout when nshapes == 1), returning 0 to skip.
int drive(const char* name, int format, clspmv_coo_matrix* mat, int ntimes,
          int device, clspmv_driver_report* report, float* y_out, int nshapes,
          const std::function<int(int, clspmv_ctx*, clspmv_dmat**)>& make,
          const std::function<bool(int, const clspmv_dmat*)>& method_ok) {
    if (!mat) return fail(CLSPMV_ERR_INVALID, "null matrix");
    int st = clspmv_validate_coo(mat);
    if (st) return st;
    clspmv_driver_report local{};
    clspmv_driver_report* rep = report ? report : &local;
    *rep = clspmv_driver_report{};
    rep->format = format;
    DriverRun R;
    st = driver_setup(R, device, mat);
    if (st) return st;
    const double nnz = mat->matinfo.nnz;
    for (int sh = 0; sh < nshapes; ++sh) {
        clspmv_dmat* A = nullptr;
        st = make(sh, R.ctx, &A);
        if (st < 0) return st;
        if (!A) continue;  // shape refused (e.g. DIA padding rule)
        for (int meth = 0; meth < clspmv_method_count(format); ++meth) {
            if (!method_ok(meth, A)) continue;
            Variant v{A, meth, A->info.bheight, A->info.bwidth};
            st = run_variant(R, v, ntimes, name, rep, y_out, nnz);
            if (st) {
                clspmv_dmat_free(A);
                return st;
            }
        }
        clspmv_dmat_free(A);
    }
    if (rep->variants == 0) return fail(CLSPMV_ERR_UNSUPPORTED, "%s: no variant ran", name);
    std::printf("%s best: method %d block %dx%d %.6f ms %.3f GFLOPS (max variant diff %.3e, "
                "max scaled error %.3e, %d variants over %.0e)\n",
                name, rep->best_method, rep->best_bheight, rep->best_bwidth, rep->best_ms,
                rep->best_gflops, rep->max_rel_diff, rep->max_scaled_error, rep->variants_over_tol, kDriverTol);
    return CLSPMV_OK;
}

const unsigned kBH[8] = {1, 2, 4, 8, 1, 2, 4, 8};
const unsigned kBW[8] = {4, 4, 4, 4, 8, 8, 8, 8};

bool all_methods(int, const clspmv_dmat*) { return true; }

}  // namespace

extern "C" {

int clspmv_spmv_csr_scalar(const char*, clspmv_coo_matrix* mat, int, int ntimes, int device,
                           clspmv_driver_report* report, float* y_out) {
    return drive("csr_scalar", CLSPMV_FMT_CSR, mat, ntimes, device, report, y_out, 1,
                 [&](int, clspmv_ctx* ctx, clspmv_dmat** A) {
                     clspmv_csr_matrix c;
                     clspmv_init_csr_matrix(&c);
                     int st = clspmv_coo2csr(mat, &c);
                     if (!st) st = clspmv_upload_csr(ctx, &c, A);
                     clspmv_free_csr_matrix(&c);
                     if (!st) (*A)->lanes = 1;  // one lane per row
                     return st;
                 },
                 [](int meth, const clspmv_dmat*) { return meth == 0; });
}

int clspmv_spmv_csr_vector(const char*, clspmv_coo_matrix* mat, int, int ntimes, int device,
                           clspmv_driver_report* report, float* y_out) {
    return drive("csr_vector", CLSPMV_FMT_CSR, mat, ntimes, device, report, y_out, 1,
                 [&](int, clspmv_ctx* ctx, clspmv_dmat** A) {
                     clspmv_csr_matrix c;
                     clspmv_init_csr_matrix(&c);
                     int st = clspmv_coo2csr(mat, &c);
                     if (!st) st = clspmv_upload_csr(ctx, &c, A);
                     clspmv_free_csr_matrix(&c);
                     return st;
                 },
                 all_methods);
}

int clspmv_spmv_coo(const char*, clspmv_coo_matrix* mat, int, int ntimes, int device,
                    clspmv_driver_report* report, float* y_out) {
    return drive("coo", CLSPMV_FMT_COO, mat, ntimes, device, report, y_out, 1,
                 [&](int, clspmv_ctx* ctx, clspmv_dmat** A) {
                     if (!coo_sorted(*mat)) coo_sort(*mat);
                     return clspmv_upload_coo(ctx, mat, A);
                 },
                 all_methods);
}

int clspmv_spmv_ell(const char*, clspmv_coo_matrix* mat, int, int ntimes, int device,
                    clspmv_driver_report* report, float* y_out) {
    return drive("ell", CLSPMV_FMT_ELL, mat, ntimes, device, report, y_out, 1,
                 [&](int, clspmv_ctx* ctx, clspmv_dmat** A) {
                     clspmv_ell_matrix e;
                     clspmv_init_ell_matrix(&e);
                     int st = clspmv_coo2ell(mat, &e, CLSPMV_GPU_ALIGNMENT, 0);
                     if (!st) st = clspmv_upload_ell(ctx, &e, A);
                     clspmv_free_ell_matrix(&e);
                     return st;
                 },
                 all_methods);
}

int clspmv_spmv_sell(const char*, clspmv_coo_matrix* mat, int, int ntimes, int device,
                     clspmv_driver_report* report, float* y_out) {
    // Two layouts: slice height WARPSIZE (method 0) and SELL_GROUP_SIZE
    // (method 1), as the Cocktail runtime pairs them (src/analyze.cpp:3099).
    return drive("sell", CLSPMV_FMT_SELL, mat, ntimes, device, report, y_out, 2,
                 [&](int sh, clspmv_ctx* ctx, clspmv_dmat** A) {
                     if (!coo_sorted(*mat)) coo_sort(*mat);
                     clspmv_sell_matrix s;
                     clspmv_init_sell_matrix(&s);
                     int st = clspmv_coo2sell(mat, &s, sh == 0 ? CLSPMV_WARPSIZE : CLSPMV_SELL_GROUP_SIZE);
                     if (!st) st = clspmv_upload_sell(ctx, &s, A);
                     clspmv_free_sell_matrix(&s);
                     return st;
                 },
                 [](int meth, const clspmv_dmat* A) {
                     return meth == (A->info.slice_height == CLSPMV_WARPSIZE ? 0 : 1);
                 });
}

int clspmv_spmv_dia(const char*, clspmv_coo_matrix* mat, int, int ntimes, int device,
                    clspmv_driver_report* report, float* y_out) {
    return drive("dia", CLSPMV_FMT_DIA, mat, ntimes, device, report, y_out, 1,
                 [&](int, clspmv_ctx* ctx, clspmv_dmat** A) {
                     clspmv_dia_matrix d;
                     clspmv_init_dia_matrix(&d);
                     int ok = clspmv_coo2dia(mat, &d, CLSPMV_GPU_ALIGNMENT);
                     if (ok < 0) return ok;
                     if (ok == 0) return fail(CLSPMV_ERR_UNSUPPORTED, "dia: too many diagonals");
                     int st = clspmv_upload_dia(ctx, &d, A);
                     clspmv_free_dia_matrix(&d);
                     return st;
                 },
                 [](int meth, const clspmv_dmat* A) {
                     int span = A->max_off - A->min_off;
                     return meth < 2 || (meth == 2 ? 256 : 1024) + span <= 8192;
                 });
}

int clspmv_spmv_bdia(const char*, clspmv_coo_matrix* mat, int, int ntimes, int device,
                     clspmv_driver_report* report, float* y_out) {
    return drive("bdia", CLSPMV_FMT_BDIA, mat, ntimes, device, report, y_out, 1,
                 [&](int, clspmv_ctx* ctx, clspmv_dmat** A) {
                     clspmv_bdia_matrix d;
                     clspmv_init_bdia_matrix(&d);
                     int ok = clspmv_coo2bdia(mat, &d, CLSPMV_GPU_ALIGNMENT);
                     if (ok < 0) return ok;
                     if (ok == 0) return fail(CLSPMV_ERR_UNSUPPORTED, "bdia: too many diagonals");
                     int st = clspmv_upload_bdia(ctx, &d, A);
                     clspmv_free_bdia_matrix(&d);
                     return st;
                 },
                 [](int meth, const clspmv_dmat* A) {
                     static const int rpt[5] = {1, 1, 2, 2, 4};
                     static const bool lds[5] = {false, true, true, false, true};
                     return !lds[meth] || rpt[meth] * 256 + A->max_band <= 4096;
                 });
}

int clspmv_spmv_bell(const char*, clspmv_coo_matrix* mat, int, int ntimes, int device,
                     clspmv_driver_report* report, float* y_out) {
    return drive("bell", CLSPMV_FMT_BELL, mat, ntimes, device, report, y_out, 8,
                 [&](int sh, clspmv_ctx* ctx, clspmv_dmat** A) {
                     clspmv_b4ell_matrix b;
                     clspmv_init_b4ell_matrix(&b);
                     int ok = clspmv_coo2b4ell_nocap(mat, &b, kBW[sh], kBH[sh], CLSPMV_GPU_ALIGNMENT, 0);
                     if (ok < 0) return ok;
                     int st = clspmv_upload_b4ell(ctx, &b, A);
                     clspmv_free_b4ell_matrix(&b);
                     return st;
                 },
                 all_methods);
}

int clspmv_spmv_sbell(const char*, clspmv_coo_matrix* mat, int, int ntimes, int device,
                      clspmv_driver_report* report, float* y_out) {
    return drive("sbell", CLSPMV_FMT_SBELL, mat, ntimes, device, report, y_out, 8,
                 [&](int sh, clspmv_ctx* ctx, clspmv_dmat** A) {
                     clspmv_sbell_matrix b;
                     clspmv_init_sbell_matrix(&b);
                     int ok = clspmv_coo2sbell_nocap(mat, &b, kBW[sh], kBH[sh], CLSPMV_WARPSIZE);
                     if (ok < 0) return ok;
                     int st = clspmv_upload_sbell(ctx, &b, A);
                     clspmv_free_sbell_matrix(&b);
                     return st;
                 },
                 all_methods);
}

int clspmv_spmv_bcsr(const char*, clspmv_coo_matrix* mat, int, int ntimes, int device,
                     clspmv_driver_report* report, float* y_out) {
    return drive("bcsr", CLSPMV_FMT_BCSR, mat, ntimes, device, report, y_out, 8,
                 [&](int sh, clspmv_ctx* ctx, clspmv_dmat** A) {
                     clspmv_b4csr_matrix b;
                     clspmv_init_b4csr_matrix(&b);
                     int ok = clspmv_coo2b4csr_nocap(mat, &b, kBW[sh], kBH[sh], CLSPMV_GPU_ALIGNMENT);
                     if (ok < 0) return ok;
                     int st = clspmv_upload_b4csr(ctx, &b, A);
                     clspmv_free_b4csr_matrix(&b);
                     return st;
                 },
                 all_methods);
}

}  // extern "C"
