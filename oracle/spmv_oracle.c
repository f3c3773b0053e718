/*
 * spmv_oracle.c -- CPU oracle of the clspmv engine.  TEST INFRASTRUCTURE
 * ONLY: called by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline as the checker.  The engine never links or calls it.
 *
 * Contents
 *   1. FP32 serial SpMV in the reference's own loop order, one routine per
 *      storage format (include/spmv_serial.h), computing y += A*x.
 *   2. FP64 SpMV from COO triples: y64 = y0 + A*x accumulated in double, and
 *      the row bound b_i = |y0_i| + sum_j |a_ij x_j| used for the tolerance
 *      |y_i - y64_i| <= tol * b_i.
 *   3. The reference's vector comparison report (src/util.cpp:93-131).
 *   4. The CSR SpMV of part 1 spread over host threads (OpenMP), for the
 *      all-core CPU baseline; identical results to the serial routine.
 *
 * How each part is pinned (see tests/test_oracle.py):
 *   - closed forms: 2D/3D Laplacian stencils with x = 1 give y_i = centre -
 *     (#neighbours); dyadic inputs make every FP32 sum exact, so every
 *     format's serial SpMV must equal the FP64 result bit for bit;
 *   - brute force: a dense matrix-vector product in numpy on tiny inputs;
 *   - the reference itself: the same routines compiled from
 *     include/spmv_serial.h into oracle/_ref (where /root/reference exists)
 *     must agree bit for bit with the FP32 routines here.
 * Format routines for b4csr and b4ell have no counterpart in the reference's
 * spmv_serial.h; they restate the kernels' arithmetic (kernels/spmv_bcsr.cl,
 * kernels/spmv_bell.cl) and are pinned by the FP64 and dyadic checks only.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- 1. FP32 serial SpMV, reference loop order --------------------------- */

/* coo_spmv, include/spmv_serial.h:184 */
void oracle_coo_spmv(int nnz, const int* row, const int* col, const float* val,
                     const float* x, float* y) {
    for (int i = 0; i < nnz; ++i) y[row[i]] += val[i] * x[col[i]];
}

/* csr_spmv, include/spmv_serial.h:199 */
void oracle_csr_spmv(int rows, const int* rp, const int* col, const float* val,
                     const float* x, float* y) {
    for (int r = 0; r < rows; ++r) {
        float acc = 0.0f;
        for (int k = rp[r]; k < rp[r + 1]; ++k) acc += val[k] * x[col[k]];
        y[r] += acc;
    }
}

/* csr_spmv over `threads` host threads (OpenMP, rows split statically):
 * every row is summed by one thread in the serial loop order, so y is the
 * same bit for bit as oracle_csr_spmv's.  The all-core CPU baseline. */
void oracle_csr_spmv_threads(int threads, int rows, const int* rp, const int* col, const float* val,
                             const float* x, float* y) {
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int r = 0; r < rows; ++r) {
        float acc = 0.0f;
        for (int k = rp[r]; k < rp[r + 1]; ++k) acc += val[k] * x[col[k]];
        y[r] += acc;
    }
}

/* Host threads OpenMP would use by default. */
int oracle_max_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ell_spmv, include/spmv_serial.h:93 -- slot-major, accumulating into y */
void oracle_ell_spmv(int rows, int hal, int ellnum, const int* col,
                     const float* val, const float* x, float* y) {
    for (int s = 0; s < ellnum; ++s)
        for (int r = 0; r < rows; ++r)
            y[r] += val[(size_t)s * hal + r] * x[col[(size_t)s * hal + r]];
}

/* sell_spmv, include/spmv_serial.h:113 */
void oracle_sell_spmv(int rows, int h, int nslice, const int* ptr,
                      const int* col, const float* val, const float* x, float* y) {
    for (int s = 0; s < nslice; ++s) {
        int w = (ptr[s + 1] - ptr[s]) / h;
        for (int k = 0; k < w; ++k)
            for (int r = s * h; r < (s + 1) * h && r < rows; ++r) {
                size_t i = (size_t)ptr[s] + (size_t)(r - s * h) + (size_t)k * h;
                y[r] += val[i] * x[col[i]];
            }
    }
}

/* dia_spmv, include/spmv_serial.h:12 -- rows up to the aligned length, so
 * y must hold `len` entries (rows past the height receive zeros). */
void oracle_dia_spmv(int len, int ndiag, const int* off, const float* data,
                     const float* x, int xlen, float* y) {
    for (int r = 0; r < len; ++r) {
        float acc = 0.0f;
        for (int j = 0; j < ndiag; ++j) {
            long long c = (long long)r + off[j];
            float xv = (c >= 0 && c < xlen) ? x[c] : 0.0f;
            acc += xv * data[(size_t)j * len + r];
        }
        y[r] += acc;
    }
}

/* bdia_spmv, include/spmv_serial.h:60 */
void oracle_bdia_spmv(int len, int nband, const int* boff, const int* bptr,
                      const float* data, const float* x, int xlen, float* y) {
    for (int r = 0; r < len; ++r) {
        float acc = 0.0f;
        for (int b = 0; b < nband; ++b) {
            long long o = boff[b];
            for (int d = bptr[b]; d < bptr[b + 1]; ++d, ++o) {
                long long c = (long long)r + o;
                float xv = (c >= 0 && c < xlen) ? x[c] : 0.0f;
                acc += xv * data[(size_t)d * len + r];
            }
        }
        y[r] += acc;
    }
}

/* bcsr_spmv, include/spmv_serial.h:220 (element-unit columns) */
void oracle_bcsr_spmv(int rows, int cols, int brows, int bh, int bw, int aligned,
                      const int* rp, const int* bcol, const float* data,
                      const float* x, float* y) {
    for (int br = 0; br < brows; ++br)
        for (int k = 0; k < bw * bh; ++k)
            for (int j = rp[br]; j < rp[br + 1]; ++j) {
                int c = bcol[j] + k % bw, r = br * bh + k / bw;
                if (c < cols && r < rows) y[r] += data[j + (size_t)k * aligned] * x[c];
            }
}

/* bell_spmv, include/spmv_serial.h:252; y must hold brows*bh entries. */
void oracle_bell_spmv(int brows, int bh, int bw, int hal, int ellnum,
                      const int* bcol, const float* data, const float* x, float* y) {
    size_t slot = (size_t)hal * bw * bh;
    for (int s = 0; s < ellnum; ++s)
        for (int r = 0; r < brows; ++r)
            for (int k = 0; k < bw * bh; ++k)
                y[r * bh + k / bw] += data[r + (size_t)k * hal + s * slot] *
                                      x[bcol[(size_t)s * hal + r] + k % bw];
}

/* sbell_spmv, include/spmv_serial.h:138 (float4-unit columns) */
void oracle_sbell_spmv(int rows, int cols, int bh, int bw, int sh, int nslice,
                       const int* sptr, const int* col4, const float* data,
                       const float* x, float* y) {
    int bs = bw * bh, bw4 = bw / 4;
    for (int s = 0; s < nslice; ++s) {
        int start = sptr[s], w = (sptr[s + 1] - start) / sh;
        for (int i = 0; i < sh; ++i) {
            size_t coff = (size_t)start + i;
            size_t doff = (size_t)start * bs + (size_t)i * 4;
            for (int k = 0; k < w; ++k) {
                int c = col4[coff] * 4;
                for (int w4 = 0; w4 < bw4; ++w4) {
                    for (int h = 0; h < bh; ++h) {
                        int r = (s * sh + i) * bh + h;
                        if (r < rows)
                            for (int q = 0; q < 4; ++q) {
                                if (c + q >= cols) break;
                                y[r] += data[doff + q] * x[c + q];
                            }
                        doff += (size_t)sh * 4;
                    }
                    c += 4;
                }
                coff += sh;
            }
        }
    }
}

/* b4csr: block j of block row br, row h, float4 column w4 holds
 * data[w4*(aligned*bh) + h*aligned + j*4 + q] (include/matrix_storage.h:
 * 1872-1895); columns in float4 units. */
void oracle_b4csr_spmv(int rows, int cols, int brows, int bh, int bw, int aligned,
                       const int* rp, const int* col4, const float* data,
                       const float* x, float* y) {
    int bw4 = bw / 4;
    for (int br = 0; br < brows; ++br)
        for (int h = 0; h < bh; ++h) {
            int r = br * bh + h;
            if (r >= rows) continue;
            float acc = 0.0f;
            for (int j = rp[br]; j < rp[br + 1]; ++j)
                for (int w4 = 0; w4 < bw4; ++w4)
                    for (int q = 0; q < 4; ++q) {
                        int c = (col4[j] + w4) * 4 + q;
                        if (c < cols)
                            acc += data[(size_t)w4 * aligned * bh + (size_t)h * aligned +
                                        (size_t)j * 4 + q] * x[c];
                    }
            y[r] += acc;
        }
}

/* b4ell: slot s, block row r, row h, float4 column w4 at
 * data[s*(f4*bw4*bh) + h*f4 + w4*(f4*bh) + r*4 + q] (matrix_storage.h:2160-2178). */
void oracle_b4ell_spmv(int rows, int cols, int brows, int bh, int bw, int hal,
                       int f4, int ellnum, const int* col4, const float* data,
                       const float* x, float* y) {
    int bw4 = bw / 4;
    for (int br = 0; br < brows; ++br)
        for (int h = 0; h < bh; ++h) {
            int r = br * bh + h;
            if (r >= rows) continue;
            float acc = 0.0f;
            for (int s = 0; s < ellnum; ++s)
                for (int w4 = 0; w4 < bw4; ++w4)
                    for (int q = 0; q < 4; ++q) {
                        int c = (col4[br + (size_t)s * hal] + w4) * 4 + q;
                        if (c < cols)
                            acc += data[(size_t)s * f4 * bw4 * bh + (size_t)h * f4 +
                                        (size_t)w4 * f4 * bh + (size_t)br * 4 + q] * x[c];
                    }
            y[r] += acc;
        }
}

/* ---- 2. FP64 reference and row bounds ------------------------------------ */

void oracle_spmv_f64(int nnz, const int* row, const int* col, const float* val,
                     const float* x, int rows, const float* y0, double* y64,
                     double* bound) {
    for (int r = 0; r < rows; ++r) {
        y64[r] = y0 ? (double)y0[r] : 0.0;
        bound[r] = y0 ? fabs((double)y0[r]) : 0.0;
    }
    for (int i = 0; i < nnz; ++i) {
        double p = (double)val[i] * (double)x[col[i]];
        y64[row[i]] += p;
        bound[row[i]] += fabs(p);
    }
}

/* Largest |y_i - y64_i| / bound_i (rows with bound 0 must match exactly:
 * they contribute +inf when they differ).  *worst receives the row. */
double oracle_scaled_error(int rows, const float* y, const double* y64,
                           const double* bound, int* worst) {
    double e = 0.0;
    int w = -1;
    for (int r = 0; r < rows; ++r) {
        double d = fabs((double)y[r] - y64[r]);
        double s;
        if (bound[r] > 0.0) s = d / bound[r];
        else s = d > 0.0 ? INFINITY : 0.0;
        if (w < 0 || s > e) {
            e = s;
            w = r;
        }
    }
    if (worst) *worst = w;
    return e;
}

/* ---- 3. two_vec_compare, src/util.cpp:93-131 ------------------------------ */
/* out[0] = L2 distance, out[1] = max |a|-|b| difference, out[2] = max ratio
 * of that difference to max(|a|,|b|), out[3] = index of the max difference,
 * out[4] = number of entries that differ. */
void oracle_two_vec_compare(const float* a, const float* b, int n, double* out) {
    double dist = 0.0, maxdiff = 0.0, maxratio = 0.0;
    int maxid = 0, count = 0;
    for (int i = 0; i < n; ++i) {
        /* Both differences are taken in float and then widened, as the
         * reference does (distance :79, two_vec_compare :108). */
        double t = (double)(float)(a[i] - b[i]);
        dist += t * t;
        float ta = a[i] < 0 ? -a[i] : a[i];
        float tb = b[i] < 0 ? -b[i] : b[i];
        double diff = (double)(float)(ta - tb);
        if (diff < 0) diff = -diff;
        float m = ta > tb ? ta : tb;
        double ratio = m > 0 ? diff / m : 0.0;
        if (diff > maxdiff) { maxdiff = diff; maxid = i; }
        if (ratio > maxratio) maxratio = ratio;
        if (a[i] != b[i]) ++count;
    }
    out[0] = sqrt(dist);
    out[1] = maxdiff;
    out[2] = maxratio;
    out[3] = (double)maxid;
    out[4] = (double)count;
}
