/*
 * clspmv_device.h -- C-ABI of the MI355X device runtime: contexts (one HIP
 * stream on one GPU), device-resident matrices of every storage format, the
 * per-format SpMV entry points and the per-format benchmark drivers.
 *
 * Every SpMV computes y += A*x (the reference's contract: each kernel
 * accumulates into result, kernels/spmv_*.cl; the caller preloads y,
 * src/eval.cpp:40-42).  Device pointers are plain float* obtained from
 * clspmv_device_malloc, hipMalloc or any other allocator of the same device.
 * Launch entry points are asynchronous on the context's stream; the driver
 * entry points block, like the reference's clFinish-terminated calls.
 */
#ifndef CLSPMV_DEVICE_H
#define CLSPMV_DEVICE_H

#include <stddef.h>
#include "clspmv_formats.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- contexts: replace cl_context + cl_command_queue -------------------
 * (include/oclcommon.h, src/oclcommon.cpp:259 initialization()).  One
 * context = one device + one in-order stream at default priority. */
typedef struct clspmv_ctx clspmv_ctx;
int   clspmv_ctx_create(int device_ordinal, clspmv_ctx** out);
int   clspmv_ctx_destroy(clspmv_ctx* ctx);
int   clspmv_ctx_synchronize(clspmv_ctx* ctx);
void* clspmv_ctx_stream(clspmv_ctx* ctx);       /* the hipStream_t */
int   clspmv_ctx_device(clspmv_ctx* ctx);
/* Name and compute-unit count of the context's device. */
int   clspmv_device_info(clspmv_ctx* ctx, char* name, int name_len, int* cu_count);

int clspmv_device_malloc(clspmv_ctx* ctx, size_t bytes, void** out);
int clspmv_device_free(clspmv_ctx* ctx, void* ptr);
int clspmv_memcpy_h2d(clspmv_ctx* ctx, void* dst, const void* src, size_t bytes);
int clspmv_memcpy_d2h(clspmv_ctx* ctx, void* dst, const void* src, size_t bytes);
int clspmv_memcpy_d2d(clspmv_ctx* ctx, void* dst, const void* src, size_t bytes);
int clspmv_memset_zero(clspmv_ctx* ctx, void* dst, size_t bytes);

/* ---- device-resident matrices ------------------------------------------ */
enum clspmv_format {
    CLSPMV_FMT_CSR   = 0,   /* csr_matrix   -> kernels/spmv_csr_*.cl   */
    CLSPMV_FMT_COO   = 1,   /* coo_matrix   -> kernels/spmv_coo.cl     */
    CLSPMV_FMT_ELL   = 2,   /* ell_matrix   -> kernels/spmv_ell.cl     */
    CLSPMV_FMT_SELL  = 3,   /* sell_matrix  -> kernels/spmv_sell.cl    */
    CLSPMV_FMT_DIA   = 4,   /* dia_matrix   -> kernels/spmv_dia.cl     */
    CLSPMV_FMT_BDIA  = 5,   /* bdia_matrix  -> kernels/spmv_bdia.cl    */
    CLSPMV_FMT_BCSR  = 6,   /* b4csr_matrix -> kernels/spmv_bcsr.cl    */
    CLSPMV_FMT_BELL  = 7,   /* b4ell_matrix -> kernels/spmv_bell.cl    */
    CLSPMV_FMT_SBELL = 8,   /* sbell_matrix -> kernels/spmv_sbell.cl   */
    CLSPMV_FMT_BSPMM = 9    /* 4x4-block BCSR, 16-bit values: SpMM only */
};

/* Kernel variants ("methods") per format.  Method numbers index the
 * per-format GFLOP/s columns of the .ben tables read by the Cocktail
 * partitioner, so the counts equal the reference's *_IMP_NUM.
 *   CSR  0: vector kernel, lanes per row = power of two near nnz/row
 *        1: adaptive row blocks (short rows share an LDS-staged block,
 *           long rows take a whole work-group)
 *   COO  0: one-pass segmented reduction + carry fix-up
 *   ELL  0: one row per lane          1: four rows per lane (128-bit loads)
 *   SELL 0/1: one row per lane (slice height is a property of the data)
 *   DIA  0: row per lane              1: two rows per lane (64-bit loads)
 *        2: row per lane, x window in LDS   3: four rows, x window in LDS
 *   BDIA 0: row per lane              1: row per lane, LDS x window per band
 *        2: two rows per lane, LDS    3: two rows per lane (64-bit loads)
 *        4: four rows per lane, LDS x window per band
 *   BCSR 0: lane group per block row (group size from blocks per row),
 *           or, in the slot-major encoding, one lane per block row
 *   BELL 0: lane per block row, scalar FMA   1: float4 accumulators
 *   SBELL 0: lane per block row within a slice                       */
#define CLSPMV_METHOD_AUTO (-1)
int clspmv_method_count(int format);

typedef struct clspmv_dmat clspmv_dmat;

typedef struct {
    int    format;          /* enum clspmv_format                       */
    int    rows, cols;      /* matinfo.height, matinfo.width            */
    int    nnz;             /* true nonzeros (matinfo.nnz)              */
    long long stored;       /* stored entries incl. padding             */
    size_t device_bytes;    /* device memory held by the matrix         */
    int    bheight, bwidth; /* block shape (block formats), else 1,1    */
    int    slice_height;    /* SELL / SBELL slice height, else 0        */
    int    panel_width;     /* COO / CSR x-panel encoding width, or 0   */
    int    panel_active;    /* 1 when launches use the panel encoding   */
    int    plain;           /* 1 when the format's own arrays are held  */
    int    dtype;           /* enum clspmv_dtype of the stored values    */
    int    panel_groups;    /* panel encoding: pass-1 work-groups        */
    int    panel_subchunks; /* panel encoding: 512-entry sub-chunks      */
    int    panel_blocks;    /* panel encoding: pass-2 row blocks         */
    int    block_slots;     /* BCSR slot-major encoding: blocks per block
                               row after padding, or 0 (row pointers)     */
    int    spmm_window;     /* block SpMM: most distinct block columns in
                               one 16-block-row slice when the X windows
                               were built (tiled X then staged in LDS),
                               else 0                                      */
} clspmv_dmat_info;

/* Upload a host matrix (the reference's cpy_<fmt>_mat,
 * src/spmv_cocktail.cpp:304-962).  The host struct is not retained. */
int clspmv_upload_csr(clspmv_ctx* ctx, const clspmv_csr_matrix* m, clspmv_dmat** out);
int clspmv_upload_coo(clspmv_ctx* ctx, const clspmv_coo_matrix* m, clspmv_dmat** out);
int clspmv_upload_ell(clspmv_ctx* ctx, const clspmv_ell_matrix* m, clspmv_dmat** out);
int clspmv_upload_sell(clspmv_ctx* ctx, const clspmv_sell_matrix* m, clspmv_dmat** out);
int clspmv_upload_dia(clspmv_ctx* ctx, const clspmv_dia_matrix* m, clspmv_dmat** out);
int clspmv_upload_bdia(clspmv_ctx* ctx, const clspmv_bdia_matrix* m, clspmv_dmat** out);
int clspmv_upload_b4csr(clspmv_ctx* ctx, const clspmv_b4csr_matrix* m, clspmv_dmat** out);
int clspmv_upload_b4ell(clspmv_ctx* ctx, const clspmv_b4ell_matrix* m, clspmv_dmat** out);
int clspmv_upload_sbell(clspmv_ctx* ctx, const clspmv_sbell_matrix* m, clspmv_dmat** out);
int clspmv_dmat_free(clspmv_dmat* m);

/* Uploads with a choice of device encoding and value type (no counterpart
 * in the reference).
 *
 * panel_width (COO and CSR): besides the format's own arrays the engine has
 * an x-panel encoding for matrices whose column indices have no locality:
 * columns are cut into panels of panel_width floats (a multiple of 8, at
 * most 32760) so that a work-group holds its panel of x in LDS, and entries
 * are stored as 16-bit in-panel columns with segment marks (6 bytes per
 * nonzero with FP32 values, 4 with 16-bit values).  Results are the same to
 * rounding; the summation order is fixed by the encoding.
 *   0                        the format's own arrays only
 *   > 0                      the panel encoding only, that width
 *   CLSPMV_PANEL_AUTO (-1)   panels when the matrix has enough nonzeros
 *                            per panel, else the own arrays
 *   CLSPMV_PANEL_BOTH (-2)   both encodings where panels pay (panel
 *                            active); clspmv_dmat_set_panel switches
 * With the panel encoding active, every method number of the format runs
 * the panel kernels.
 *
 * dtype: CLSPMV_DT_F32 (the reference's values), or CLSPMV_DT_BF16 /
 * CLSPMV_DT_F16: values rounded to 16 bits (nearest even) and widened to
 * FP32 in the kernels, which accumulate in FP32 as before (FP64 for rows
 * longer than 4096 entries on the CSR and panel paths).  16-bit values
 * exist for DIA (every method, at most 4096 diagonals), for the COO / CSR
 * panel encoding (COO: AUTO then always uses panels, BOTH is refused) and
 * for CSR's own arrays (every method runs the multi-vector kernel with
 * k = 1), alone (panel_width 0) or beside the panels (BOTH).
 *
 * clspmv_upload_coo is clspmv_upload_coo_ex with {CLSPMV_PANEL_AUTO, F32},
 * clspmv_upload_csr is clspmv_upload_csr_ex with {0, F32}, clspmv_upload_dia
 * is clspmv_upload_dia_ex with F32; opts = NULL means those defaults. */
enum clspmv_dtype { CLSPMV_DT_F32 = 0, CLSPMV_DT_BF16 = 1, CLSPMV_DT_F16 = 2 };
#define CLSPMV_PANEL_AUTO (-1)
#define CLSPMV_PANEL_BOTH (-2)
typedef struct {
    int panel_width;
    int dtype;
    int panel_group;  /* sub-chunks (512 entries) per pass-1 work-group: 8..512,
                         or 0 to size the work-groups so that they fill the
                         GPU's CUs in whole rounds */
} clspmv_upload_opts;
int clspmv_upload_coo_ex(clspmv_ctx* ctx, const clspmv_coo_matrix* m,
                         const clspmv_upload_opts* opts, clspmv_dmat** out);
int clspmv_upload_csr_ex(clspmv_ctx* ctx, const clspmv_csr_matrix* m,
                         const clspmv_upload_opts* opts, clspmv_dmat** out);
int clspmv_upload_dia_ex(clspmv_ctx* ctx, const clspmv_dia_matrix* m,
                         const clspmv_upload_opts* opts, clspmv_dmat** out);
/* Selects the panel encoding (1) or the format's own arrays (0) of a matrix
 * that holds both. */
int clspmv_dmat_set_panel(clspmv_dmat* m, int use_panel);
int clspmv_dmat_get_info(const clspmv_dmat* m, clspmv_dmat_info* out);

/* y += A*x on the context's stream (asynchronous).  x holds cols floats,
 * y holds rows floats.  method = CLSPMV_METHOD_AUTO picks by shape. */
int clspmv_spmv(clspmv_ctx* ctx, const clspmv_dmat* A, int method,
                const float* d_x, float* d_y);

/* Multi-vector SpMV (no counterpart in the reference): Y[:, q] += A X[:, q]
 * for q < k (k <= 16) in one pass over the matrix.  X is row-major: the k
 * values of column c are X[c*k .. c*k+k) (cols * k floats), so one gather
 * serves every right-hand side; Y is column-major with leading dimension
 * ldy >= rows.  DIA matrices (any value type) and CSR matrices holding
 * their own arrays (FP32 or 16-bit values) only.  Asynchronous. */
int clspmv_spmv_multi(clspmv_ctx* ctx, const clspmv_dmat* A, int k, const float* d_X,
                      float* d_Y, int ldy);

/* Times ntimes back-to-back launches of clspmv_spmv after warmup
 * untimed ones, with HIP events on the context's stream; y keeps
 * accumulating.  Blocks.  *ms_per_call = average kernel-to-kernel time. */
int clspmv_time_spmv(clspmv_ctx* ctx, const clspmv_dmat* A, int method,
                     const float* d_x, float* d_y, int warmup, int ntimes,
                     double* ms_per_call);

/* The same for clspmv_spmv_multi over k right-hand sides (no counterpart in
 * the reference): ntimes back-to-back passes after warmup untimed ones, HIP
 * events on the context's stream, Y accumulating.  Blocks. */
int clspmv_time_spmv_multi(clspmv_ctx* ctx, const clspmv_dmat* A, int k, const float* d_X,
                           float* d_Y, int ldy, int warmup, int ntimes, double* ms_per_call);

/* ---- mixed-precision block SpMM (no counterpart in the reference) -------
 * Y += A X for k right-hand sides with A's values stored in BF16 or FP16
 * and the products accumulated in FP32 on the matrix cores (MFMA 4x4x4: one
 * 4x4 block of A times a 4x4 tile of X per group of four lanes).
 *
 * clspmv_upload_bcsr_lowp: A from a bcsr_matrix with 4x4 blocks
 *   (include/matrix_storage.h:178, from clspmv_coo2bcsr(..., 4, 4, ...)),
 *   values rounded to `dtype` (round to nearest even).  The result has
 *   format CLSPMV_FMT_BSPMM and runs only through clspmv_spmm.
 * clspmv_spmm: X holds cols x k values of the same dtype, column-major with
 *   leading dimension ldx; Y holds rows x k FP32 values, column-major with
 *   leading dimension ldy.  ldx >= cols and ldy >= rows, both multiples of
 *   4; X 8-byte and Y 16-byte aligned.  Asynchronous on the context's
 *   stream.  Each Y element is written by one lane: no atomics, a fixed
 *   summation order.
 * clspmv_spmm_tiled: the same product with X in 4-row tiles: the values of
 *   rows 4c..4c+3 of right-hand side j at X4[(c * k + j) * 4 + 0..3], for
 *   c < ceil(cols / 4), zeros past the last column (8-byte aligned).  A
 *   lane group then reads one contiguous run per 4 right-hand sides instead
 *   of 4 columns, a quarter of the cache lines of the column-major form;
 *   the results are identical to clspmv_spmm's (same products, same order).
 * clspmv_spmm_pack_x: column-major X (k columns of `cols` values, leading
 *   dimension ldx, a multiple of 4) into that tiled layout, on the GPU
 *   (asynchronous on the context's stream).
 * clspmv_convert_from_f32 / clspmv_convert_to_f32: host conversions of n
 *   values between float and dtype (round to nearest even), for building X
 *   and for exact references. */
int clspmv_upload_bcsr_lowp(clspmv_ctx* ctx, const clspmv_bcsr_matrix* m, int dtype,
                            clspmv_dmat** out);
int clspmv_spmm(clspmv_ctx* ctx, const clspmv_dmat* A, int k, const void* d_X, int ldx,
                float* d_Y, int ldy);
int clspmv_spmm_tiled(clspmv_ctx* ctx, const clspmv_dmat* A, int k, const void* d_X4, float* d_Y, int ldy);
/* clspmv_spmm_tiled with accumulate = 1, or, with accumulate = 0, Y = A X:
 * Y's previous contents are not read (the result equals a zeroed Y plus
 * A X bit for bit), which saves Y's read traffic. */
int clspmv_spmm_tiled_ex(clspmv_ctx* ctx, const clspmv_dmat* A, int k, const void* d_X4, float* d_Y, int ldy,
                         int accumulate);
int clspmv_spmm_pack_x(clspmv_ctx* ctx, int k, int cols, const void* d_X, int ldx, void* d_X4);
int clspmv_convert_from_f32(const float* src, void* dst, size_t n, int dtype);
int clspmv_convert_to_f32(const void* src, float* dst, size_t n, int dtype);

/* ---- verification and per-matrix error (src/util.cpp:85-131) ------------
 * The reference recomputes y = A x serially on the host (spmv_only,
 * src/util.cpp:85) and compares with two_vec_compare (:93-131), from
 * src/eval.cpp:79-97.  Here the check runs on the GPU in FP64 over an FP32
 * CSR reference `ref` (clspmv_upload_csr of the original values; it must
 * hold its own arrays: panel off or CLSPMV_PANEL_BOTH): per row
 *     y64 = y0 + sum double(a) * double(x),  bound = |y0| + sum |a x|,
 *     e   = |y - y64| / bound   (0 / +inf when bound is 0; NaN -> +inf),
 * the row-scaled metric of the FP32 tolerance.  Blocks until done; the
 * report is the same on every run (fixed reduction order).
 *
 * clspmv_verify_spmv: scores a result d_y of y = y0 + A x that the caller
 *   computed (d_y0 may be NULL for y0 = 0).
 * clspmv_dmat_error_report: the error of a device matrix A (any format and
 *   value type: FP32, BF16 or FP16 storage) against the FP64 SpMV of `ref`
 *   (the same matrix before rounding), on x_i = uniform(seed, 0, i) (the
 *   values of clspmv_request_x): y = A x by clspmv_spmv with the automatic
 *   method, or by clspmv_spmm with k = 1 for CLSPMV_FMT_BSPMM (x rounded to
 *   the matrix's 16-bit type, so that rounding is part of the error). */
typedef struct {
    double    max_scaled_error;  /* max_i e_i                             */
    double    rms_scaled_error;  /* sqrt(mean_i e_i^2)                    */
    double    max_abs_error;     /* max_i |y_i - y64_i|                   */
    int       worst_row;         /* first row reaching max_scaled_error   */
    int       rows;
    long long rows_over_tol;     /* rows with e_i > tol                   */
    double    tol;
} clspmv_error_report;

int clspmv_verify_spmv(clspmv_ctx* ctx, const clspmv_dmat* ref, const float* d_x, const float* d_y0,
                       const float* d_y, double tol, clspmv_error_report* rep);
int clspmv_dmat_error_report(clspmv_ctx* ctx, const clspmv_dmat* A, const clspmv_dmat* ref,
                             unsigned long long seed, double tol, clspmv_error_report* rep);

/* clspmv_verify_spmv_noalloc: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc): clspmv_verify_spmv's check with no
 * allocation, no memset and no host synchronisation.  The report (same
 * values as clspmv_verify_spmv's) is written to d_report, a device pointer
 * to one clspmv_error_report, asynchronously on the context's stream.
 * The per-work-group results live in a fixed device array inside the
 * library, one per GPU, sized for CLSPMV_NOALLOC_MAX_ROW_BLOCKS row blocks
 * of `ref` (32 bytes each, 4 MiB; a CSR matrix of r rows and z entries has
 * at most r/256 + 3z/1024 + 1 blocks; the 27-point stencil
 * on a 128^3 grid has 55,586).  Larger references return CLSPMV_ERR_TOO_LARGE.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (two
 * contexts' streams, or threads) race and must be ordered by the caller. */
#define CLSPMV_NOALLOC_MAX_ROW_BLOCKS 131072
int clspmv_verify_spmv_noalloc(clspmv_ctx* ctx, const clspmv_dmat* ref, const float* d_x, const float* d_y0,
                               const float* d_y, double tol, clspmv_error_report* d_report);

/* ---- bandwidth probe (src/mem_bandwidth.cpp bandwidth_test :14) ---------
 * Copies `bytes` (rounded down to 16) between two fresh device buffers
 * `ntimes` times after one warm-up copy and returns the average copy time
 * and GB/s counted as bytes read + bytes written. */
int clspmv_bandwidth_probe(clspmv_ctx* ctx, size_t bytes, int ntimes,
                           double* ms_per_copy, double* gbps);

/* Read probe (no counterpart in the reference): streams `bytes` (rounded
 * down to 16) once per call with 128-bit non-temporal loads, each thread
 * writing one float; the fastest of a few launch shapes, as GB/s read --
 * the ceiling of kernels that mostly read.  clspmv_bandwidth_probe also
 * reports the fastest of its one-float4-per-thread copy and unrolled
 * copies with non-temporal stores over the same shapes. */
int clspmv_read_probe(clspmv_ctx* ctx, size_t bytes, int ntimes, double* ms_per_read, double* gbps);

/* Random-gather probe: n reads of x[hash(i) % nx] (no locality) written to
 * a dense y.  Returns the average kernel time and billions of gathers per
 * second: the ceiling of SpMV kernels whose column indices are random. */
int clspmv_gather_probe(clspmv_ctx* ctx, int nx, long long n, int ntimes,
                        double* ms_per_call, double* ggathers_per_s);

/* ---- per-format drivers (include/spmv_<fmt>.h:7) ------------------------
 * spmv_<fmt>(char* oclfilename, coo_matrix<int,float>* mat, int dim2Size,
 *            int ntimes, cl_device_type) -- the reference converts the COO
 * matrix, runs every kernel variant, times ntimes launches of each and
 * prints the best (e.g. src/spmv_bdia.cpp:413-440).  Here kernel_file is
 * accepted and ignored (kernels are compiled in), device_ordinal replaces
 * the OpenCL device type, and the outcome is also returned in report.
 * y_out (optional, rows floats) receives A*x of the best variant for
 * x[i] = 1 (the all-ones vector of src/eval.cpp:119), y starting at zero.
 * As in the reference, every variant is checked against the serial COO
 * SpMV on the host and the comparison printed (src/spmv_bdia.cpp:118);
 * the report carries it and an FP64 tolerance check.  The COO matrix is
 * sorted in place when it is not (every reference conversion does so).
 * Block formats sweep the eight reference block shapes. */
typedef struct {
    int    format;
    int    variants;              /* variants run                        */
    int    best_method;
    int    best_bheight, best_bwidth;
    double best_ms;               /* per launch                          */
    double best_gflops;           /* 2*nnz / t                           */
    double ms[64];                /* per variant, in run order           */
    int    method[64];
    int    bheight[64], bwidth[64];
    double max_rel_diff;          /* largest relative difference between
                                     any variant's y and the first's    */
    /* The reference's check of each variant against the serial COO SpMV
     * on the host (spmv_only + two_vec_compare, src/util.cpp:85-131):
     * largest distance, max diff and max diff ratio over the variants. */
    double check_distance, check_max_diff, check_max_ratio;
    /* Largest row-scaled error |y - y64| / (sum_j |a_ij x_j|) of any
     * variant against an FP64 host product, and the number of variants
     * above 1e-5. */
    double max_scaled_error;
    int    variants_over_tol;
} clspmv_driver_report;

int clspmv_spmv_csr_scalar(const char* kernel_file, clspmv_coo_matrix* mat, int dim2Size, int ntimes, int device_ordinal, clspmv_driver_report* report, float* y_out);
int clspmv_spmv_csr_vector(const char* kernel_file, clspmv_coo_matrix* mat, int dim2Size, int ntimes, int device_ordinal, clspmv_driver_report* report, float* y_out);
int clspmv_spmv_coo(const char* kernel_file, clspmv_coo_matrix* mat, int dim2Size, int ntimes, int device_ordinal, clspmv_driver_report* report, float* y_out);
int clspmv_spmv_ell(const char* kernel_file, clspmv_coo_matrix* mat, int dim2Size, int ntimes, int device_ordinal, clspmv_driver_report* report, float* y_out);
int clspmv_spmv_sell(const char* kernel_file, clspmv_coo_matrix* mat, int dim2Size, int ntimes, int device_ordinal, clspmv_driver_report* report, float* y_out);
int clspmv_spmv_dia(const char* kernel_file, clspmv_coo_matrix* mat, int dim2Size, int ntimes, int device_ordinal, clspmv_driver_report* report, float* y_out);
int clspmv_spmv_bdia(const char* kernel_file, clspmv_coo_matrix* mat, int dim2Size, int ntimes, int device_ordinal, clspmv_driver_report* report, float* y_out);
int clspmv_spmv_bell(const char* kernel_file, clspmv_coo_matrix* mat, int dim2Size, int ntimes, int device_ordinal, clspmv_driver_report* report, float* y_out);
int clspmv_spmv_sbell(const char* kernel_file, clspmv_coo_matrix* mat, int dim2Size, int ntimes, int device_ordinal, clspmv_driver_report* report, float* y_out);
int clspmv_spmv_bcsr(const char* kernel_file, clspmv_coo_matrix* mat, int dim2Size, int ntimes, int device_ordinal, clspmv_driver_report* report, float* y_out);

#ifdef __cplusplus
}
#endif
#endif /* CLSPMV_DEVICE_H */
