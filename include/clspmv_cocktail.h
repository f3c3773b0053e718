/*
 * clspmv_cocktail.h -- C-ABI of the Cocktail multi-format SpMV: the
 * partitioner (analyze_matrix) and the device runtime that runs every part
 * of a partitioned matrix into one y.
 *
 * Mirrors include/analyze.h (cocktail struct, analyze_matrix) and
 * include/spmv_cocktail.h (init_mat_kernels, do_spmv, cpy_*, evaluate) of
 * the reference.  OpenCL handles are replaced as follows:
 *   cocktail_gpu&                      -> clspmv_cocktail_gpu*     (opaque)
 *   cocktail_kernels& (+ its context
 *   and command queue arguments)       -> clspmv_cocktail_kernels* (opaque;
 *                                         owns one HIP stream on one GPU)
 *   cl_mem                             -> float* device pointer
 *   bool use_tex                       -> int, accepted and ignored (there
 *                                         is no texture path on CDNA)
 */
#ifndef CLSPMV_COCKTAIL_H
#define CLSPMV_COCKTAIL_H

#include <stdbool.h>
#include "clspmv_device.h"
#include "clspmv_formats.h"

#ifdef __cplusplus
extern "C" {
#endif

/* cocktail<int,int,float> (include/analyze.h:105-144), member for member.
 * Part arrays are malloc()ed by clspmv_analyze_matrix and released by
 * clspmv_free_cocktail. */
typedef struct {
    bool ifusebdia;
    bool ifusedia;
    bool ifusesbell;
    bool ifusebell;
    bool ifusebcsr;
    bool ifusesell;
    bool ifuseell;
    bool ifusecsr;
    bool ifusecoo;

    int bdia_meth_num;
    int dia_meth_num;
    int sbell_meth_num;
    int bell_meth_num;
    int bcsr_meth_num;
    int sell_meth_num;
    int ell_meth_num;
    int csr_meth_num;
    int coo_meth_num;

    int csr_group_num;
    int coo_group_num;

    int mat_width;
    int mat_height;
    int mat_nnz;

    clspmv_bdia_matrix  bdia;
    clspmv_dia_matrix   dia;
    clspmv_sbell_matrix sbell;
    clspmv_b4ell_matrix bell;
    clspmv_b4csr_matrix bcsr;
    clspmv_sell_matrix  sell;
    clspmv_ell_matrix   ell;
    clspmv_csr_matrix   csr;
    clspmv_coo_matrix   coo;
} clspmv_cocktail;

/* init_cocktail / free_cocktail, include/analyze.h:146-206 */
void clspmv_init_cocktail(clspmv_cocktail* mat);
void clspmv_free_cocktail(clspmv_cocktail* mat);

/* analyze_matrix, include/analyze.h:257 / src/analyze.cpp:3714.
 * Splits coomat into at most nine parts (BDIA, DIA, SBELL, BELL, BCSR,
 * SELL, ELL, CSR, COO) by the reference's cost model over per-format
 * GFLOP/s tables.  The tables are read from <ben_dir>/benchmark/<fmt>.ben
 * (the reference reads $CLSPMVPATH/benchmark, src/analyze.cpp:179); a NULL
 * ben_dir falls back to $CLSPMVPATH.  coomat must be sorted by (row, col)
 * without duplicates.  Returns CLSPMV_OK or an error (missing or empty
 * tables are reported instead of read out of bounds). */
int clspmv_analyze_matrix(clspmv_coo_matrix* coomat, clspmv_cocktail* mat,
                          float dia_priority_multiplier,
                          int estimate_flat_partial, int estimate_flat_full,
                          const char* ben_dir);

/* Optional log of the partitioner's decisions (one line per decision);
 * NULL disables.  Not thread safe. */
void clspmv_set_analyze_log(void (*log)(const char* line));

/* ---- autotuner (replaces src/bench_*.cpp and the benchmarking script) ----
 * Times every kernel variant of every format on synthetic matrices
 * generated in device memory (dimensions 1024 .. 2M rows) and writes the ten
 * tables to out_dir/benchmark/NAME.ben in the reference's grammar, ready for
 * clspmv_analyze_matrix.  level 0: up to 64M stored entries per probe and 8
 * timed launches; level 1: up to 256M entries and 20 launches; level -1: a
 * reduced grid (1024, 8192 and 65536 rows, 1M entries, 3 launches) for
 * checking the tables' form.  Each column holds the best of the knobs it
 * stands for: CSR method 0 and BCSR over lanes per row (powers of two
 * around the shape's own), CSR and COO over the format's own arrays and
 * the x-panel encoding, SELL's two columns are its slice heights 32 and
 * 128, and the block tables' rows are the eight block shapes. */
int clspmv_autotune(int device_ordinal, const char* out_dir, int level, int verbose);

/* ---- device runtime, include/spmv_cocktail.h:166-175 ------------------- */
typedef struct clspmv_cocktail_gpu clspmv_cocktail_gpu;
typedef struct clspmv_cocktail_kernels clspmv_cocktail_kernels;

/* init_mat_kernels (src/spmv_cocktail.cpp:1590): uploads every part of
 * cpumat to device `device_ordinal`, x from vec (mat_width floats) and y
 * from res (mat_height floats), and prepares the launch sequence, captured
 * once into a HIP graph.  CSR and COO parts are uploaded with
 * CLSPMV_PANEL_BOTH (the panel encoding active where it pays) until
 * clspmv_cocktail_tune keeps one.  The cocktail is not retained. */
int clspmv_init_mat_kernels(const clspmv_cocktail* cpumat, const float* vec,
                            const float* res, clspmv_cocktail_gpu** gpumat,
                            clspmv_cocktail_kernels** kernels, int use_tex);
/* Same with an explicit device. */
int clspmv_init_mat_kernels_on(int device_ordinal, const clspmv_cocktail* cpumat,
                               const float* vec, const float* res,
                               clspmv_cocktail_gpu** gpumat,
                               clspmv_cocktail_kernels** kernels);

/* do_spmv (src/spmv_cocktail.cpp:1088): ntimes passes of y += A*x over all
 * parts in the reference's order BDIA, DIA, SBELL, BELL, BCSR, SELL, ELL,
 * CSR, COO; blocks until done.  In graph mode passes go out 16 to a graph
 * launch while ntimes allows (the same kernels in the same order). */
int clspmv_do_spmv(clspmv_cocktail_kernels* kernels, int ntimes);
/* Same without the final wait (work stays queued on the kernels' stream). */
int clspmv_do_spmv_async(clspmv_cocktail_kernels* kernels, int ntimes);

/* cpy_* (src/spmv_cocktail.cpp:1459-1587) */
int clspmv_cpy_vector_from_cpu(clspmv_cocktail_gpu* gpumat, clspmv_cocktail_kernels* kernels, const float* vec, int vec_size);
int clspmv_cpy_vector_from_gpu(clspmv_cocktail_gpu* gpumat, clspmv_cocktail_kernels* kernels, const float* d_vec, int vec_size);
int clspmv_cpy_result_from_cpu(clspmv_cocktail_gpu* gpumat, clspmv_cocktail_kernels* kernels, const float* res, int res_size);
int clspmv_cpy_result_from_gpu(clspmv_cocktail_gpu* gpumat, clspmv_cocktail_kernels* kernels, const float* d_res, int res_size);
int clspmv_cpy_result_to_cpu(clspmv_cocktail_gpu* gpumat, clspmv_cocktail_kernels* kernels, float* res, int res_size);
int clspmv_cpy_result_to_gpu(clspmv_cocktail_gpu* gpumat, clspmv_cocktail_kernels* kernels, float* d_res, int res_size);

/* free_mat_kernels (src/spmv_cocktail.cpp:1613) */
int clspmv_free_mat_kernels(clspmv_cocktail_gpu* gpumat, clspmv_cocktail_kernels* kernels);

/* gpumat.vec / gpumat.res of the reference: device x and y. */
float* clspmv_cocktail_gpu_vec(clspmv_cocktail_gpu* gpumat);
float* clspmv_cocktail_gpu_res(clspmv_cocktail_gpu* gpumat);
/* The context (device + stream) the kernels run on. */
clspmv_ctx* clspmv_cocktail_kernels_ctx(clspmv_cocktail_kernels* kernels);

/* Per-part timing: each part's kernels timed alone with HIP events over
 * ntimes launches, then the whole captured pass.  Parts are reported in
 * launch order. */
typedef struct {
    int    format;     /* enum clspmv_format                   */
    int    method;
    int    nnz;        /* true nonzeros of the part            */
    long long stored;  /* stored entries of the part           */
    double ms;         /* average per launch of this part      */
} clspmv_part_timing;

int clspmv_cocktail_profile(clspmv_cocktail_kernels* kernels, int ntimes,
                            clspmv_part_timing* parts, int max_parts,
                            int* nparts, double* graph_ms);

/* Cold-cache timing (no counterpart in the reference): ntimes passes in the
 * kernels' launch mode, each after a read-only stream over a buffer of its
 * own of scrub_bytes (evicting the L2s and the memory-side cache without
 * leaving write-backs behind), each pass alone between HIP events;
 * *ms_per_pass is the average.  y is preserved.  Blocks. */
int clspmv_cocktail_time_cold(clspmv_cocktail_kernels* kernels, int ntimes, size_t scrub_bytes,
                              double* ms_per_pass);

/* Measured launch tuning, the device-side counterpart of the reference's
 * runtime variant probe (init_mat_kernels with use_tex,
 * src/spmv_cocktail.cpp:1590): every method a part's shape allows (and,
 * for CSR method 0 and BCSR, the neighbouring lane counts) is timed with
 * HIP events over ntimes whole passes (best of three such batches; the
 * other parts run their current choices) and the fastest kept (CSR and COO
 * parts
 * are uploaded in both their own and the x-panel encoding, and the losing
 * encoding is freed); then the captured
 * graphs are re-built, and graph launches (as clspmv_do_spmv issues them)
 * and plain stream launches of the same pass are timed, each with the x-panel first passes on a second
 * stream (clspmv_cocktail_set_overlap), on one stream, and with a DIA part
 * fused into a first-pass launch (clspmv_cocktail_set_fuse); the fastest is
 * kept.  y is preserved.  Tuning picks
 * by timing, so two processes may settle on different methods whose
 * results differ by rounding; untuned kernels follow the tables' choice.
 * `graph` (may be null) receives 1 when passes run as a graph launch. */
int clspmv_cocktail_tune(clspmv_cocktail_kernels* kernels, int ntimes, int* graph);
/* Sets the launch mode by hand: 1 = one graph launch per pass (the default
 * after init), 0 = plain launches of the parts' kernels. */
int clspmv_cocktail_set_graph(clspmv_cocktail_kernels* kernels, int graph);
/* Sets whether the x-panel parts' first pass (which reads only x) runs on
 * a second stream beside the parts launched before them, the default after
 * init; 0 = every kernel on one stream.  Either way each row of y receives
 * the parts' contributions in the same order, so results are identical.
 * Ignored while a DIA part is fused (clspmv_cocktail_set_fuse).
 * clspmv_cocktail_tune times both and keeps the faster.  Re-builds the
 * graph. */
int clspmv_cocktail_set_overlap(clspmv_cocktail_kernels* kernels, int overlap);
/* Sets whether an FP32 DIA part directly followed by an x-panel part is
 * summed by extra work-groups of that part's first-pass launch (filling the
 * first pass's ragged end) instead of its own launch; off after init.  The
 * DIA rows then get DIA method 1's arithmetic whatever the part's method,
 * and still reach y before the x-panel part's.  clspmv_cocktail_tune times
 * it with the other modes.  Re-builds the graph. */
int clspmv_cocktail_set_fuse(clspmv_cocktail_kernels* kernels, int fuse);
/* The launch modes in force: *graph = 1 for one graph launch per pass;
 * *overlap = 1 when a pass actually runs a first pass on the second stream
 * (overlap set, and an x-panel part follows another part); *fuse = 1 when
 * a DIA part rides in a first-pass launch.  Any pointer may be null. */
int clspmv_cocktail_launch_modes(const clspmv_cocktail_kernels* kernels, int* graph, int* overlap, int* fuse);

/* evaluate (src/spmv_cocktail.cpp:1431): runs one SpMV from res, compares
 * with coores (the caller's expected y, e.g. from a serial COO SpMV) using
 * the reference's report (src/util.cpp:93), then times 1000 passes and
 * prints "Cocktail ... GFLOPS" with GFLOP/s = 2*nnz/t. */
typedef struct {
    double ms_per_spmv;
    double gflops;
    double l2_distance;
    double max_diff;
    double max_ratio;
} clspmv_eval_report;
int clspmv_evaluate(const clspmv_cocktail* cpumat, const float* vec, float* res,
                    const float* coores, clspmv_eval_report* report);

#ifdef __cplusplus
}
#endif
#endif /* CLSPMV_COCKTAIL_H */
