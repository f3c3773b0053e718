/*
 * clspmv_formats.h -- C-ABI of the host-side storage formats and COO<->X
 * conversions of the MI355X SpMV engine.
 *
 * Every struct below is a C-layout twin of the reference's <int,float>
 * (or <int,int,float>) template instantiation in include/matrix_storage.h:
 * same members, same order, same types, so a caller that built the reference
 * structs can hand them over by pointer.  Arrays are allocated with malloc()
 * by the conversion routines and released with free() by the clspmv_free_*
 * routines, exactly like the reference's init and free routines
 * (include/matrix_storage.h:370-652).
 *
 * Conventions shared by every entry point:
 *   - Return values that were bool in the reference are int 1/0 here.
 *   - Routines that were void in the reference return a status code
 *     (CLSPMV_OK == 0, negative on error) instead of asserting.
 *   - All indices are 32-bit int, as in the reference (nnz < 2^31).
 *   - A conversion that sorts its COO source does so in place, like the
 *     reference (include/matrix_storage.h:1271-1274).
 */
#ifndef CLSPMV_FORMATS_H
#define CLSPMV_FORMATS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CLSPMV_OK                 0
#define CLSPMV_ERR_INVALID       -1  /* bad argument or malformed matrix   */
#define CLSPMV_ERR_HIP           -2  /* a HIP runtime call failed          */
#define CLSPMV_ERR_ALLOC         -3  /* host or device allocation failed   */
#define CLSPMV_ERR_UNSUPPORTED   -4  /* method / shape not implemented     */
#define CLSPMV_ERR_TOO_LARGE     -5  /* exceeds a documented size limit    */
#define CLSPMV_ERR_IO            -6  /* file could not be read             */

/* Text of the last error raised on the calling thread ("" if none). */
const char* clspmv_last_error(void);

/* ---- compile-time constants kept from include/constant.h:1-55 ---------- */
#define CLSPMV_GPU_ALIGNMENT      32
#define CLSPMV_COO_ALIGNMENT      32
#define CLSPMV_WARPSIZE           32   /* slice height of "warp" SELL/SBELL */
#define CLSPMV_WORK_GROUP_SIZE    256
#define CLSPMV_MAX_BAND_NUM       31
#define CLSPMV_MAX_BAND_WIDTH     128
#define CLSPMV_MAX_DIA_NUM        128
#define CLSPMV_CSR_VEC_GROUP_SIZE 64
#define CLSPMV_BCSR_VEC_GROUP_SIZE 128
#define CLSPMV_BELL_GROUP_SIZE    256
#define CLSPMV_SELL_GROUP_SIZE    128
#define CLSPMV_COO_GROUP_SIZE     64
#define CLSPMV_BDIA_IMP_NUM       5
#define CLSPMV_DIA_IMP_NUM        4
#define CLSPMV_SBELL_IMP_NUM      1
#define CLSPMV_BELL_IMP_NUM       2
#define CLSPMV_BCSR_IMP_NUM       1
#define CLSPMV_SELL_IMP_NUM       2
#define CLSPMV_ELL_IMP_NUM        2
#define CLSPMV_CSR_IMP_NUM        2
#define CLSPMV_COO_IMP_NUM        1
/* The reference refuses block formats above 256 MiB (include/constant.h:10);
 * the same limit is applied by the b4csr/b4ell/sbell converters so their
 * success/failure matches.  The device runtime has no such limit. */
#define CLSPMV_MAX_MEM_OBJ        268435456

/* ---- structs (include/matrix_storage.h:15-365) ------------------------- */
typedef struct { int width; int height; int nnz; } clspmv_matinfo;   /* :15 */

typedef struct {                                   /* dia_matrix :30 */
    clspmv_matinfo matinfo;
    int   dia_num;
    int   dia_length;
    int   dia_length_aligned;
    int*  dia_offsets;
    float* dia_data;          /* dia_data[j*dia_length_aligned + row] */
} clspmv_dia_matrix;

typedef struct {                                   /* dia_ext_matrix :52 */
    clspmv_matinfo matinfo;
    int   dia_num;
    int   dia_length;
    int   dia_length_aligned;
    int*  dia_offsets;        /* offsets taken modulo the width, >= 0 */
    float* dia_data;
} clspmv_dia_ext_matrix;

typedef struct {                                   /* bdia_matrix :72 */
    clspmv_matinfo matinfo;
    int   bdia_band_num;
    int   bdia_length;
    int   bdia_length_aligned;
    int*  bdia_offsets;       /* first offset of each band               */
    int*  bdia_bptr;          /* band b holds diagonals [bptr[b],bptr[b+1]) */
    float* bdia_data;
} clspmv_bdia_matrix;

typedef struct {                                   /* ell_matrix :96 */
    clspmv_matinfo matinfo;
    int   ell_num;
    int   ell_height_aligned;
    int*  ell_col_id;         /* [slot*ell_height_aligned + row] */
    float* ell_data;
} clspmv_ell_matrix;

typedef struct {                                   /* sell_matrix :122 */
    clspmv_matinfo matinfo;
    unsigned int sell_slice_height;
    int   sell_slice_num;
    int*  sell_slice_ptr;     /* element offsets, size slice_num+1 */
    int*  sell_col_id;        /* [ptr[s] + slot*h + row_in_slice] */
    float* sell_data;
} clspmv_sell_matrix;

typedef struct {                                   /* coo_matrix :141 */
    clspmv_matinfo matinfo;
    int*  coo_row_id;
    int*  coo_col_id;
    float* coo_data;
} clspmv_coo_matrix;

typedef struct {                                   /* csr_matrix :155 */
    clspmv_matinfo matinfo;
    int*  csr_row_ptr;
    int*  csr_col_id;
    float* csr_data;
} clspmv_csr_matrix;

typedef struct {                                   /* bcsr_matrix :178 */
    clspmv_matinfo matinfo;
    unsigned int bcsr_bwidth;
    unsigned int bcsr_bheight;
    int   bcsr_row_num;
    int   bcsr_aligned_size;
    int   bcsr_block_num;
    int*  bcsr_row_ptr;
    int*  bcsr_col_id;        /* element column of the block's left edge */
    float* bcsr_data;         /* [k*bcsr_aligned_size + block], k = h*bw+w */
} clspmv_bcsr_matrix;

typedef struct {                                   /* bell_matrix :215 */
    clspmv_matinfo matinfo;
    unsigned int bell_bwidth;
    unsigned int bell_bheight;
    int   bell_row_num;
    int   bell_block_num;
    int   bell_height_aligned;
    int   bell_float4_aligned;
    int*  bell_col_id;
    float* bell_data;
} clspmv_bell_matrix;

typedef struct {                                   /* b4csr_matrix :266 */
    clspmv_matinfo matinfo;
    unsigned int b4csr_bwidth;
    unsigned int b4csr_bheight;
    int   b4csr_row_num;
    int   b4csr_aligned_size;
    int   b4csr_block_num;
    int*  b4csr_row_ptr;
    int*  b4csr_col_id;       /* block column in float4 units */
    float* b4csr_data;        /* [w4*(aligned*bh) + h*aligned + blk*4 + w] */
} clspmv_b4csr_matrix;

typedef struct {                                   /* b4ell_matrix :311 */
    clspmv_matinfo matinfo;
    unsigned int b4ell_bwidth;
    unsigned int b4ell_bheight;
    int   b4ell_row_num;
    int   b4ell_block_num;
    int   b4ell_height_aligned;
    int   b4ell_float4_aligned;
    int*  b4ell_col_id;       /* [brow + slot*height_aligned], float4 units */
    float* b4ell_data;
} clspmv_b4ell_matrix;

typedef struct {                                   /* sbell_matrix :344 */
    clspmv_matinfo matinfo;
    unsigned int sbell_bwidth;
    unsigned int sbell_bheight;
    unsigned int sbell_slice_height;
    int   sbell_slice_num;
    int   sbell_row_num;
    int*  sbell_slice_ptr;    /* block offsets, size slice_num+1 */
    int*  sbell_col_id;       /* float4 units */
    float* sbell_data;
} clspmv_sbell_matrix;

/* ---- init / free (include/matrix_storage.h:370-652) -------------------- */
void clspmv_init_coo_matrix(clspmv_coo_matrix* m);
void clspmv_init_csr_matrix(clspmv_csr_matrix* m);
void clspmv_init_ell_matrix(clspmv_ell_matrix* m);
void clspmv_init_sell_matrix(clspmv_sell_matrix* m);
void clspmv_init_dia_matrix(clspmv_dia_matrix* m);
void clspmv_init_dia_ext_matrix(clspmv_dia_ext_matrix* m);
void clspmv_init_bdia_matrix(clspmv_bdia_matrix* m);
void clspmv_init_bcsr_matrix(clspmv_bcsr_matrix* m);
void clspmv_init_bell_matrix(clspmv_bell_matrix* m);
void clspmv_init_b4csr_matrix(clspmv_b4csr_matrix* m);
void clspmv_init_b4ell_matrix(clspmv_b4ell_matrix* m);
void clspmv_init_sbell_matrix(clspmv_sbell_matrix* m);
void clspmv_free_coo_matrix(clspmv_coo_matrix* m);
void clspmv_free_csr_matrix(clspmv_csr_matrix* m);
void clspmv_free_ell_matrix(clspmv_ell_matrix* m);
void clspmv_free_sell_matrix(clspmv_sell_matrix* m);
void clspmv_free_dia_matrix(clspmv_dia_matrix* m);
void clspmv_free_dia_ext_matrix(clspmv_dia_ext_matrix* m);
void clspmv_free_bdia_matrix(clspmv_bdia_matrix* m);
void clspmv_free_bcsr_matrix(clspmv_bcsr_matrix* m);
void clspmv_free_bell_matrix(clspmv_bell_matrix* m);
void clspmv_free_b4csr_matrix(clspmv_b4csr_matrix* m);
void clspmv_free_b4ell_matrix(clspmv_b4ell_matrix* m);
void clspmv_free_sbell_matrix(clspmv_sbell_matrix* m);

/* ---- ordering helpers (include/matrix_storage.h:732-823) -------------- */
/* if_sorted_coo :732 -- 1 when (row,col) is non-decreasing. */
int clspmv_if_sorted_coo(const clspmv_coo_matrix* m);
/* sort_coo :745 -- sorts by (row,col) in place; returns 1.  Entries with
 * equal (row,col) keep their input order.  Unlike the reference's
 * fixed-depth quicksort this never fails. */
int clspmv_sort_coo(clspmv_coo_matrix* m);
/* aligned_length :818 */
int clspmv_aligned_length(int length, int alignment);
/* Checks a COO matrix: indices in range, nnz >= 0.  CLSPMV_OK or error. */
int clspmv_validate_coo(const clspmv_coo_matrix* m);

/* ---- COO -> X (include/matrix_storage.h:826-2409) ---------------------- */
/* coo2csr :1269 (sorts src in place when unsorted) */
int clspmv_coo2csr(clspmv_coo_matrix* src, clspmv_csr_matrix* dst);
/* coo2ell :1341 (ellnum == 0 -> max row length) */
int clspmv_coo2ell(clspmv_coo_matrix* src, clspmv_ell_matrix* dst, int alignment, int ellnum);
/* coo2sell :1398 (src must be sorted, as in the reference) */
int clspmv_coo2sell(const clspmv_coo_matrix* src, clspmv_sell_matrix* dst, unsigned int sliceheight);
/* coo2dia :937 -- returns 1, or 0 when the padded storage would reach
 * twice nnz (the reference's refusal rule, :972). */
int clspmv_coo2dia(const clspmv_coo_matrix* src, clspmv_dia_matrix* dst, int alignment);
/* coo2diaext :1054 */
int clspmv_coo2diaext(const clspmv_coo_matrix* src, clspmv_dia_ext_matrix* dst, int alignment);
/* coo2bdia :1185 */
int clspmv_coo2bdia(const clspmv_coo_matrix* src, clspmv_bdia_matrix* dst, int alignment);
/* coo2bcsr :1595 (sorts src) */
int clspmv_coo2bcsr(clspmv_coo_matrix* src, clspmv_bcsr_matrix* dst, unsigned int bwidth, unsigned int bheight, int alignment);
/* coo2b4csr :1741 -- 0 above CLSPMV_MAX_MEM_OBJ (the reference's rule :1863) */
int clspmv_coo2b4csr(clspmv_coo_matrix* src, clspmv_b4csr_matrix* dst, unsigned int bwidth, unsigned int bheight, int alignment);
/* coo2bell :1941 */
int clspmv_coo2bell(clspmv_coo_matrix* src, clspmv_bell_matrix* dst, unsigned int bwidth, unsigned int bheight, int alignment, int ellnum);
/* coo2b4ell :2010 -- 0 above CLSPMV_MAX_MEM_OBJ (:2141) */
int clspmv_coo2b4ell(clspmv_coo_matrix* src, clspmv_b4ell_matrix* dst, unsigned int bwidth, unsigned int bheight, int alignment, int ellnum);
/* coo2sbell :2203 -- 0 above CLSPMV_MAX_MEM_OBJ (:2350) */
int clspmv_coo2sbell(clspmv_coo_matrix* src, clspmv_sbell_matrix* dst, unsigned int bwidth, unsigned int bheight, unsigned int sliceheight);

/* Variants of the three block converters without the 256 MiB refusal:
 * the device runtime has 288 GB of HBM, so the engine lifts the cap.
 * Same layouts as the capped versions. */
int clspmv_coo2b4csr_nocap(clspmv_coo_matrix* src, clspmv_b4csr_matrix* dst, unsigned int bwidth, unsigned int bheight, int alignment);
int clspmv_coo2b4ell_nocap(clspmv_coo_matrix* src, clspmv_b4ell_matrix* dst, unsigned int bwidth, unsigned int bheight, int alignment, int ellnum);
int clspmv_coo2sbell_nocap(clspmv_coo_matrix* src, clspmv_sbell_matrix* dst, unsigned int bwidth, unsigned int bheight, unsigned int sliceheight);

/* ---- X -> COO: sorted by (row, col); csr2coo copies every stored entry,
 * the others skip stored zeros (so the formats' padding goes), as the
 * reference's loops do, and set nnz to the entries kept ---------------- */
int clspmv_csr2coo(const clspmv_csr_matrix* src, clspmv_coo_matrix* dst);     /* :1240 */
int clspmv_ell2coo(const clspmv_ell_matrix* src, clspmv_coo_matrix* dst);     /* :1310 */
int clspmv_dia2coo(const clspmv_dia_matrix* src, clspmv_coo_matrix* dst);     /* :826  */
int clspmv_diaext2coo(const clspmv_dia_ext_matrix* src, clspmv_coo_matrix* dst); /* :1017 */
int clspmv_bdia2coo(const clspmv_bdia_matrix* src, clspmv_coo_matrix* dst);   /* :1137 */
int clspmv_bcsr2coo(const clspmv_bcsr_matrix* src, clspmv_coo_matrix* dst);   /* :1546 */
int clspmv_bell2coo(const clspmv_bell_matrix* src, clspmv_coo_matrix* dst);   /* :1901 */

/* ---- Matrix Market ingest (src/fileio.cpp:10) ------------------------- */
/* ReadMMF: coordinate files; 1-based -> 0-based, pattern -> 1.0, values
 * read as double and rounded to float, symmetric (and skew-symmetric, with
 * negated mirror values) expansion without doubling the diagonal, result
 * sorted by (row, col), duplicates kept.  Complex and array files are
 * refused (CLSPMV_ERR_UNSUPPORTED); indices outside the stated size give
 * CLSPMV_ERR_INVALID.  dst's arrays are malloc()ed (clspmv_free_coo_matrix). */
int clspmv_read_mmf(const char* filename, clspmv_coo_matrix* dst);

#ifdef __cplusplus
}
#endif
#endif /* CLSPMV_FORMATS_H */
