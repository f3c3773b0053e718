// ref_harness.cpp -- C entry points into the reference's own CPU code
// (include/matrix_storage.h, include/spmv_serial.h, src/analyze.cpp,
// src/util.cpp, src/fileio.cpp), compiled from the sources under /root/reference into
// oracle/_ref/libclspmv_ref.so by oracle/Makefile.  TEST INFRASTRUCTURE
// ONLY: used to pin the oracle and to write golden fixtures; the engine
// never links it.
//
// The engine's C structs are layout twins of the reference's templates; the
// static_asserts below check that member for member, so pointers are passed
// across by reinterpret_cast.

#include <fcntl.h>
#include <stddef.h>
#include <stdlib.h>
#include <unistd.h>

#include "analyze.h"          // reference
#include "fileio.h"           // reference
#include "matrix_storage.h"   // reference
#include "spmv_serial.h"      // reference
#include "util.h"             // reference

#include "clspmv_cocktail.h"  // engine C-ABI

#define SAME(A, B, f) static_assert(offsetof(A, f) == offsetof(B, f), #A "." #f)
#define SAME_SIZE(A, B) static_assert(sizeof(A) == sizeof(B), #A " size")

typedef coo_matrix<int, float> rcoo;
typedef csr_matrix<int, float> rcsr;
typedef ell_matrix<int, float> rell;
typedef sell_matrix<int, float> rsell;
typedef dia_matrix<int, int, float> rdia;
typedef dia_ext_matrix<int, int, float> rdiaext;
typedef bdia_matrix<int, int, float> rbdia;
typedef bcsr_matrix<int, float> rbcsr;
typedef bell_matrix<int, float> rbell;
typedef b4csr_matrix<int, float> rb4csr;
typedef b4ell_matrix<int, float> rb4ell;
typedef sbell_matrix<int, float> rsbell;
typedef cocktail<int, int, float> rcocktail;

SAME_SIZE(clspmv_coo_matrix, rcoo);
SAME(clspmv_coo_matrix, rcoo, coo_row_id); SAME(clspmv_coo_matrix, rcoo, coo_data);
SAME_SIZE(clspmv_csr_matrix, rcsr);
SAME(clspmv_csr_matrix, rcsr, csr_row_ptr); SAME(clspmv_csr_matrix, rcsr, csr_data);
SAME_SIZE(clspmv_ell_matrix, rell);
SAME(clspmv_ell_matrix, rell, ell_num); SAME(clspmv_ell_matrix, rell, ell_data);
SAME_SIZE(clspmv_sell_matrix, rsell);
SAME(clspmv_sell_matrix, rsell, sell_slice_height); SAME(clspmv_sell_matrix, rsell, sell_slice_ptr);
SAME(clspmv_sell_matrix, rsell, sell_data);
SAME_SIZE(clspmv_dia_matrix, rdia);
SAME(clspmv_dia_matrix, rdia, dia_length_aligned); SAME(clspmv_dia_matrix, rdia, dia_data);
SAME_SIZE(clspmv_dia_ext_matrix, rdiaext);
SAME(clspmv_dia_ext_matrix, rdiaext, dia_length_aligned); SAME(clspmv_dia_ext_matrix, rdiaext, dia_data);
SAME_SIZE(clspmv_bdia_matrix, rbdia);
SAME(clspmv_bdia_matrix, rbdia, bdia_bptr); SAME(clspmv_bdia_matrix, rbdia, bdia_data);
SAME_SIZE(clspmv_bcsr_matrix, rbcsr);
SAME(clspmv_bcsr_matrix, rbcsr, bcsr_block_num); SAME(clspmv_bcsr_matrix, rbcsr, bcsr_data);
SAME_SIZE(clspmv_bell_matrix, rbell);
SAME(clspmv_bell_matrix, rbell, bell_float4_aligned); SAME(clspmv_bell_matrix, rbell, bell_data);
SAME_SIZE(clspmv_b4csr_matrix, rb4csr);
SAME(clspmv_b4csr_matrix, rb4csr, b4csr_aligned_size); SAME(clspmv_b4csr_matrix, rb4csr, b4csr_data);
SAME_SIZE(clspmv_b4ell_matrix, rb4ell);
SAME(clspmv_b4ell_matrix, rb4ell, b4ell_float4_aligned); SAME(clspmv_b4ell_matrix, rb4ell, b4ell_data);
SAME_SIZE(clspmv_sbell_matrix, rsbell);
SAME(clspmv_sbell_matrix, rsbell, sbell_slice_num); SAME(clspmv_sbell_matrix, rsbell, sbell_data);
SAME_SIZE(clspmv_cocktail, rcocktail);
SAME(clspmv_cocktail, rcocktail, ifusecoo); SAME(clspmv_cocktail, rcocktail, bdia_meth_num);
SAME(clspmv_cocktail, rcocktail, coo_group_num); SAME(clspmv_cocktail, rcocktail, mat_nnz);
SAME(clspmv_cocktail, rcocktail, bdia); SAME(clspmv_cocktail, rcocktail, sbell);
SAME(clspmv_cocktail, rcocktail, bell); SAME(clspmv_cocktail, rcocktail, bcsr);
SAME(clspmv_cocktail, rcocktail, csr); SAME(clspmv_cocktail, rcocktail, coo);

template <class T, class U>
static T* R(U* p) { return reinterpret_cast<T*>(p); }

// Runs `f` with stdout sent to /dev/null (the reference prints its
// progress on every call).
template <class F>
static auto quiet(F&& f) -> decltype(f()) {
    fflush(stdout);
    int saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    if (devnull >= 0) { dup2(devnull, 1); close(devnull); }
    struct Restore {
        int fd;
        ~Restore() { fflush(stdout); if (fd >= 0) { dup2(fd, 1); close(fd); } }
    } restore{saved};
    return f();
}

extern "C" {

int ref_sort_coo(clspmv_coo_matrix* m) { return sort_coo<int, float>(R<rcoo>(m)) ? 1 : 0; }
int ref_if_sorted_coo(clspmv_coo_matrix* m) { return if_sorted_coo<int, float>(R<rcoo>(m)) ? 1 : 0; }

void ref_coo2csr(clspmv_coo_matrix* s, clspmv_csr_matrix* d) { coo2csr<int, float>(R<rcoo>(s), R<rcsr>(d)); }
void ref_coo2ell(clspmv_coo_matrix* s, clspmv_ell_matrix* d, int alignment, int ellnum) {
    coo2ell<int, float>(R<rcoo>(s), R<rell>(d), alignment, ellnum);
}
void ref_coo2sell(clspmv_coo_matrix* s, clspmv_sell_matrix* d, unsigned sh) {
    coo2sell<int, float>(R<rcoo>(s), R<rsell>(d), sh);
}
int ref_coo2dia(clspmv_coo_matrix* s, clspmv_dia_matrix* d, int alignment) {
    return quiet([&] { return coo2dia<int, int, float>(R<rcoo>(s), R<rdia>(d), alignment) ? 1 : 0; });
}
int ref_coo2bdia(clspmv_coo_matrix* s, clspmv_bdia_matrix* d, int alignment) {
    return quiet([&] { return coo2bdia<int, int, float>(R<rcoo>(s), R<rbdia>(d), alignment) ? 1 : 0; });
}
void ref_coo2bcsr(clspmv_coo_matrix* s, clspmv_bcsr_matrix* d, unsigned bw, unsigned bh, int alignment) {
    coo2bcsr<int, float>(R<rcoo>(s), R<rbcsr>(d), bw, bh, alignment);
}
int ref_coo2b4csr(clspmv_coo_matrix* s, clspmv_b4csr_matrix* d, unsigned bw, unsigned bh, int alignment) {
    return quiet([&] { return coo2b4csr<int, float>(R<rcoo>(s), R<rb4csr>(d), bw, bh, alignment) ? 1 : 0; });
}
void ref_coo2bell(clspmv_coo_matrix* s, clspmv_bell_matrix* d, unsigned bw, unsigned bh, int alignment, int ellnum) {
    coo2bell<int, float>(R<rcoo>(s), R<rbell>(d), bw, bh, alignment, ellnum);
}
int ref_coo2b4ell(clspmv_coo_matrix* s, clspmv_b4ell_matrix* d, unsigned bw, unsigned bh, int alignment, int ellnum) {
    return quiet([&] { return coo2b4ell<int, float>(R<rcoo>(s), R<rb4ell>(d), bw, bh, alignment, ellnum) ? 1 : 0; });
}
int ref_coo2sbell(clspmv_coo_matrix* s, clspmv_sbell_matrix* d, unsigned bw, unsigned bh, unsigned sh) {
    return quiet([&] { return coo2sbell<int, float>(R<rcoo>(s), R<rsbell>(d), bw, bh, sh) ? 1 : 0; });
}

// X -> COO (the inverses).  csr2coo is left out: its body calls
// if_sorted_coo on the CSR struct (include/matrix_storage.h:1241-1244), so
// the template does not compile when instantiated.
int ref_coo2diaext(clspmv_coo_matrix* s, clspmv_dia_ext_matrix* d, int alignment) {
    return quiet([&] { return coo2diaext<int, int, float>(R<rcoo>(s), R<rdiaext>(d), alignment) ? 1 : 0; });
}
void ref_diaext2coo(clspmv_dia_ext_matrix* s, clspmv_coo_matrix* d) { diaext2coo<int, int, float>(R<rdiaext>(s), R<rcoo>(d)); }
void ref_dia2coo(clspmv_dia_matrix* s, clspmv_coo_matrix* d) { dia2coo<int, int, float>(R<rdia>(s), R<rcoo>(d)); }
void ref_bdia2coo(clspmv_bdia_matrix* s, clspmv_coo_matrix* d) { bdia2coo<int, int, float>(R<rbdia>(s), R<rcoo>(d)); }
void ref_ell2coo(clspmv_ell_matrix* s, clspmv_coo_matrix* d) { ell2coo<int, float>(R<rell>(s), R<rcoo>(d)); }
void ref_bcsr2coo(clspmv_bcsr_matrix* s, clspmv_coo_matrix* d) { bcsr2coo<int, float>(R<rbcsr>(s), R<rcoo>(d)); }
void ref_bell2coo(clspmv_bell_matrix* s, clspmv_coo_matrix* d) { bell2coo<int, float>(R<rbell>(s), R<rcoo>(d)); }

void ref_coo_spmv(clspmv_coo_matrix* m, float* x, float* y) { coo_spmv<int, float>(R<rcoo>(m), x, y, m->matinfo.width); }
void ref_csr_spmv(clspmv_csr_matrix* m, float* x, float* y) { csr_spmv<int, float>(R<rcsr>(m), x, y, m->matinfo.width); }
void ref_ell_spmv(clspmv_ell_matrix* m, float* x, float* y) { ell_spmv<int, float>(R<rell>(m), x, y, m->matinfo.width); }
void ref_sell_spmv(clspmv_sell_matrix* m, float* x, float* y) { sell_spmv<int, float>(R<rsell>(m), x, y, m->matinfo.width); }
void ref_dia_spmv(clspmv_dia_matrix* m, float* x, float* y) { dia_spmv<int, int, float>(R<rdia>(m), x, y, m->matinfo.width); }
void ref_bdia_spmv(clspmv_bdia_matrix* m, float* x, float* y) { bdia_spmv<int, int, float>(R<rbdia>(m), x, y, m->matinfo.width); }
void ref_bcsr_spmv(clspmv_bcsr_matrix* m, float* x, float* y) { bcsr_spmv<int, float>(R<rbcsr>(m), x, y, m->matinfo.width); }
void ref_bell_spmv(clspmv_bell_matrix* m, float* x, float* y) { bell_spmv<int, float>(R<rbell>(m), x, y, m->matinfo.width); }
void ref_sbell_spmv(clspmv_sbell_matrix* m, float* x, float* y) { sbell_spmv<int, float>(R<rsbell>(m), x, y, m->matinfo.width); }
void ref_spmv_only(clspmv_coo_matrix* m, float* x, float* y) { spmv_only(R<rcoo>(m), x, y); }
// ReadMMF (src/fileio.cpp:10); arrays malloc()ed, release with free().
void ref_read_mmf(const char* path, clspmv_coo_matrix* m) { ReadMMF(const_cast<char*>(path), R<rcoo>(m)); }

// analyze_matrix with the tables under <ben_dir>/benchmark (the reference
// reads $CLSPMVPATH, src/analyze.cpp:179).
void ref_analyze_matrix(clspmv_coo_matrix* coo, clspmv_cocktail* mat, float mult,
                        int partial, int full, const char* ben_dir) {
    setenv("CLSPMVPATH", ben_dir, 1);
    init_cocktail(*R<rcocktail>(mat));
    quiet([&] {
        analyze_matrix(*R<rcoo>(coo), *R<rcocktail>(mat), mult, partial != 0, full != 0);
        return 0;
    });
}

void ref_free_cocktail(clspmv_cocktail* mat) { free_cocktail(*R<rcocktail>(mat)); }
void ref_free_coo(clspmv_coo_matrix* m) {
    free(m->coo_row_id);
    free(m->coo_col_id);
    free(m->coo_data);
    m->coo_row_id = m->coo_col_id = nullptr;
    m->coo_data = nullptr;
}

}  // extern "C"
