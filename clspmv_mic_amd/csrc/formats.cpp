// Host-side storage formats and COO<->X conversions of the clspmv engine.
//
// Each converter reproduces the layout and padding rules of the reference's
// converter of the same name in include/matrix_storage.h (line cited at each
// function), so that arrays produced here are element-for-element equal to
// the reference's on matrices without duplicate (row, col) entries.  The code
// is organised around two shared building blocks instead of per-format copies:
// `row_pointer` (CSR row starts of a row-sorted COO) and `build_block_csr`
// (the block-row grouping every block format starts from).
//
// Documented deviations (no effect on valid inputs):
//   * sort_coo never fails and keeps equal keys in input order; the
//     reference's quicksort gives up after 1000 nested partitions.
//   * Indices are validated up front and an error is returned where the
//     reference would read or write out of bounds (rows >= min(width, height)
//     in coo2dia, unsorted input to coo2sell, rows longer than ellnum).
//   * Padding that the reference leaves uninitialised (b4ell rows beyond
//     row_num) is zero here.

#include <algorithm>
#include <cassert>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdint>
#include <numeric>

#include "blocks.hpp"
#include "common.hpp"

namespace clspmv {

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

void row_pointer(const int* rows, int nnz, int height, std::vector<int>& rp) {
    rp.assign((size_t)height + 1, 0);
    for (int i = 0; i < nnz; ++i) rp[rows[i] + 1]++;
    for (int r = 0; r < height; ++r) rp[r + 1] += rp[r];
}

VecCompare two_vec_compare(const float* coovec, const float* newvec, int n) {
    VecCompare c;
    double dist = 0.0;
    for (int i = 0; i < n; ++i) {
        const double t = (double)(float)(coovec[i] - newvec[i]);
        dist += t * t;
        const float a = std::fabs(coovec[i]), b = std::fabs(newvec[i]);
        const double d = std::fabs((double)(float)(a - b));
        const float mx = a > b ? a : b;
        if (d > c.max_diff) c.max_diff = d;
        if (mx > 0 && d / mx > c.max_ratio) c.max_ratio = d / mx;
    }
    c.distance = std::sqrt(dist);
    return c;
}

bool coo_sorted(const clspmv_coo_matrix& m) {
    for (int i = 0; i + 1 < m.matinfo.nnz; ++i) {
        int r0 = m.coo_row_id[i], r1 = m.coo_row_id[i + 1];
        if (r0 > r1 || (r0 == r1 && m.coo_col_id[i] > m.coo_col_id[i + 1]))
            return false;
    }
    return true;
}

void coo_sort(clspmv_coo_matrix& m) {
    int n = m.matinfo.nnz;
    if (n <= 1) return;
    // Sort a permutation by 64-bit (row, col) keys, then gather.
    std::vector<uint64_t> key(n);
    for (int i = 0; i < n; ++i)
        key[i] = ((uint64_t)(uint32_t)m.coo_row_id[i] << 32) |
                 (uint32_t)m.coo_col_id[i];
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int a, int b) { return key[a] < key[b]; });
    std::vector<int> ri(n), ci(n);
    std::vector<float> vi(n);
    for (int i = 0; i < n; ++i) {
        ri[i] = m.coo_row_id[perm[i]];
        ci[i] = m.coo_col_id[perm[i]];
        vi[i] = m.coo_data[perm[i]];
    }
    std::copy(ri.begin(), ri.end(), m.coo_row_id);
    std::copy(ci.begin(), ci.end(), m.coo_col_id);
    std::copy(vi.begin(), vi.end(), m.coo_data);
}

void build_block_csr(const int* rows, const int* cols, const float* vals,
                     int nnz, int height, int bheight, int bwidth,
                     BlockCSR& out) {
    std::vector<int> rp;
    row_pointer(rows, nnz, height, rp);
    int nbrows = (height + bheight - 1) / bheight;
    out.bheight = bheight;
    out.bwidth = bwidth;
    out.rowptr.assign((size_t)nbrows + 1, 0);
    out.bcol.clear();
    out.tiles.clear();
    out.bcol.reserve(nnz / 2 + 1);
    out.tiles.reserve((size_t)nnz * 2 + 1);
    BlockRowScratch scratch;
    const int bs = bheight * bwidth;
    for (int br = 0; br < nbrows; ++br) {
        int count = 0;
        for_each_block(rows, cols, vals, rp, height, br, bheight, bwidth,
                       scratch, [&](int bcol, const BlockEntry* b,
                                    const BlockEntry* e) {
                           size_t off = out.tiles.size();
                           out.tiles.resize(off + bs);
                           fill_tile(b, e, bheight, bwidth, &out.tiles[off]);
                           out.bcol.push_back(bcol);
                           ++count;
                       });
        out.rowptr[br + 1] = out.rowptr[br] + count;
    }
}

}  // namespace clspmv

using namespace clspmv;

// The definitions below take C linkage from their declarations in
// clspmv_formats.h.

const char* clspmv_last_error(void) { return g_last_error.c_str(); }

// ---- init / free -----------------------------------------------------------

#define ZERO_INIT(T)                                   \
    void clspmv_init_##T(clspmv_##T* m) {              \
        if (m) std::memset(m, 0, sizeof(*m));          \
    }
ZERO_INIT(coo_matrix)
ZERO_INIT(csr_matrix)
ZERO_INIT(ell_matrix)
ZERO_INIT(sell_matrix)
ZERO_INIT(dia_matrix)
ZERO_INIT(dia_ext_matrix)
ZERO_INIT(bdia_matrix)
ZERO_INIT(bcsr_matrix)
ZERO_INIT(bell_matrix)
ZERO_INIT(b4csr_matrix)
ZERO_INIT(b4ell_matrix)
ZERO_INIT(sbell_matrix)
#undef ZERO_INIT

// Frees the arrays and clears the pointers; the header fields are kept, as in
// the reference (include/matrix_storage.h:534).
#define FREE_PTR(p) do { std::free(p); (p) = nullptr; } while (0)
void clspmv_free_coo_matrix(clspmv_coo_matrix* m) {
    if (!m) return;
    FREE_PTR(m->coo_row_id); FREE_PTR(m->coo_col_id); FREE_PTR(m->coo_data);
}
void clspmv_free_csr_matrix(clspmv_csr_matrix* m) {
    if (!m) return;
    FREE_PTR(m->csr_row_ptr); FREE_PTR(m->csr_col_id); FREE_PTR(m->csr_data);
}
void clspmv_free_ell_matrix(clspmv_ell_matrix* m) {
    if (!m) return;
    FREE_PTR(m->ell_col_id); FREE_PTR(m->ell_data);
}
void clspmv_free_sell_matrix(clspmv_sell_matrix* m) {
    if (!m) return;
    FREE_PTR(m->sell_slice_ptr); FREE_PTR(m->sell_col_id); FREE_PTR(m->sell_data);
}
void clspmv_free_dia_matrix(clspmv_dia_matrix* m) {
    if (!m) return;
    FREE_PTR(m->dia_offsets); FREE_PTR(m->dia_data);
}
void clspmv_free_dia_ext_matrix(clspmv_dia_ext_matrix* m) {
    if (!m) return;
    FREE_PTR(m->dia_offsets); FREE_PTR(m->dia_data);
}
void clspmv_free_bdia_matrix(clspmv_bdia_matrix* m) {
    if (!m) return;
    FREE_PTR(m->bdia_offsets); FREE_PTR(m->bdia_bptr); FREE_PTR(m->bdia_data);
}
void clspmv_free_bcsr_matrix(clspmv_bcsr_matrix* m) {
    if (!m) return;
    FREE_PTR(m->bcsr_row_ptr); FREE_PTR(m->bcsr_col_id); FREE_PTR(m->bcsr_data);
}
void clspmv_free_bell_matrix(clspmv_bell_matrix* m) {
    if (!m) return;
    FREE_PTR(m->bell_col_id); FREE_PTR(m->bell_data);
}
void clspmv_free_b4csr_matrix(clspmv_b4csr_matrix* m) {
    if (!m) return;
    FREE_PTR(m->b4csr_row_ptr); FREE_PTR(m->b4csr_col_id); FREE_PTR(m->b4csr_data);
}
void clspmv_free_b4ell_matrix(clspmv_b4ell_matrix* m) {
    if (!m) return;
    FREE_PTR(m->b4ell_col_id); FREE_PTR(m->b4ell_data);
}
void clspmv_free_sbell_matrix(clspmv_sbell_matrix* m) {
    if (!m) return;
    FREE_PTR(m->sbell_slice_ptr); FREE_PTR(m->sbell_col_id); FREE_PTR(m->sbell_data);
}
#undef FREE_PTR

// ---- ordering helpers --------------------------------------------------------

int clspmv_if_sorted_coo(const clspmv_coo_matrix* m) {
    return (m && coo_sorted(*m)) ? 1 : 0;
}

int clspmv_sort_coo(clspmv_coo_matrix* m) {
    if (!m) return 0;
    coo_sort(*m);
    return 1;
}

int clspmv_aligned_length(int length, int alignment) {
    return alignment > 0 ? round_up(length, alignment) : length;
}

int clspmv_validate_coo(const clspmv_coo_matrix* m) {
    if (!m) return fail(CLSPMV_ERR_INVALID, "null COO matrix");
    const clspmv_matinfo& mi = m->matinfo;
    if (mi.width < 0 || mi.height < 0 || mi.nnz < 0)
        return fail(CLSPMV_ERR_INVALID, "negative matrix dimension");
    if (mi.nnz > 0 && (!m->coo_row_id || !m->coo_col_id || !m->coo_data))
        return fail(CLSPMV_ERR_INVALID, "COO arrays missing");
    for (int i = 0; i < mi.nnz; ++i) {
        int r = m->coo_row_id[i], c = m->coo_col_id[i];
        if (r < 0 || r >= mi.height || c < 0 || c >= mi.width)
            return fail(CLSPMV_ERR_INVALID,
                        "entry %d (%d,%d) outside %dx%d matrix", i, r, c,
                        mi.height, mi.width);
    }
    return CLSPMV_OK;
}

// ---- COO -> CSR (include/matrix_storage.h:1269) ----------------------------

int clspmv_coo2csr(clspmv_coo_matrix* src, clspmv_csr_matrix* dst) {
    int st = clspmv_validate_coo(src);
    if (st) return st;
    if (!dst) return fail(CLSPMV_ERR_INVALID, "null CSR destination");
    if (!coo_sorted(*src)) coo_sort(*src);
    const int nnz = src->matinfo.nnz, h = src->matinfo.height;
    dst->matinfo = src->matinfo;
    std::vector<int> rp;
    row_pointer(src->coo_row_id, nnz, h, rp);
    dst->csr_row_ptr = host_copy(rp);
    dst->csr_col_id = host_alloc<int>(nnz);
    dst->csr_data = host_alloc<float>(nnz);
    if (!dst->csr_row_ptr || !dst->csr_col_id || !dst->csr_data)
        return fail(CLSPMV_ERR_ALLOC, "coo2csr: out of host memory");
    if (nnz) {
        std::memcpy(dst->csr_col_id, src->coo_col_id, sizeof(int) * nnz);
        std::memcpy(dst->csr_data, src->coo_data, sizeof(float) * nnz);
    }
    return CLSPMV_OK;
}

// ---- COO -> ELL (include/matrix_storage.h:1341) ----------------------------
// Column-major slots of height round_up(height, alignment).  Padding slots
// repeat the row's last real column (0 for an empty row) with value 0.

int clspmv_coo2ell(clspmv_coo_matrix* src, clspmv_ell_matrix* dst,
                   int alignment, int ellnum) {
    int st = clspmv_validate_coo(src);
    if (st) return st;
    if (!dst || alignment <= 0 || ellnum < 0)
        return fail(CLSPMV_ERR_INVALID, "coo2ell: bad arguments");
    if (!coo_sorted(*src)) coo_sort(*src);
    const int h = src->matinfo.height, nnz = src->matinfo.nnz;
    std::vector<int> rp;
    row_pointer(src->coo_row_id, nnz, h, rp);
    int maxlen = 0;
    for (int r = 0; r < h; ++r) maxlen = std::max(maxlen, rp[r + 1] - rp[r]);
    if (ellnum == 0) ellnum = maxlen;
    if (maxlen > ellnum)
        return fail(CLSPMV_ERR_INVALID, "coo2ell: row of %d entries > ellnum %d",
                    maxlen, ellnum);
    const int hal = round_up(h, alignment);
    dst->matinfo = src->matinfo;
    dst->ell_num = ellnum;
    dst->ell_height_aligned = hal;
    size_t total = (size_t)hal * ellnum;
    dst->ell_col_id = host_alloc_zero<int>(total);
    dst->ell_data = host_alloc_zero<float>(total);
    if (!dst->ell_col_id || !dst->ell_data)
        return fail(CLSPMV_ERR_ALLOC, "coo2ell: out of host memory");
    for (int r = 0; r < h; ++r) {
        int last = 0, slot = 0;
        for (int k = rp[r]; k < rp[r + 1]; ++k, ++slot) {
            dst->ell_col_id[(size_t)slot * hal + r] = src->coo_col_id[k];
            dst->ell_data[(size_t)slot * hal + r] = src->coo_data[k];
            last = src->coo_col_id[k];
        }
        for (; slot < ellnum; ++slot) dst->ell_col_id[(size_t)slot * hal + r] = last;
    }
    return CLSPMV_OK;
}

// ---- COO -> SELL (include/matrix_storage.h:1398) ---------------------------
// Slices of `sliceheight` rows, each as wide as its longest row.  A padding
// slot copies the column of the previous slot of the same row; an empty row's
// slot 0 copies slot 0 of the slice's first row.

int clspmv_coo2sell(const clspmv_coo_matrix* src, clspmv_sell_matrix* dst,
                    unsigned int sliceheight) {
    int st = clspmv_validate_coo(src);
    if (st) return st;
    if (!dst || sliceheight == 0)
        return fail(CLSPMV_ERR_INVALID, "coo2sell: bad arguments");
    if (!coo_sorted(*src))
        return fail(CLSPMV_ERR_INVALID, "coo2sell: source must be sorted");
    const int h = src->matinfo.height, nnz = src->matinfo.nnz;
    const int sh = (int)sliceheight;
    const int nslice = (h + sh - 1) / sh;
    std::vector<int> rp;
    row_pointer(src->coo_row_id, nnz, h, rp);
    dst->matinfo = src->matinfo;
    dst->sell_slice_height = sliceheight;
    dst->sell_slice_num = nslice;
    std::vector<int> ptr((size_t)nslice + 1, 0), width(nslice, 0);
    for (int s = 0; s < nslice; ++s) {
        int w = 0;
        for (int r = s * sh; r < std::min(h, (s + 1) * sh); ++r)
            w = std::max(w, rp[r + 1] - rp[r]);
        width[s] = w;
        ptr[s + 1] = ptr[s] + w * sh;
    }
    dst->sell_slice_ptr = host_copy(ptr);
    dst->sell_col_id = host_alloc_zero<int>(ptr[nslice]);
    dst->sell_data = host_alloc_zero<float>(ptr[nslice]);
    if (!dst->sell_slice_ptr || !dst->sell_col_id || !dst->sell_data)
        return fail(CLSPMV_ERR_ALLOC, "coo2sell: out of host memory");
    for (int s = 0; s < nslice; ++s) {
        int* col = dst->sell_col_id + ptr[s];
        float* val = dst->sell_data + ptr[s];
        for (int i = 0; i < sh; ++i) {       // row within slice (tail rows too)
            int r = s * sh + i;
            int len = r < h ? rp[r + 1] - rp[r] : 0;
            for (int k = 0; k < width[s]; ++k) {
                int at = k * sh + i;
                if (k < len) {
                    col[at] = src->coo_col_id[rp[r] + k];
                    val[at] = src->coo_data[rp[r] + k];
                } else {
                    col[at] = k > 0 ? col[at - sh] : col[0];
                }
            }
        }
    }
    return CLSPMV_OK;
}

// ---- COO -> DIA / DIA-ext / BDIA (include/matrix_storage.h:937,1054,1185) --

namespace {

// Collects diagonals in the order of first appearance, refusing a new one
// once the padded storage would reach twice nnz (matrix_storage.h:972).
// `wrap` folds negative offsets into [0, width) (the DIA-ext convention).
struct DiaBuild {
    std::vector<int> offsets;               // discovery order
    std::vector<std::vector<float>> data;   // per diagonal, length `len`
};

int collect_diagonals(const clspmv_coo_matrix* src, int alignment, bool wrap,
                      DiaBuild& b, int& dia_length, int& len, bool& refused) {
    const clspmv_matinfo& mi = src->matinfo;
    dia_length = std::min(mi.width, mi.height);
    len = round_up(dia_length, alignment);
    refused = false;
    std::vector<int> slot_of;  // offset -> index (dense table over offsets)
    const int span = mi.width + mi.height + 1;
    slot_of.assign(span, -1);
    for (int i = 0; i < mi.nnz; ++i) {
        int r = src->coo_row_id[i], c = src->coo_col_id[i];
        if (r >= len)
            return fail(CLSPMV_ERR_INVALID,
                        "coo2dia: row %d beyond diagonal length %d", r, len);
        int off = c - r;
        if (wrap && off < 0) off += mi.width;
        int key = off + mi.height;
        int j = slot_of[key];
        if (j < 0) {
            if ((long long)b.offsets.size() * len >= 2LL * mi.nnz) {
                refused = true;
                return CLSPMV_OK;
            }
            j = (int)b.offsets.size();
            slot_of[key] = j;
            b.offsets.push_back(off);
            b.data.emplace_back((size_t)len, 0.0f);
        }
        b.data[j][r] = src->coo_data[i];
    }
    return CLSPMV_OK;
}

// Offsets ascending; returns the permutation into discovery order.
std::vector<int> order_by_offset(const std::vector<int>& offsets) {
    std::vector<int> order(offsets.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return offsets[a] < offsets[b]; });
    return order;
}

template <class DiaT>
int dia_like(const clspmv_coo_matrix* src, DiaT* dst, int alignment, bool wrap) {
    int st = clspmv_validate_coo(src);
    if (st) return st;
    if (!dst || alignment <= 0) return fail(CLSPMV_ERR_INVALID, "coo2dia: bad arguments");
    DiaBuild b;
    int dia_length = 0, len = 0;
    bool refused = false;
    st = collect_diagonals(src, alignment, wrap, b, dia_length, len, refused);
    if (st) return st;  // error text already recorded
    dst->matinfo = src->matinfo;
    dst->dia_length = dia_length;
    dst->dia_length_aligned = len;
    dst->dia_num = (int)b.offsets.size();
    if (refused) {
        dst->dia_offsets = nullptr;
        dst->dia_data = nullptr;
        return 0;
    }
    std::vector<int> order = order_by_offset(b.offsets);
    dst->dia_offsets = host_alloc<int>(order.size());
    dst->dia_data = host_alloc<float>(order.size() * (size_t)len);
    if (!dst->dia_offsets || !dst->dia_data)
        return fail(CLSPMV_ERR_ALLOC, "coo2dia: out of host memory");
    for (size_t j = 0; j < order.size(); ++j) {
        dst->dia_offsets[j] = b.offsets[order[j]];
        std::copy(b.data[order[j]].begin(), b.data[order[j]].end(),
                  dst->dia_data + j * (size_t)len);
    }
    return 1;
}

}  // namespace

int clspmv_coo2dia(const clspmv_coo_matrix* src, clspmv_dia_matrix* dst, int alignment) {
    return dia_like(src, dst, alignment, false);
}

int clspmv_coo2diaext(const clspmv_coo_matrix* src, clspmv_dia_ext_matrix* dst, int alignment) {
    return dia_like(src, dst, alignment, true);
}

int clspmv_coo2bdia(const clspmv_coo_matrix* src, clspmv_bdia_matrix* dst, int alignment) {
    clspmv_dia_matrix d;
    clspmv_init_dia_matrix(&d);
    int ok = clspmv_coo2dia(src, &d, alignment);
    if (ok != 1) return ok;
    dst->matinfo = src->matinfo;
    // Consecutive offsets (difference 1) form one band (matrix_storage.h:1196).
    std::vector<int> first, count;
    for (int j = 0; j < d.dia_num; ++j) {
        if (j > 0 && d.dia_offsets[j] - d.dia_offsets[j - 1] == 1) {
            count.back()++;
        } else {
            first.push_back(d.dia_offsets[j]);
            count.push_back(1);
        }
    }
    dst->bdia_band_num = (int)first.size();
    dst->bdia_length = d.dia_length;
    dst->bdia_length_aligned = d.dia_length_aligned;
    dst->bdia_offsets = host_copy(first);
    std::vector<int> bptr(first.size() + 1, 0);
    for (size_t b = 0; b < count.size(); ++b) bptr[b + 1] = bptr[b] + count[b];
    dst->bdia_bptr = host_copy(bptr);
    dst->bdia_data = d.dia_data;  // same diagonal-major storage
    d.dia_data = nullptr;
    clspmv_free_dia_matrix(&d);
    return 1;
}

// ---- block formats (include/matrix_storage.h:1595,1741,1941,2010,2203) -----

namespace {

int block_prologue(clspmv_coo_matrix* src, unsigned bwidth, unsigned bheight,
                   bool need_w4, BlockCSR& bc) {
    int st = clspmv_validate_coo(src);
    if (st) return st;
    if (bwidth == 0 || bheight == 0 || (need_w4 && bwidth % 4 != 0))
        return fail(CLSPMV_ERR_INVALID, "block shape %ux%u not supported",
                    bheight, bwidth);
    if (!coo_sorted(*src)) coo_sort(*src);
    build_block_csr(src->coo_row_id, src->coo_col_id, src->coo_data,
                    src->matinfo.nnz, src->matinfo.height, (int)bheight,
                    (int)bwidth, bc);
    return CLSPMV_OK;
}

int max_blocks_per_row(const BlockCSR& bc) {
    int m = 0;
    for (int r = 0; r < bc.nbrows(); ++r) m = std::max(m, bc.rowptr[r + 1] - bc.rowptr[r]);
    return m;
}

bool over_cap(long long blocks, unsigned bw, unsigned bh) {
    return blocks > (long long)(CLSPMV_MAX_MEM_OBJ / (bw * bh * sizeof(float)));
}

}  // namespace

int clspmv_coo2bcsr(clspmv_coo_matrix* src, clspmv_bcsr_matrix* dst,
                    unsigned int bwidth, unsigned int bheight, int alignment) {
    BlockCSR bc;
    int st = block_prologue(src, bwidth, bheight, false, bc);
    if (st) return st;
    if (!dst || alignment <= 0) return fail(CLSPMV_ERR_INVALID, "coo2bcsr: bad arguments");
    const int nb = bc.nblocks(), bs = (int)(bwidth * bheight);
    const int aligned = round_up(nb, alignment);
    dst->matinfo = src->matinfo;
    dst->bcsr_bwidth = bwidth;
    dst->bcsr_bheight = bheight;
    dst->bcsr_row_num = bc.nbrows();
    dst->bcsr_aligned_size = aligned;
    dst->bcsr_block_num = nb;
    dst->bcsr_row_ptr = host_copy(bc.rowptr);
    dst->bcsr_col_id = host_copy(bc.bcol);
    dst->bcsr_data = host_alloc_zero<float>((size_t)aligned * bs);
    if (!dst->bcsr_row_ptr || !dst->bcsr_col_id || !dst->bcsr_data)
        return fail(CLSPMV_ERR_ALLOC, "coo2bcsr: out of host memory");
    for (int i = 0; i < nb; ++i)
        for (int k = 0; k < bs; ++k)
            dst->bcsr_data[(size_t)k * aligned + i] = bc.tiles[(size_t)i * bs + k];
    return CLSPMV_OK;
}

static int coo2b4csr_impl(clspmv_coo_matrix* src, clspmv_b4csr_matrix* dst,
                          unsigned bwidth, unsigned bheight, int alignment,
                          bool cap) {
    BlockCSR bc;
    int st = block_prologue(src, bwidth, bheight, true, bc);
    if (st) return st;
    if (!dst || alignment <= 0) return fail(CLSPMV_ERR_INVALID, "coo2b4csr: bad arguments");
    dst->matinfo = src->matinfo;
    dst->b4csr_bwidth = bwidth;
    dst->b4csr_bheight = bheight;
    const int nb = bc.nblocks(), bs = (int)(bwidth * bheight);
    if (cap && over_cap(nb, bwidth, bheight)) {
        fail(CLSPMV_ERR_TOO_LARGE, "b4csr: %d blocks exceed 256 MiB", nb);
        return 0;
    }
    const long long aligned = round_up_ll(4LL * nb, alignment);
    const int bw4 = (int)bwidth / 4, bh = (int)bheight;
    if (aligned > INT_MAX) return fail(CLSPMV_ERR_TOO_LARGE, "b4csr: too many blocks");
    dst->b4csr_row_num = bc.nbrows();
    dst->b4csr_aligned_size = (int)aligned;
    dst->b4csr_block_num = nb;
    dst->b4csr_row_ptr = host_copy(bc.rowptr);
    dst->b4csr_col_id = host_alloc<int>(nb);
    dst->b4csr_data = host_alloc_zero<float>((size_t)aligned * bh * bw4);
    if (!dst->b4csr_row_ptr || !dst->b4csr_col_id || !dst->b4csr_data)
        return fail(CLSPMV_ERR_ALLOC, "coo2b4csr: out of host memory");
    const size_t plane = (size_t)aligned * bh;  // one float4 column of blocks
    for (int i = 0; i < nb; ++i) {
        dst->b4csr_col_id[i] = bc.bcol[i] / 4;
        const float* t = &bc.tiles[(size_t)i * bs];
        for (int hh = 0; hh < bh; ++hh)
            for (int w4 = 0; w4 < bw4; ++w4)
                for (int w = 0; w < 4; ++w)
                    dst->b4csr_data[w4 * plane + (size_t)hh * aligned + (size_t)i * 4 + w] =
                        t[hh * bwidth + w4 * 4 + w];
    }
    return 1;
}

int clspmv_coo2b4csr(clspmv_coo_matrix* src, clspmv_b4csr_matrix* dst,
                     unsigned int bwidth, unsigned int bheight, int alignment) {
    return coo2b4csr_impl(src, dst, bwidth, bheight, alignment, true);
}
int clspmv_coo2b4csr_nocap(clspmv_coo_matrix* src, clspmv_b4csr_matrix* dst,
                           unsigned int bwidth, unsigned int bheight, int alignment) {
    return coo2b4csr_impl(src, dst, bwidth, bheight, alignment, false);
}

int clspmv_coo2bell(clspmv_coo_matrix* src, clspmv_bell_matrix* dst,
                    unsigned int bwidth, unsigned int bheight, int alignment,
                    int ellnum) {
    BlockCSR bc;
    int st = block_prologue(src, bwidth, bheight, false, bc);
    if (st) return st;
    if (!dst || alignment <= 0 || ellnum < 0)
        return fail(CLSPMV_ERR_INVALID, "coo2bell: bad arguments");
    int maxb = max_blocks_per_row(bc);
    if (ellnum == 0) ellnum = maxb;
    if (maxb > ellnum)
        return fail(CLSPMV_ERR_INVALID, "coo2bell: %d blocks in a row > ellnum %d", maxb, ellnum);
    const int nbr = bc.nbrows(), bs = (int)(bwidth * bheight);
    const int hal = round_up(nbr, alignment);
    dst->matinfo = src->matinfo;
    dst->bell_bwidth = bwidth;
    dst->bell_bheight = bheight;
    dst->bell_block_num = ellnum;
    dst->bell_row_num = nbr;
    dst->bell_height_aligned = hal;
    dst->bell_float4_aligned = 0;  // not set by the reference's coo2bell
    dst->bell_col_id = host_alloc_zero<int>((size_t)hal * ellnum);
    dst->bell_data = host_alloc_zero<float>((size_t)hal * ellnum * bs);
    if (!dst->bell_col_id || !dst->bell_data)
        return fail(CLSPMV_ERR_ALLOC, "coo2bell: out of host memory");
    const size_t slot_stride = (size_t)hal * bs;
    for (int r = 0; r < nbr; ++r) {
        int last = 0, slot = 0;
        for (int j = bc.rowptr[r]; j < bc.rowptr[r + 1]; ++j, ++slot) {
            dst->bell_col_id[r + (size_t)slot * hal] = bc.bcol[j];
            for (int k = 0; k < bs; ++k)
                dst->bell_data[r + (size_t)k * hal + slot * slot_stride] =
                    bc.tiles[(size_t)j * bs + k];
            last = bc.bcol[j];
        }
        for (; slot < ellnum; ++slot) dst->bell_col_id[r + (size_t)slot * hal] = last;
    }
    return CLSPMV_OK;
}

static int coo2b4ell_impl(clspmv_coo_matrix* src, clspmv_b4ell_matrix* dst,
                          unsigned bwidth, unsigned bheight, int alignment,
                          int ellnum, bool cap) {
    BlockCSR bc;
    int st = block_prologue(src, bwidth, bheight, true, bc);
    if (st) return st;
    if (!dst || alignment <= 0 || ellnum < 0)
        return fail(CLSPMV_ERR_INVALID, "coo2b4ell: bad arguments");
    dst->matinfo = src->matinfo;
    dst->b4ell_bwidth = bwidth;
    dst->b4ell_bheight = bheight;
    int maxb = max_blocks_per_row(bc);
    if (ellnum == 0) ellnum = maxb;
    if (maxb > ellnum)
        return fail(CLSPMV_ERR_INVALID, "coo2b4ell: %d blocks in a row > ellnum %d", maxb, ellnum);
    const int nbr = bc.nbrows();
    if (cap && over_cap((long long)nbr * ellnum, bwidth, bheight)) {
        fail(CLSPMV_ERR_TOO_LARGE, "b4ell: %d x %d blocks exceed 256 MiB", nbr, ellnum);
        return 0;
    }
    const int bs = (int)(bwidth * bheight), bw4 = (int)bwidth / 4, bh = (int)bheight;
    const int hal = round_up(nbr, alignment);
    const int f4 = round_up(4 * nbr, alignment);
    dst->b4ell_row_num = nbr;
    dst->b4ell_block_num = ellnum;
    dst->b4ell_height_aligned = hal;
    dst->b4ell_float4_aligned = f4;
    dst->b4ell_col_id = host_alloc_zero<int>((size_t)hal * ellnum);
    dst->b4ell_data = host_alloc_zero<float>((size_t)f4 * bh * bw4 * ellnum);
    if (!dst->b4ell_col_id || !dst->b4ell_data)
        return fail(CLSPMV_ERR_ALLOC, "coo2b4ell: out of host memory");
    const size_t slot_stride = (size_t)f4 * bw4 * bh;
    const size_t w4_stride = (size_t)f4 * bh;
    for (int r = 0; r < nbr; ++r) {
        int last = 0, slot = 0;
        for (int j = bc.rowptr[r]; j < bc.rowptr[r + 1]; ++j, ++slot) {
            last = bc.bcol[j] / 4;
            dst->b4ell_col_id[r + (size_t)slot * hal] = last;
            const float* t = &bc.tiles[(size_t)j * bs];
            for (int hh = 0; hh < bh; ++hh)
                for (int w4 = 0; w4 < bw4; ++w4)
                    for (int w = 0; w < 4; ++w)
                        dst->b4ell_data[slot * slot_stride + (size_t)hh * f4 +
                                        w4 * w4_stride + (size_t)r * 4 + w] =
                            t[hh * bwidth + w4 * 4 + w];
        }
        for (; slot < ellnum; ++slot) dst->b4ell_col_id[r + (size_t)slot * hal] = last;
    }
    return 1;
}

int clspmv_coo2b4ell(clspmv_coo_matrix* src, clspmv_b4ell_matrix* dst,
                     unsigned int bwidth, unsigned int bheight, int alignment, int ellnum) {
    return coo2b4ell_impl(src, dst, bwidth, bheight, alignment, ellnum, true);
}
int clspmv_coo2b4ell_nocap(clspmv_coo_matrix* src, clspmv_b4ell_matrix* dst,
                           unsigned int bwidth, unsigned int bheight, int alignment, int ellnum) {
    return coo2b4ell_impl(src, dst, bwidth, bheight, alignment, ellnum, false);
}

static int coo2sbell_impl(clspmv_coo_matrix* src, clspmv_sbell_matrix* dst,
                          unsigned bwidth, unsigned bheight, unsigned sliceheight,
                          bool cap) {
    BlockCSR bc;
    int st = block_prologue(src, bwidth, bheight, true, bc);
    if (st) return st;
    if (!dst || sliceheight == 0) return fail(CLSPMV_ERR_INVALID, "coo2sbell: bad arguments");
    const int nbr = bc.nbrows(), sh = (int)sliceheight;
    const int nslice = (nbr + sh - 1) / sh;
    dst->matinfo = src->matinfo;
    dst->sbell_bwidth = bwidth;
    dst->sbell_bheight = bheight;
    dst->sbell_slice_num = nslice;
    dst->sbell_row_num = nbr;
    dst->sbell_slice_height = sliceheight;
    std::vector<int> ptr((size_t)nslice + 1, 0), width(nslice, 0);
    for (int s = 0; s < nslice; ++s) {
        int w = 0;
        for (int r = s * sh; r < std::min(nbr, (s + 1) * sh); ++r)
            w = std::max(w, bc.rowptr[r + 1] - bc.rowptr[r]);
        width[s] = w;
        ptr[s + 1] = ptr[s] + w * sh;
    }
    const int total = ptr[nslice];
    if (cap && over_cap(total, bwidth, bheight)) {
        dst->sbell_slice_ptr = nullptr;
        fail(CLSPMV_ERR_TOO_LARGE, "sbell: %d blocks exceed 256 MiB", total);
        return 0;
    }
    const int bs = (int)(bwidth * bheight), bw4 = (int)bwidth / 4, bh = (int)bheight;
    dst->sbell_slice_ptr = host_copy(ptr);
    dst->sbell_col_id = host_alloc_zero<int>(total);
    dst->sbell_data = host_alloc_zero<float>((size_t)total * bs);
    if (!dst->sbell_slice_ptr || !dst->sbell_col_id || !dst->sbell_data)
        return fail(CLSPMV_ERR_ALLOC, "coo2sbell: out of host memory");
    const size_t slot_stride = (size_t)sh * bw4 * bh * 4;
    const size_t h_stride = (size_t)sh * 4;
    const size_t w4_stride = (size_t)sh * bh * 4;
    for (int r = 0; r < nbr; ++r) {
        const int s = r / sh, i = r % sh;
        const size_t colstart = ptr[s];
        float* data = dst->sbell_data + colstart * bs;
        int last = 0, slot = 0;
        for (int j = bc.rowptr[r]; j < bc.rowptr[r + 1]; ++j, ++slot) {
            last = bc.bcol[j] / 4;
            dst->sbell_col_id[colstart + i + (size_t)slot * sh] = last;
            const float* t = &bc.tiles[(size_t)j * bs];
            for (int hh = 0; hh < bh; ++hh)
                for (int w4 = 0; w4 < bw4; ++w4)
                    for (int w = 0; w < 4; ++w)
                        data[slot * slot_stride + hh * h_stride + w4 * w4_stride +
                             (size_t)i * 4 + w] = t[hh * bwidth + w4 * 4 + w];
        }
        for (; slot < width[s]; ++slot) dst->sbell_col_id[colstart + i + (size_t)slot * sh] = last;
    }
    return 1;
}

int clspmv_coo2sbell(clspmv_coo_matrix* src, clspmv_sbell_matrix* dst,
                     unsigned int bwidth, unsigned int bheight, unsigned int sliceheight) {
    return coo2sbell_impl(src, dst, bwidth, bheight, sliceheight, true);
}
int clspmv_coo2sbell_nocap(clspmv_coo_matrix* src, clspmv_sbell_matrix* dst,
                           unsigned int bwidth, unsigned int bheight, unsigned int sliceheight) {
    return coo2sbell_impl(src, dst, bwidth, bheight, sliceheight, false);
}

// ---- X -> COO ----------------------------------------------------------------
// Each inverse gathers the stored nonzeros (explicit zeros dropped, as in the
// reference) and returns them sorted by (row, col).

namespace {
struct Triples {
    std::vector<int> r, c;
    std::vector<float> v;
    void add(int row, int col, float val) {
        if (val != 0.0f) { r.push_back(row); c.push_back(col); v.push_back(val); }
    }
};

int emit_coo(Triples& t, const clspmv_matinfo& mi, clspmv_coo_matrix* dst) {
    if (!dst) return fail(CLSPMV_ERR_INVALID, "null COO destination");
    dst->matinfo = mi;
    dst->matinfo.nnz = (int)t.r.size();
    dst->coo_row_id = host_copy(t.r);
    dst->coo_col_id = host_copy(t.c);
    dst->coo_data = host_copy(t.v);
    if (!dst->coo_row_id || !dst->coo_col_id || !dst->coo_data)
        return fail(CLSPMV_ERR_ALLOC, "out of host memory");
    if (!coo_sorted(*dst)) coo_sort(*dst);
    return CLSPMV_OK;
}
}  // namespace

int clspmv_csr2coo(const clspmv_csr_matrix* src, clspmv_coo_matrix* dst) {
    if (!src || !dst) return fail(CLSPMV_ERR_INVALID, "csr2coo: null argument");
    // csr2coo keeps every stored entry, zeros included (matrix_storage.h:1240).
    const int nnz = src->matinfo.nnz, h = src->matinfo.height;
    dst->matinfo = src->matinfo;
    dst->coo_row_id = host_alloc<int>(nnz);
    dst->coo_col_id = host_alloc<int>(nnz);
    dst->coo_data = host_alloc<float>(nnz);
    if (!dst->coo_row_id || !dst->coo_col_id || !dst->coo_data)
        return fail(CLSPMV_ERR_ALLOC, "csr2coo: out of host memory");
    for (int r = 0; r < h; ++r)
        for (int k = src->csr_row_ptr[r]; k < src->csr_row_ptr[r + 1]; ++k)
            dst->coo_row_id[k] = r;
    if (nnz) {
        std::memcpy(dst->coo_col_id, src->csr_col_id, sizeof(int) * nnz);
        std::memcpy(dst->coo_data, src->csr_data, sizeof(float) * nnz);
    }
    return CLSPMV_OK;
}

int clspmv_ell2coo(const clspmv_ell_matrix* src, clspmv_coo_matrix* dst) {
    if (!src) return fail(CLSPMV_ERR_INVALID, "ell2coo: null argument");
    Triples t;
    const int hal = src->ell_height_aligned;
    for (int r = 0; r < src->matinfo.height; ++r)
        for (int s = 0; s < src->ell_num; ++s)
            t.add(r, src->ell_col_id[(size_t)s * hal + r], src->ell_data[(size_t)s * hal + r]);
    return emit_coo(t, src->matinfo, dst);
}

int clspmv_dia2coo(const clspmv_dia_matrix* src, clspmv_coo_matrix* dst) {
    if (!src) return fail(CLSPMV_ERR_INVALID, "dia2coo: null argument");
    Triples t;
    const int len = src->dia_length_aligned;
    for (int r = 0; r < len; ++r)
        for (int j = 0; j < src->dia_num; ++j) {
            long long c = (long long)r + src->dia_offsets[j];
            if (c >= 0 && c < src->dia_length)
                t.add(r, (int)c, src->dia_data[(size_t)j * len + r]);
        }
    return emit_coo(t, src->matinfo, dst);
}

int clspmv_diaext2coo(const clspmv_dia_ext_matrix* src, clspmv_coo_matrix* dst) {
    if (!src) return fail(CLSPMV_ERR_INVALID, "diaext2coo: null argument");
    Triples t;
    const int len = src->dia_length_aligned, w = src->matinfo.width;
    for (int r = 0; r < len; ++r)
        for (int j = 0; j < src->dia_num; ++j)
            t.add(r, (int)(((long long)r + src->dia_offsets[j]) % w),
                  src->dia_data[(size_t)j * len + r]);
    return emit_coo(t, src->matinfo, dst);
}

int clspmv_bdia2coo(const clspmv_bdia_matrix* src, clspmv_coo_matrix* dst) {
    if (!src) return fail(CLSPMV_ERR_INVALID, "bdia2coo: null argument");
    Triples t;
    const int len = src->bdia_length_aligned;
    for (int b = 0; b < src->bdia_band_num; ++b)
        for (int j = src->bdia_bptr[b]; j < src->bdia_bptr[b + 1]; ++j) {
            int off = src->bdia_offsets[b] + (j - src->bdia_bptr[b]);
            for (int r = 0; r < len; ++r) {
                long long c = (long long)r + off;
                if (c >= 0 && c < src->bdia_length)
                    t.add(r, (int)c, src->bdia_data[(size_t)j * len + r]);
            }
        }
    return emit_coo(t, src->matinfo, dst);
}

int clspmv_bcsr2coo(const clspmv_bcsr_matrix* src, clspmv_coo_matrix* dst) {
    if (!src) return fail(CLSPMV_ERR_INVALID, "bcsr2coo: null argument");
    Triples t;
    const int bw = (int)src->bcsr_bwidth, bh = (int)src->bcsr_bheight;
    for (int br = 0; br < src->bcsr_row_num; ++br)
        for (int k = 0; k < bw * bh; ++k)
            for (int j = src->bcsr_row_ptr[br]; j < src->bcsr_row_ptr[br + 1]; ++j)
                t.add(br * bh + k / bw, src->bcsr_col_id[j] + k % bw,
                      src->bcsr_data[j + (size_t)k * src->bcsr_aligned_size]);
    return emit_coo(t, src->matinfo, dst);
}

int clspmv_bell2coo(const clspmv_bell_matrix* src, clspmv_coo_matrix* dst) {
    if (!src) return fail(CLSPMV_ERR_INVALID, "bell2coo: null argument");
    Triples t;
    const int bw = (int)src->bell_bwidth, bh = (int)src->bell_bheight;
    const int hal = src->bell_height_aligned;
    const size_t slot_stride = (size_t)hal * bw * bh;
    for (int s = 0; s < src->bell_block_num; ++s)
        for (int r = 0; r < src->bell_row_num; ++r)
            for (int k = 0; k < bw * bh; ++k)
                t.add(r * bh + k / bw, src->bell_col_id[(size_t)s * hal + r] + k % bw,
                      src->bell_data[r + (size_t)k * hal + s * slot_stride]);
    return emit_coo(t, src->matinfo, dst);
}

// ---- Matrix Market (src/fileio.cpp:10) --------------------------------------

int clspmv_read_mmf(const char* filename, clspmv_coo_matrix* dst) {
    if (!filename || !dst) return fail(CLSPMV_ERR_INVALID, "read_mmf: null argument");
    FILE* f = std::fopen(filename, "r");
    if (!f) return fail(CLSPMV_ERR_IO, "read_mmf: cannot open %s", filename);
    char banner[5][128] = {{0}};
    for (int i = 0; i < 5; ++i)
        if (std::fscanf(f, "%127s", banner[i]) != 1) {
            std::fclose(f);
            return fail(CLSPMV_ERR_IO, "read_mmf: truncated banner");
        }
    // Banner: %%MatrixMarket matrix coordinate <field> <symmetry>; field 4
    // is the value type, field 5 the symmetry (src/fileio.cpp:15-25).  The
    // reference parses values only for "real" and mirrors only "symmetric";
    // here integer values are parsed too, skew-symmetric matrices are
    // mirrored with negated values, and complex or dense (array) files are
    // refused instead of misread.
    if (std::strcmp(banner[2], "coordinate") != 0) {
        std::fclose(f);
        return fail(CLSPMV_ERR_UNSUPPORTED, "read_mmf: only coordinate files are supported (got %s)", banner[2]);
    }
    if (std::strcmp(banner[3], "complex") == 0) {
        std::fclose(f);
        return fail(CLSPMV_ERR_UNSUPPORTED, "read_mmf: complex values are not supported");
    }
    const bool has_values = std::strcmp(banner[3], "pattern") != 0;
    const bool skew = std::strcmp(banner[4], "skew-symmetric") == 0;
    const bool symmetric = skew || std::strcmp(banner[4], "symmetric") == 0;
    char line[4096];
    int height = 0, width = 0, nnz = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '%' || line[0] == '\n') continue;
        if (std::sscanf(line, "%d %d %d", &height, &width, &nnz) == 3) break;
    }
    if (height <= 0 || width <= 0 || nnz < 0) {
        std::fclose(f);
        return fail(CLSPMV_ERR_IO, "read_mmf: bad size line");
    }
    std::vector<int> r, c;
    std::vector<float> v;
    r.reserve(symmetric ? 2 * (size_t)nnz : nnz);
    c.reserve(r.capacity());
    v.reserve(r.capacity());
    for (int i = 0; i < nnz; ++i) {
        int ri = 0, ci = 0;
        double val = 1.0;
        if (std::fscanf(f, "%d %d", &ri, &ci) != 2 ||
            (has_values && std::fscanf(f, "%lf", &val) != 1)) {
            std::fclose(f);
            return fail(CLSPMV_ERR_IO, "read_mmf: truncated entry %d", i);
        }
        r.push_back(ri - 1); c.push_back(ci - 1); v.push_back((float)val);
        if (symmetric && ri != ci) {
            r.push_back(ci - 1); c.push_back(ri - 1); v.push_back(skew ? -(float)val : (float)val);
        }
    }
    std::fclose(f);
    dst->matinfo.width = width;
    dst->matinfo.height = height;
    dst->matinfo.nnz = (int)r.size();
    dst->coo_row_id = host_copy(r);
    dst->coo_col_id = host_copy(c);
    dst->coo_data = host_copy(v);
    int st = clspmv_validate_coo(dst);
    if (st) {
        clspmv_free_coo_matrix(dst);
        return st;
    }
    if (!coo_sorted(*dst)) coo_sort(*dst);
    return CLSPMV_OK;
}
