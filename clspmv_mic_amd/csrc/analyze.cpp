// Cocktail partitioner: analyze_matrix (src/analyze.cpp:3714-3900).
//
// Splits a sorted COO matrix into at most nine parts -- BDIA, DIA (diagonal
// family), SBELL, BELL, BCSR (block family), SELL, ELL, CSR, COO (flat
// family) -- by a cost model over measured per-format GFLOP/s tables.  The
// decisions (which diagonals, which block shape, which widths, which format
// takes the remainder) are the reference's, restated here so that the same
// tables and the same matrix give the same partition, index for index.  Every
// floating-point expression of the model keeps the reference's types and
// evaluation order (float/double mix, truncating casts, the ms-vs-us time
// units of src/analyze.cpp:1401-1420), because a changed rounding can flip a
// decision.
//
// The implementation differs in structure, not in outcome:
//   * block counting sorts packed (block row, block column) keys per 8-row
//     strip instead of filling eight std::maps per strip (count_blocks,
//     src/analyze.cpp:1582);
//   * the extract_*_part routines share one block-row walker;
//   * tables are validated when read; conditions under which the reference
//     reads out of bounds or aborts on an assert are reported as errors or
//     noted below.
// Known divergences, all outside well-formed inputs: the reference reads
// bw[8] past the end of its shape table while locating the last block shape
// (src/analyze.cpp:523); the walk here stops at the eighth shape.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <utility>

#include "blocks.hpp"
#include "cocktail.hpp"
#include "common.hpp"

namespace clspmv {
namespace ckt {
namespace {

const int kBH[8] = {1, 2, 4, 8, 1, 2, 4, 8};
const int kBW[8] = {4, 4, 4, 4, 8, 8, 8, 8};

void (*g_log)(const char*) = nullptr;

void logf(const char* fmt, ...) {
    if (!g_log) return;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_log(buf);
}

// maxflop_info (include/analyze.h:208-245).
struct FlopInfo {
    int bdia_meth_num = 0; double bdia_max_flop = 0;
    int dia_meth_num = 0; double dia_max_flop = 0;
    int sbell_meth_num[8] = {0}; double sbell_max_flop[8] = {0};
    int bell_meth_num[8] = {0}; double bell_max_flop[8] = {0};
    int bcsr_meth_num[8] = {0}; double bcsr_max_flop[8] = {0};
    int sell_meth_num = 0; double sell_max_flop = 0;
    int ell_meth_num = 0; double ell_max_flop = 0;
    int csr_meth_num = 0; int csr_groupnum = 0; double csr_max_flop = 0;
    int coo_meth_num = 0; int coo_groupnum = 0; double coo_max_flop = 0;
    double max_dia_base_flop = 0, max_block_base_flop = 0, max_flat_base_flop = 0;
};

// ---- table model (src/analyze.cpp:294-430) ---------------------------------

// Rows (ss, sl, ls, ll) bracketing (dim, width): smaller/larger dimension x
// smaller/larger width, clamped to the table's edges (find_fourid :294).
struct Bracket { int ss, sl, ls, ll; };

struct Key { unsigned dim, width; };

void width_bracket(const std::vector<Key>& k, int lo, int hi, int width, int& s, int& l) {
    // Widths compare as unsigned, like the reference's int-vs-unsigned tests.
    const unsigned w = (unsigned)width;
    if (w < k[lo].width) {
        s = l = lo;
    } else if (w > k[hi - 1].width) {
        s = l = hi - 1;
    } else {
        for (int i = lo; i < hi; ++i) {
            if (w == k[i].width) {
                s = l = i;
                break;
            }
            if (i < hi - 1 && k[i].width < w && k[i + 1].width > w) {
                s = i;
                l = i + 1;
            }
        }
    }
}

// Start offsets of the runs of equal dimension (plus the end).
void dim_runs(const std::vector<Key>& k, std::vector<int>& dims, std::vector<int>& offs) {
    dims.assign(1, (int)k[0].dim);
    offs.assign(1, 0);
    for (size_t i = 0; i < k.size(); ++i)
        if ((int)k[i].dim != dims.back()) {
            dims.push_back((int)k[i].dim);
            offs.push_back((int)i);
        }
    offs.push_back((int)k.size());
}

void dim_bracket(const std::vector<int>& dims, int dim, int& s, int& l) {
    s = l = 0;
    if (dim < dims[0]) {
        s = l = 0;
    } else if (dim > dims.back()) {
        s = l = (int)dims.size() - 1;
    } else {
        for (size_t i = 0; i < dims.size(); ++i) {
            if (dim == dims[i]) {
                s = l = (int)i;
                break;
            }
            if (i + 1 < dims.size() && dims[i] < dim && dims[i + 1] > dim) {
                s = (int)i;
                l = (int)i + 1;
                break;
            }
        }
    }
}

Bracket bracket(const std::vector<Key>& k, int dim, int width) {
    std::vector<int> dims, offs;
    dim_runs(k, dims, offs);
    int ds, dl;
    dim_bracket(dims, dim, ds, dl);
    Bracket b{0, 0, 0, 0};
    width_bracket(k, offs[ds], offs[ds + 1], width, b.ss, b.sl);
    width_bracket(k, offs[dl], offs[dl + 1], width, b.ls, b.ll);
    return b;
}

// linear_inter (:401)
double lerp(int small, double sv, int large, double lv, int at) {
    if (small == large) return lv;
    return sv + ((double)(at - small)) * (lv - sv) / ((double)(large - small));
}

// Bilinear value of column `m` at (dim, width); gf(i) returns row i's value.
double bilinear(const std::vector<Key>& k, const Bracket& b, int dim, int width,
                const std::function<double(int)>& gf) {
    double sv = lerp((int)k[b.ss].width, gf(b.ss), (int)k[b.sl].width, gf(b.sl), width);
    double lv = lerp((int)k[b.ls].width, gf(b.ls), (int)k[b.ll].width, gf(b.ll), width);
    return lerp((int)k[b.ss].dim, sv, (int)k[b.ll].dim, lv, dim);
}

class Model {
  public:
    explicit Model(const Tables& t) : T(t) {}

    // find_overhead (:413): launch overhead in microseconds.
    double overhead(int groups, int threads) const {
        std::vector<Key> k(T.overhead.size());
        for (size_t i = 0; i < k.size(); ++i) k[i] = {T.overhead[i].groups, T.overhead[i].threads};
        Bracket b = bracket(k, groups, threads);
        const auto& o = T.overhead;
        double sv = lerp((int)o[b.ss].threads, o[b.ss].us, (int)o[b.sl].threads, o[b.sl].us, threads);
        double lv = lerp((int)o[b.ls].threads, o[b.ls].us, (int)o[b.ll].threads, o[b.ll].us, threads);
        return lerp((int)o[b.ss].groups, sv, (int)o[b.ll].groups, lv, groups);
    }

    // maxflop_{bdia,dia,sell,ell} (:431-545): best method and its GFLOP/s.
    template <class Row, int N>
    void best_width(const std::vector<Row>& t, int rows, int width, int& meth, double& best) const {
        std::vector<Key> k(t.size());
        for (size_t i = 0; i < t.size(); ++i) k[i] = {t[i].dim, t[i].width};
        Bracket b = bracket(k, rows, width);
        meth = 0;
        best = 0.0f;
        for (int m = 0; m < N; ++m) {
            double v = bilinear(k, b, rows, width, [&](int i) { return t[i].gf[m]; });
            if (v > best) {
                meth = m;
                best = v;
            }
        }
    }

    // maxflop_{sbell,bell,bcsr} (:547-700): per block shape, the width is
    // the rounded number of blocks per row (nnzperrow / bw, rounded in
    // double, :566).
    template <class Row, int N>
    void best_block(const std::vector<Row>& t, int rows, int nnzperrow, int* meth, double* best) const {
        int off[9] = {0};
        int b = 1;
        for (size_t i = 0; i < t.size() && b < 8; ++i) {
            if ((int)t[i].bw == kBW[b] && (int)t[i].bh == kBH[b]) {
                off[b] = (int)i;
                ++b;
            }
        }
        off[8] = (int)t.size();
        for (b = 0; b < 8; ++b) {
            const int blocknum = (int)(((double)nnzperrow) / ((double)kBW[b]) + 0.5);
            std::vector<Key> k(off[b + 1] - off[b]);
            for (int i = off[b]; i < off[b + 1]; ++i) k[i - off[b]] = {t[i].dim, t[i].blocks};
            Bracket br = bracket(k, rows, blocknum);
            meth[b] = 0;
            best[b] = 0.0f;
            for (int m = 0; m < N; ++m) {
                double v = bilinear(k, br, rows, blocknum, [&](int i) { return t[off[b] + i].gf[m]; });
                if (v > best[b]) {
                    meth[b] = m;
                    best[b] = v;
                }
            }
        }
    }

    // maxflop_{csr,coo} (:702-802): rows come in (dim, width) groups, each
    // listing the same sequence of work-group counts.
    template <class Row, int N>
    void best_group(const std::vector<Row>& t, int rows, int width, int& meth, int& groups,
                    double& best) const {
        std::vector<int> off{0};
        for (size_t i = 1; i < t.size(); ++i)
            if (t[i].dim != t[i - 1].dim || t[i].width != t[i - 1].width) off.push_back((int)i);
        off.push_back((int)t.size());
        std::vector<Key> k(off.size() - 1);
        for (size_t i = 0; i + 1 < off.size(); ++i) k[i] = {t[off[i]].dim, t[off[i]].width};
        Bracket b = bracket(k, rows, width);
        meth = 0;
        best = 0.0f;
        const int choices = off[1] - off[0];
        for (int g = 0; g < choices; ++g) {
            for (int m = 0; m < N; ++m) {
                const Row &ss = t[off[b.ss] + g], &sl = t[off[b.sl] + g];
                const Row &ls = t[off[b.ls] + g], &ll = t[off[b.ll] + g];
                double sv = lerp((int)ss.width, ss.gf[m], (int)sl.width, sl.gf[m], width);
                double lv = lerp((int)ls.width, ls.gf[m], (int)ll.width, ll.gf[m], width);
                double v = lerp((int)ss.dim, sv, (int)ll.dim, lv, rows);
                if (v > best) {
                    meth = m;
                    groups = (int)ss.groups;
                    best = v;
                }
            }
        }
    }

    const Tables& T;
};

// nnz per row rounded in float, as at src/analyze.cpp:1313.
int nnz_per_row(long long nnz, long long rows) {
    return (int)(((float)nnz) / ((float)rows) + 0.5);
}

// ---- owned COO (the remainders passed between the extraction stages) ------

struct Coo {
    int width = 0, height = 0;
    std::vector<int> row, col;
    std::vector<float> val;
    bool allocated = false;  // mirrors "cooremain.coo_row_id != NULL"
    int nnz() const { return (int)row.size(); }
};

Coo from_c(const clspmv_coo_matrix& m) {
    Coo c;
    c.width = m.matinfo.width;
    c.height = m.matinfo.height;
    c.row.assign(m.coo_row_id, m.coo_row_id + m.matinfo.nnz);
    c.col.assign(m.coo_col_id, m.coo_col_id + m.matinfo.nnz);
    c.val.assign(m.coo_data, m.coo_data + m.matinfo.nnz);
    c.allocated = true;
    return c;
}

void sort_rows(Coo& c) {
    clspmv_coo_matrix m;
    m.matinfo = {c.width, c.height, c.nnz()};
    m.coo_row_id = c.row.data();
    m.coo_col_id = c.col.data();
    m.coo_data = c.val.data();
    if (!coo_sorted(m)) coo_sort(m);
}

void rowptr_of(const Coo& c, std::vector<int>& rp) { row_pointer(c.row.data(), c.nnz(), c.height, rp); }

class Partitioner {
  public:
    Partitioner(const Tables& t, clspmv_cocktail& out) : M(t), T(t), mat(out) {}

    void run(const Coo& A, float dia_mult, bool flat_partial, bool flat_full);

  private:
    Model M;
    const Tables& T;
    clspmv_cocktail& mat;
    FlopInfo info;

    double overhead_ms(int groups, int threads) const { return M.overhead(groups, threads) / 1000.0; }

    // update_info_* (:1311-1363)
    void update_dia(const Coo& A) {
        int w = nnz_per_row(A.nnz(), A.height);
        M.best_width<WidthRow<CLSPMV_BDIA_IMP_NUM>, CLSPMV_BDIA_IMP_NUM>(T.bdia, A.height, w, info.bdia_meth_num, info.bdia_max_flop);
        M.best_width<WidthRow<CLSPMV_DIA_IMP_NUM>, CLSPMV_DIA_IMP_NUM>(T.dia, A.height, w, info.dia_meth_num, info.dia_max_flop);
        double best = info.bdia_max_flop;
        if (info.dia_max_flop > best) best = info.dia_max_flop;
        info.max_dia_base_flop = best;
    }
    void block_flops(FlopInfo& f, int height, int w) const {
        M.best_block<BlockRow<CLSPMV_SBELL_IMP_NUM>, CLSPMV_SBELL_IMP_NUM>(T.sbell, height, w, f.sbell_meth_num, f.sbell_max_flop);
        M.best_block<BlockRow<CLSPMV_BELL_IMP_NUM>, CLSPMV_BELL_IMP_NUM>(T.bell, height, w, f.bell_meth_num, f.bell_max_flop);
        M.best_block<BlockRow<CLSPMV_BCSR_IMP_NUM>, CLSPMV_BCSR_IMP_NUM>(T.bcsr, height, w, f.bcsr_meth_num, f.bcsr_max_flop);
    }
    void update_block(const Coo& A) {
        block_flops(info, A.height, nnz_per_row(A.nnz(), A.height));
        double best = 0.0f;
        for (int i = 0; i < 8; ++i) {
            if (info.sbell_max_flop[i] > best) best = info.sbell_max_flop[i];
            if (info.bell_max_flop[i] > best) best = info.bell_max_flop[i];
            if (info.bcsr_max_flop[i] > best) best = info.bcsr_max_flop[i];
        }
        info.max_block_base_flop = best;
    }
    void sell_flops(FlopInfo& f, int rows, int w) const {
        M.best_width<WidthRow<CLSPMV_SELL_IMP_NUM>, CLSPMV_SELL_IMP_NUM>(T.sell, rows, w, f.sell_meth_num, f.sell_max_flop);
    }
    void ell_flops(FlopInfo& f, int rows, int w) const {
        M.best_width<WidthRow<CLSPMV_ELL_IMP_NUM>, CLSPMV_ELL_IMP_NUM>(T.ell, rows, w, f.ell_meth_num, f.ell_max_flop);
    }
    void csr_flops(FlopInfo& f, int rows, int w) const {
        M.best_group<GroupRow<CLSPMV_CSR_IMP_NUM>, CLSPMV_CSR_IMP_NUM>(T.csr, rows, w, f.csr_meth_num, f.csr_groupnum, f.csr_max_flop);
    }
    void coo_flops(FlopInfo& f, int rows, int w) const {
        M.best_group<GroupRow<CLSPMV_COO_IMP_NUM>, CLSPMV_COO_IMP_NUM>(T.coo, rows, w, f.coo_meth_num, f.coo_groupnum, f.coo_max_flop);
    }
    void flat_flops(FlopInfo& f, int rows, int w) const {
        sell_flops(f, rows, w);
        ell_flops(f, rows, w);
        csr_flops(f, rows, w);
        coo_flops(f, rows, w);
        double best = f.sell_max_flop;
        if (f.ell_max_flop > best) best = f.ell_max_flop;
        if (f.csr_max_flop > best) best = f.csr_max_flop;
        if (f.coo_max_flop > best) best = f.coo_max_flop;
        f.max_flat_base_flop = best;
    }
    void update_flat(const Coo& A) { flat_flops(info, A.height, nnz_per_row(A.nnz(), A.height)); }

    // --- diagonal family (:804-1524)
    bool extract_dia(const Coo& A, Coo& rem);
    void bdia_dia_split(const Coo& A, Coo& rem, const std::vector<std::pair<int, int>>& sel,
                        int threshold, int mode, int least_band, int bdia_meth, int dia_meth);

    // --- block family (:1531-2844)
    bool extract_block(const Coo& A, Coo& rem, bool flat_partial, bool flat_full);
    void take_blocks(const Coo& A, Coo& rem, const std::vector<int>& rp, int wh, int threshold,
                     int fmt, const std::vector<int>& limit, int nnz_taken_hint);
    void sbell_or_bell(const Coo& A, Coo& rem, const std::vector<int>& rp, int wh,
                       const std::vector<std::pair<int, int>>& count, std::pair<int, int> sum,
                       bool full);

    // --- flat family (:2846-3711)
    double estimate_coo_time(int nnz, const FlopInfo& f) const;
    double estimate_csr_time(const std::vector<int>& rownnz, int nnz, const FlopInfo& f) const;
    double estimate_csr_coo_time(const std::vector<int>& rownnz, int nnz, FlopInfo& f) const;
    double estimate_sell_ell_time(const std::vector<int>& rownnz, int nnz, const FlopInfo& f,
                                  std::vector<int>& remain, int& nnzremain) const;
    double estimate_flat_time(const std::vector<int>& rownnz, int nnz) const;
    void extract_flat(const Coo& A);
    void extract_sell_ell(const Coo& A, Coo& rem);
    void extract_sell(const Coo& A, Coo& rem, const std::vector<int>& count, int sh,
                      const std::vector<int>& width, int slice_nnz);
    void extract_ell(const Coo& A, Coo& rem, const std::vector<int>& count, int ellnum, int ellnnz);
    void extract_csr_coo(const Coo& A);
    void extract_csr(const Coo& A);
    void extract_coo(const Coo& A);
};

// ---- diagonals -------------------------------------------------------------

struct BdiaVsDia {
    int min_width = 1000000, range_s = 0, range_l = 0;
    double bdia_single = 0, dia_single = 0;
};

// find_min_bdia_width (:828): in the table dimension nearest the matrix's,
// the narrowest band width for which BDIA beats DIA.
BdiaVsDia min_bdia_width(const Tables& T, int dimension) {
    std::vector<Key> k(T.bdia.size());
    for (size_t i = 0; i < k.size(); ++i) k[i] = {T.bdia[i].dim, T.bdia[i].width};
    std::vector<int> dims, offs;
    dim_runs(k, dims, offs);
    int ds, dl;
    dim_bracket(dims, dimension, ds, dl);
    int near = ds;
    if (ds != dl) {
        int dist_s = (int)((unsigned)dimension - T.bdia[offs[ds]].dim);
        int dist_l = (int)(T.bdia[offs[dl]].dim - (unsigned)dimension);
        near = dist_s < dist_l ? ds : dl;
    }
    BdiaVsDia r;
    r.range_s = offs[near];
    r.range_l = offs[near + 1];
    for (int i = r.range_s; i < r.range_l; ++i) {
        double mb = 0.0f, md = 0.0f;
        for (int j = 0; j < CLSPMV_BDIA_IMP_NUM; ++j)
            if (T.bdia[i].gf[j] > mb) mb = T.bdia[i].gf[j];
        for (int j = 0; j < CLSPMV_DIA_IMP_NUM; ++j)
            if (T.dia[i].gf[j] > md) md = T.dia[i].gf[j];
        if (T.bdia[i].width == 1) {
            r.bdia_single = mb;
            r.dia_single = md;
        }
        if (mb > md) {
            r.min_width = (int)T.bdia[i].width;
            return r;
        }
    }
    r.min_width = 1000000;
    return r;
}

// bdia_gflop / dia_gflop (:913, :951): best method of the row at `width`
// (clamped, no interpolation).
template <class Row, int N>
void row_best(const std::vector<Row>& t, int rs, int rl, int width, double& gflop, int& meth) {
    const unsigned w = (unsigned)width;
    int index = 0;
    if (w < t[rs].width) {
        index = rs;
    } else if (w > t[rl - 1].width) {
        index = rl - 1;
    } else {
        for (int i = rs; i < rl; ++i) {
            if (w == t[i].width) {
                index = i;
                break;
            }
            if (i < rl - 1 && t[i].width < w && t[i + 1].width > w) {
                index = i;
                break;
            }
        }
    }
    gflop = 0.0f;
    meth = 0;
    for (int m = 0; m < N; ++m)
        if (t[index].gf[m] > gflop) {
            gflop = t[index].gf[m];
            meth = m;
        }
}

bool Partitioner::extract_dia(const Coo& A, Coo& rem) {
    if (A.nnz() == 0) return false;
    update_dia(A);
    update_flat(A);
    const double margin = 1.05;
    const double ratio = info.max_flat_base_flop / info.max_dia_base_flop;
    const int least = (int)(ratio * (double)A.height * margin);
    if (least > A.height) {
        logf("dia: none (a diagonal needs more than %d nonzeros)", least);
        return false;
    }
    const int threshold = A.height - least;
    // count_dia_nnz (:804): entries per diagonal offset in [-threshold, threshold].
    std::vector<int> count(2 * (size_t)threshold + 1, 0);
    for (int i = 0; i < A.nnz(); ++i) {
        long long o = (long long)A.col[i] - A.row[i] + threshold;
        if (o >= 0 && o < (long long)count.size()) count[o]++;
    }
    std::vector<std::pair<int, int>> valid;  // (count, index)
    for (size_t i = 0; i < count.size(); ++i)
        if (count[i] > least) valid.push_back({count[i], (int)i});
    std::sort(valid.begin(), valid.end(), std::greater<std::pair<int, int>>());
    const double ovh = M.overhead(A.height / CLSPMV_WORK_GROUP_SIZE, CLSPMV_WORK_GROUP_SIZE) / 1000.0;

    // How many of the densest diagonals to take (:1409-1430).
    int current = 0, dia_num = 0;
    double best = 1000000.0f;
    for (int i = 0; i <= CLSPMV_MAX_DIA_NUM && i <= (int)valid.size(); ++i) {
        if (i > 0) current += valid[i - 1].first;
        const int remain = A.nnz() - current;
        double t = ((double)i) * ((double)A.height) / 1000000.0 / info.max_dia_base_flop +
                   ((double)remain) / 1000000.0 / info.max_flat_base_flop;
        if (i > 0) t += ovh;
        if (remain > 0) t += ovh;
        if (t < best) {
            best = t;
            dia_num = i;
        }
    }
    logf("dia: %d candidate diagonals, taking %d", (int)valid.size(), dia_num);
    if (dia_num == 0) return false;

    std::vector<std::pair<int, int>> sel;  // (index, count), ascending offset
    for (int i = 0; i < dia_num; ++i) sel.push_back({valid[i].second, valid[i].first});
    std::sort(sel.begin(), sel.end());

    BdiaVsDia bvd = min_bdia_width(T, A.height);
    const int least_band = bvd.min_width;
    int band_num = 0, cur = 1;
    for (size_t i = 1; i < sel.size(); ++i) {
        if (sel[i].first == sel[i - 1].first + 1) {
            cur++;
        } else {
            if (cur >= least_band) band_num += cur;
            cur = 1;
        }
        if (i == sel.size() - 1 && cur >= least_band) band_num += cur;
    }
    const int remain_dia = dia_num - band_num;
    double bdia_part = 0.0001f, dia_part = 0.0001f, dia_all = 0.0001f;
    int bdia_meth = 0, dia_part_meth = 0, dia_all_meth = 0;
    if (band_num > 0)
        row_best<WidthRow<CLSPMV_BDIA_IMP_NUM>, CLSPMV_BDIA_IMP_NUM>(T.bdia, bvd.range_s, bvd.range_l, band_num, bdia_part, bdia_meth);
    if (remain_dia > 0)
        row_best<WidthRow<CLSPMV_DIA_IMP_NUM>, CLSPMV_DIA_IMP_NUM>(T.dia, bvd.range_s, bvd.range_l, remain_dia, dia_part, dia_part_meth);
    row_best<WidthRow<CLSPMV_DIA_IMP_NUM>, CLSPMV_DIA_IMP_NUM>(T.dia, bvd.range_s, bvd.range_l, dia_num, dia_all, dia_all_meth);
    const double h = (double)A.height;
    const double bdia_only = ((double)band_num * h) / 1000000.0 / bdia_part +
                             ((double)remain_dia * h) / 1000000.0 /
                                 (dia_part * bvd.bdia_single / bvd.dia_single) + ovh;
    const double dia_only = ((double)dia_num * h) / 1000000.0 / dia_all + ovh;
    const double both = ((double)band_num * h) / 1000000.0 / bdia_part +
                        ((double)remain_dia * h) / 1000000.0 / dia_part + 2 * ovh;
    if (bdia_only <= dia_only && bdia_only <= both) {
        logf("dia: BDIA only (%d diagonals in bands of >= %d)", band_num, least_band);
        bdia_dia_split(A, rem, sel, threshold, 0, least_band, bdia_meth, 0);
    } else if (dia_only <= bdia_only && dia_only <= both) {
        logf("dia: DIA only (%d diagonals)", dia_num);
        bdia_dia_split(A, rem, sel, threshold, 1, least_band, 0, dia_all_meth);
    } else {
        logf("dia: BDIA for bands of >= %d, DIA for the rest", least_band);
        bdia_dia_split(A, rem, sel, threshold, 2, least_band, bdia_meth, dia_part_meth);
    }
    return true;
}

// extract_bdia_only (mode 0, :989), extract_dia_only (mode 1, :1084),
// extract_bdia_and_dia (mode 2, :1147).  Selected diagonals become BDIA
// bands (runs of consecutive offsets) and/or DIA diagonals; the remainder
// keeps the input order.
void Partitioner::bdia_dia_split(const Coo& A, Coo& rem, const std::vector<std::pair<int, int>>& sel,
                                 int threshold, int mode, int least_band, int bdia_meth, int dia_meth) {
    const int len = round_up(A.height, CLSPMV_GPU_ALIGNMENT);
    std::vector<int> off(sel.size());
    int taken = 0;
    for (size_t i = 0; i < sel.size(); ++i) {
        off[i] = sel[i].first - threshold;
        taken += sel[i].second;
    }
    // Bands over all selected offsets.
    std::vector<int> band_first, band_count, band_nnz;
    for (size_t i = 0; i < off.size(); ++i) {
        if (i > 0 && off[i] - off[i - 1] == 1) {
            band_count.back()++;
            band_nnz.back() += sel[i].second;
        } else {
            band_first.push_back(off[i]);
            band_count.push_back(1);
            band_nnz.push_back(sel[i].second);
        }
    }
    std::vector<int> dia_off, bdia_off, bdia_first, bdia_ptr{0};
    int dia_nnz = 0, bdia_nnz = 0;
    for (size_t b = 0; b < band_first.size(); ++b) {
        const bool to_bdia = mode == 0 || (mode == 2 && band_count[b] >= least_band);
        for (int j = 0; j < band_count[b]; ++j) (to_bdia ? bdia_off : dia_off).push_back(band_first[b] + j);
        if (to_bdia) {
            bdia_first.push_back(band_first[b]);
            bdia_ptr.push_back(bdia_ptr.back() + band_count[b]);
            bdia_nnz += band_nnz[b];
        } else {
            dia_nnz += band_nnz[b];
        }
    }
    // Offset -> slot lookup (DIA first, then BDIA, like the reference's
    // two linear searches).
    std::vector<std::pair<int, int>> dia_slot, bdia_slot;  // (offset, slot)
    for (size_t j = 0; j < dia_off.size(); ++j) dia_slot.push_back({dia_off[j], (int)j});
    for (size_t j = 0; j < bdia_off.size(); ++j) bdia_slot.push_back({bdia_off[j], (int)j});
    std::sort(dia_slot.begin(), dia_slot.end());
    std::sort(bdia_slot.begin(), bdia_slot.end());
    auto find = [](const std::vector<std::pair<int, int>>& v, int o) {
        auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(o, INT32_MIN));
        return (it != v.end() && it->first == o) ? it->second : -1;
    };
    std::vector<float> dia_data(dia_off.size() * (size_t)len, 0.0f);
    std::vector<float> bdia_data(bdia_off.size() * (size_t)len, 0.0f);
    rem = Coo();
    rem.width = A.width;
    rem.height = A.height;
    rem.allocated = true;
    const int keep = A.nnz() - taken;
    rem.row.reserve(keep);
    rem.col.reserve(keep);
    rem.val.reserve(keep);
    for (int i = 0; i < A.nnz(); ++i) {
        const int o = A.col[i] - A.row[i];
        int j = find(dia_slot, o);
        if (j >= 0) {
            dia_data[(size_t)j * len + A.row[i]] = A.val[i];
            continue;
        }
        j = find(bdia_slot, o);
        if (j >= 0) {
            bdia_data[(size_t)j * len + A.row[i]] = A.val[i];
            continue;
        }
        rem.row.push_back(A.row[i]);
        rem.col.push_back(A.col[i]);
        rem.val.push_back(A.val[i]);
    }
    if (mode == 0 || mode == 2) {
        mat.ifusebdia = true;
        mat.bdia_meth_num = bdia_meth;
        clspmv_bdia_matrix& b = mat.bdia;
        b.matinfo = {A.width, A.height, bdia_nnz};
        b.bdia_length = A.height;
        b.bdia_length_aligned = len;
        b.bdia_band_num = (int)bdia_first.size();
        b.bdia_offsets = host_copy(bdia_first);
        b.bdia_bptr = host_copy(bdia_ptr);
        b.bdia_data = host_copy(bdia_data);
    }
    if (mode == 1 || mode == 2) {
        mat.ifusedia = true;
        mat.dia_meth_num = dia_meth;
        clspmv_dia_matrix& d = mat.dia;
        d.matinfo = {A.width, A.height, dia_nnz};
        d.dia_length = A.height;
        d.dia_length_aligned = len;
        d.dia_num = (int)dia_off.size();
        d.dia_offsets = host_copy(dia_off);
        d.dia_data = host_copy(dia_data);
    }
}

// ---- blocks ------------------------------------------------------------------

struct BlockInfo { int partial_blocks = 0, partial_nnz = 0, full_blocks = 0, full_nnz = 0; };

// count_blocks (:1582): per block shape and block row, the number of blocks
// and entries, all blocks ("full") and those with more than threshold[i]
// entries ("partial", i.e. dense enough).
void count_blocks(const Coo& A, const std::vector<int>& rp, const std::vector<int>& threshold,
                  std::vector<std::vector<BlockInfo>>& out) {
    out.assign(8, {});
    for (int i = 0; i < 8; ++i) out[i].assign((A.height + kBH[i] - 1) / kBH[i], BlockInfo());
    const int strips = (A.height + 7) / 8;
    std::vector<uint64_t> keys;
    for (int s = 0; s < strips; ++s) {
        const int start = rp[s * 8];
        const int end = rp[std::min(s * 8 + 8, A.height)];
        if (end <= start) continue;
        for (int i = 0; i < 8; ++i) {
            keys.resize(end - start);
            for (int k = start; k < end; ++k)
                keys[k - start] = ((uint64_t)(uint32_t)(A.row[k] / kBH[i]) << 32) |
                                  (uint32_t)(A.col[k] / kBW[i]);
            std::sort(keys.begin(), keys.end());
            size_t a = 0;
            while (a < keys.size()) {
                size_t b = a + 1;
                while (b < keys.size() && keys[b] == keys[a]) ++b;
                const int brow = (int)(keys[a] >> 32);
                const int n = (int)(b - a);
                BlockInfo& bi = out[i][brow];
                bi.full_blocks++;
                bi.full_nnz += n;
                if (n > threshold[i]) {
                    bi.partial_blocks++;
                    bi.partial_nnz += n;
                }
                a = b;
            }
        }
    }
}

bool Partitioner::extract_block(const Coo& A, Coo& rem, bool flat_partial, bool flat_full) {
    if (A.nnz() == 0) return false;
    update_block(A);
    update_flat(A);
    std::vector<int> rp;
    rowptr_of(A, rp);
    // find_thre_max_flop (:1558)
    std::vector<double> best_flop(8, 0.0f);
    std::vector<int> threshold(8, 0);
    for (int i = 0; i < 8; ++i) {
        if (info.sbell_max_flop[i] > best_flop[i]) best_flop[i] = info.sbell_max_flop[i];
        if (info.bell_max_flop[i] > best_flop[i]) best_flop[i] = info.bell_max_flop[i];
        if (info.bcsr_max_flop[i] > best_flop[i]) best_flop[i] = info.bcsr_max_flop[i];
        threshold[i] = (int)((double)kBH[i] * (double)kBW[i] * info.max_flat_base_flop / best_flop[i]);
        if (threshold[i] > kBH[i] * kBW[i] - 1) threshold[i] = kBH[i] * kBW[i] - 1;
    }
    std::vector<std::vector<BlockInfo>> bc;
    count_blocks(A, rp, threshold, bc);

    std::vector<double> block_over(8);
    for (int i = 0; i < 8; ++i)
        block_over[i] = M.overhead(A.height / (kBH[i] * CLSPMV_SELL_GROUP_SIZE), CLSPMV_SELL_GROUP_SIZE) / 1000.0;
    const double flatover = M.overhead(A.height / CLSPMV_WORK_GROUP_SIZE, CLSPMV_WORK_GROUP_SIZE) / 1000.0;
    std::vector<BlockInfo> sum(8);
    for (int i = 0; i < 8; ++i)
        for (const BlockInfo& b : bc[i]) {
            sum[i].partial_blocks += b.partial_blocks;
            sum[i].partial_nnz += b.partial_nnz;
            sum[i].full_blocks += b.full_blocks;
            sum[i].full_nnz += b.full_nnz;
        }
    double best_full = 1000000.0f, best_partial = 1000000.0f;
    int full_id = 0, partial_id = 0;
    for (int i = 0; i < 8; ++i) {
        const double full = ((double)sum[i].full_blocks) * ((double)kBH[i]) * ((double)kBW[i]) /
                                1000000.0 / best_flop[i] + block_over[i];
        if (full < best_full) {
            best_full = full;
            full_id = i;
        }
        // Entries outside dense blocks, spread evenly over each block row's rows.
        std::vector<int> rownnz(A.height, 0);
        int flatnnz = 0;
        for (size_t j = 0; j < bc[i].size(); ++j) {
            const int diff = bc[i][j].full_nnz - bc[i][j].partial_nnz;
            const int r0 = (int)j * kBH[i];
            const int r1 = std::min((int)(j + 1) * kBH[i], A.height);
            const int n = r1 - r0;
            const int avg = diff / n;
            for (int k = r0; k < r1 - 1; ++k) rownnz[k] = avg;
            rownnz[r1 - 1] = diff - avg * (n - 1);
            flatnnz += diff;
        }
        double partial = ((double)sum[i].partial_blocks) * ((double)kBH[i]) * ((double)kBW[i]) /
                             1000000.0 / best_flop[i] + block_over[i];
        if (flat_partial)
            partial += estimate_flat_time(rownnz, flatnnz);
        else
            partial += ((double)(A.nnz() - sum[i].partial_nnz)) / 1000000.0 / info.max_flat_base_flop + flatover;
        if (partial < best_partial) {
            best_partial = partial;
            partial_id = i;
        }
    }
    double fullflat = ((double)A.nnz()) / 1000000.0 / info.max_flat_base_flop + flatover;
    fullflat *= 1.02;
    if (flat_full) {
        std::vector<int> rownnz(A.height);
        for (int i = 0; i < A.height; ++i) rownnz[i] = rp[i + 1] - rp[i];
        fullflat = estimate_flat_time(rownnz, A.nnz());
    }
    if (fullflat <= best_full && fullflat <= best_partial) {
        logf("block: none (flat estimate %.6f ms)", fullflat);
        return false;
    }
    const bool full = best_full <= best_partial;
    const int wh = full ? full_id : partial_id;
    logf("block: %s blocks of %dx%d", full ? "all" : "dense", kBH[wh], kBW[wh]);
    const BlockInfo& s = sum[wh];
    // extract_full_block (:2676) / extract_partial_block (:2700)
    if (info.bcsr_max_flop[wh] >= info.sbell_max_flop[wh] && info.bcsr_max_flop[wh] >= info.bell_max_flop[wh]) {
        take_blocks(A, rem, rp, wh, full ? 0 : threshold[wh], CLSPMV_FMT_BCSR, {},
                    full ? s.full_nnz : s.partial_nnz);
        return true;
    }
    std::vector<std::pair<int, int>> count(bc[wh].size());
    for (size_t j = 0; j < bc[wh].size(); ++j)
        count[j] = full ? std::make_pair(bc[wh][j].full_blocks, bc[wh][j].full_nnz)
                        : std::make_pair(bc[wh][j].partial_blocks, bc[wh][j].partial_nnz);
    std::pair<int, int> total = full ? std::make_pair(s.full_blocks, s.full_nnz)
                                     : std::make_pair(s.partial_blocks, s.partial_nnz);
    sbell_or_bell(A, rem, rp, wh, count, total, full);
    return true;
}

// extract_full_sbell_bell (:2435, full = true) and extract_sbell_bell
// (:2521, full = false): choose the block slots per row (BELL) or per slice
// of 32 block rows (SBELL).
void Partitioner::sbell_or_bell(const Coo& A, Coo& rem, const std::vector<int>& rp, int wh,
                                const std::vector<std::pair<int, int>>& count,
                                std::pair<int, int> sum, bool full) {
    if (A.nnz() == 0) return;
    const int w = nnz_per_row(sum.second, A.height);
    M.best_block<BlockRow<CLSPMV_SBELL_IMP_NUM>, CLSPMV_SBELL_IMP_NUM>(T.sbell, A.height, w, info.sbell_meth_num, info.sbell_max_flop);
    M.best_block<BlockRow<CLSPMV_BELL_IMP_NUM>, CLSPMV_BELL_IMP_NUM>(T.bell, A.height, w, info.bell_meth_num, info.bell_max_flop);
    const int bh = kBH[wh], bw = kBW[wh], bw4 = bw / 4;
    int maxb = 0;
    for (const auto& c : count) maxb = std::max(maxb, c.first);
    int brownum = A.height / bh + (A.height % bh != 0);
    const double bell_over = M.overhead(brownum / CLSPMV_BELL_GROUP_SIZE, CLSPMV_BELL_GROUP_SIZE) / 1000.0;
    const int newlength = round_up(4 * brownum, CLSPMV_GPU_ALIGNMENT);
    const double avgnnz = ((double)sum.second / (double)sum.first);

    double best_ell = 1000000.0;
    int best_ell_num = 0;
    double remain_flop = info.max_flat_base_flop, remain_over = 0.0;
    if (full) {
        best_ell = 0.0f;
        best_ell += ((double)maxb * (double)newlength * (double)bh * (double)bw4) / 1000000.0 /
                        info.bell_max_flop[wh] + bell_over;
        best_ell_num = maxb;
    } else {
        remain_over = M.overhead(A.height / CLSPMV_WORK_GROUP_SIZE, CLSPMV_WORK_GROUP_SIZE) / 1000.0;
        std::vector<int> hist(maxb + 1, 0);
        for (const auto& c : count) hist[c.first]++;
        int bellnnz = 0;
        for (size_t i = 0; i < hist.size(); ++i) {
            brownum -= hist[i];
            double t = 0.0f;
            if (i > 0)
                t += ((double)i * (double)newlength * (double)bh * (double)bw4) / 1000000.0 /
                         info.bell_max_flop[wh] + bell_over;
            if (bellnnz < A.nnz())
                t += ((double)(A.nnz() - bellnnz)) / 1000000.0 / remain_flop + remain_over;
            if (t < best_ell) {
                best_ell = t;
                best_ell_num = (int)i;
            }
            bellnnz += (int)((double)brownum * avgnnz);
        }
    }

    const int sh = CLSPMV_WARPSIZE;
    brownum = A.height / bh + (A.height % bh != 0);
    const int nslice = brownum / sh + (brownum % sh != 0);
    std::vector<int> w_full(nslice, 0), w_part(nslice, 0);
    int part_nnz = 0;
    const int full_nnz = sum.second;
    for (int i = 0; i < nslice; ++i) {
        std::vector<int> sw;
        for (int j = i * sh; j < (i + 1) * sh && j < (int)count.size(); ++j) sw.push_back(count[j].first);
        std::sort(sw.begin(), sw.end());
        w_full[i] = sw.back();
        if (full) continue;
        int partial = (int)((double)sw[0] * (double)sw.size() * avgnnz);
        int current = partial;
        w_part[i] = sw[0];
        for (size_t j = 1; j < sw.size(); ++j) {
            partial += (int)((double)(sw[j] - sw[j - 1]) * (double)(sw.size() - j) * avgnnz);
            const double ratio = remain_flop / info.sbell_max_flop[wh];
            const size_t cells = sw.size() * (size_t)sw[j] * (size_t)bw * (size_t)bh;
            if (partial > (int)(((double)cells) * ratio)) {
                w_part[i] = sw[j];
                current = partial;
            } else {
                break;
            }
        }
        part_nnz += current;
    }
    const double sbell_over = M.overhead(brownum / CLSPMV_SELL_GROUP_SIZE, CLSPMV_SELL_GROUP_SIZE) / 1000.0;
    int sb_full = 0, sb_part = 0;
    for (int i = 0; i < nslice; ++i) {
        sb_full += sh * w_full[i];
        sb_part += sh * w_part[i];
    }
    double t_full = ((double)(sb_full * bw * bh)) / 1000000.0 / info.sbell_max_flop[wh] + sbell_over;
    if (full) {
        if (best_ell < t_full) {
            logf("block: BELL, %d slots", best_ell_num);
            take_blocks(A, rem, rp, wh, 0, CLSPMV_FMT_BELL, {best_ell_num}, 0);
        } else {
            logf("block: SBELL with full slices");
            take_blocks(A, rem, rp, wh, 0, CLSPMV_FMT_SBELL, w_full, 0);
        }
        return;
    }
    t_full += ((double)(A.nnz() - full_nnz)) / 1000000.0 / remain_flop + remain_over;
    const double t_part = ((double)(sb_part * bw * bh)) / 1000000.0 / info.sbell_max_flop[wh] + sbell_over +
                          ((double)(A.nnz() - part_nnz)) / 1000000.0 / remain_flop + remain_over;
    if (best_ell < t_full && best_ell < t_part) {
        logf("block: BELL, %d slots", best_ell_num);
        take_blocks(A, rem, rp, wh, 0, CLSPMV_FMT_BELL, {best_ell_num}, 0);
    } else if (t_part < t_full) {
        logf("block: SBELL with trimmed slices");
        take_blocks(A, rem, rp, wh, 0, CLSPMV_FMT_SBELL, w_part, 0);
    } else {
        logf("block: SBELL with full slices");
        take_blocks(A, rem, rp, wh, 0, CLSPMV_FMT_SBELL, w_full, 0);
    }
}

// extract_bcsr_part (:2113), extract_bell_part (:1888), extract_sbell_part
// (:1660): walk each block row's blocks left to right; a block is kept when
// it holds more than `threshold` entries and (BELL/SBELL) the row still has
// a free slot (limit[0] slots per row for BELL, limit[slice] for SBELL).
// Entries of the other blocks form the remainder, sorted by (row, col).
void Partitioner::take_blocks(const Coo& A, Coo& rem, const std::vector<int>& rp, int wh,
                              int threshold, int fmt, const std::vector<int>& limit,
                              int nnz_taken_hint) {
    if (A.nnz() == 0) return;
    const int bh = kBH[wh], bw = kBW[wh], bs = bh * bw, bw4 = bw / 4;
    rem = Coo();
    rem.width = A.width;
    rem.height = A.height;
    if (fmt == CLSPMV_FMT_BELL && limit[0] == 0) {
        rem.row = A.row;
        rem.col = A.col;
        rem.val = A.val;
        rem.allocated = true;
        logf("block: BELL with no slots, nothing taken");
        return;
    }
    rem.allocated = true;
    const int nbr = A.height / bh + (A.height % bh != 0);
    const int sh = CLSPMV_WARPSIZE;
    std::vector<int> brp(nbr + 1, 0), bcol;
    std::vector<float> tiles;
    BlockRowScratch scratch;
    for (int br = 0; br < nbr; ++br) {
        int kept = 0;
        const int slot_limit = fmt == CLSPMV_FMT_BELL ? limit[0]
                             : fmt == CLSPMV_FMT_SBELL ? limit[br / sh] : INT32_MAX;
        for_each_block(A.row.data(), A.col.data(), A.val.data(), rp, A.height, br, bh, bw, scratch,
                       [&](int c, const BlockEntry* b, const BlockEntry* e) {
                           const int n = (int)(e - b);
                           if (n > threshold && kept + 1 <= slot_limit) {
                               ++kept;
                               bcol.push_back(c);
                               size_t off = tiles.size();
                               tiles.resize(off + bs);
                               fill_tile(b, e, bh, bw, &tiles[off]);
                           } else {
                               for (const BlockEntry* p = b; p != e; ++p) {
                                   rem.row.push_back(p->row);
                                   rem.col.push_back(p->col);
                                   rem.val.push_back(p->val);
                               }
                           }
                       });
        brp[br + 1] = brp[br] + kept;
    }
    const int nblocks = (int)bcol.size();
    const int taken = A.nnz() - rem.nnz();
    (void)nnz_taken_hint;
    sort_rows(rem);

    if (fmt == CLSPMV_FMT_BCSR) {
        mat.ifusebcsr = true;
        mat.bcsr_meth_num = info.bcsr_meth_num[wh];
        clspmv_b4csr_matrix& m = mat.bcsr;
        m.matinfo = {A.width, A.height, taken};
        m.b4csr_bwidth = bw;
        m.b4csr_bheight = bh;
        m.b4csr_row_num = nbr;
        const int aligned = round_up(nblocks * 4, CLSPMV_GPU_ALIGNMENT);
        m.b4csr_aligned_size = aligned;
        m.b4csr_block_num = nblocks;
        m.b4csr_row_ptr = host_copy(brp);
        m.b4csr_col_id = host_alloc<int>(nblocks);
        m.b4csr_data = host_alloc_zero<float>((size_t)aligned * bh * bw4);
        const size_t plane = (size_t)aligned * bh;
        for (int i = 0; i < nblocks; ++i) {
            m.b4csr_col_id[i] = bcol[i] / 4;
            for (int h = 0; h < bh; ++h)
                for (int w4 = 0; w4 < bw4; ++w4)
                    for (int q = 0; q < 4; ++q)
                        m.b4csr_data[w4 * plane + (size_t)h * aligned + (size_t)i * 4 + q] =
                            tiles[(size_t)i * bs + h * bw + w4 * 4 + q];
        }
    } else if (fmt == CLSPMV_FMT_BELL) {
        const int ell = limit[0];
        mat.ifusebell = true;
        mat.bell_meth_num = info.bell_meth_num[wh];
        clspmv_b4ell_matrix& m = mat.bell;
        m.matinfo = {A.width, A.height, taken};
        m.b4ell_bwidth = bw;
        m.b4ell_bheight = bh;
        m.b4ell_row_num = nbr;
        m.b4ell_block_num = ell;
        const int hal = round_up(nbr, CLSPMV_GPU_ALIGNMENT);
        const int f4 = round_up(4 * nbr, CLSPMV_GPU_ALIGNMENT);
        m.b4ell_height_aligned = hal;
        m.b4ell_float4_aligned = f4;
        m.b4ell_col_id = host_alloc_zero<int>((size_t)hal * ell);
        m.b4ell_data = host_alloc_zero<float>((size_t)f4 * bh * bw4 * ell);
        const size_t slot = (size_t)f4 * bw4 * bh, w4s = (size_t)f4 * bh;
        for (int r = 0; r < nbr; ++r) {
            int last = 0, s = 0;
            for (int j = brp[r]; j < brp[r + 1]; ++j, ++s) {
                last = bcol[j] / 4;
                m.b4ell_col_id[r + (size_t)s * hal] = last;
                for (int h = 0; h < bh; ++h)
                    for (int w4 = 0; w4 < bw4; ++w4)
                        for (int q = 0; q < 4; ++q)
                            m.b4ell_data[s * slot + (size_t)h * f4 + w4 * w4s + (size_t)r * 4 + q] =
                                tiles[(size_t)j * bs + h * bw + w4 * 4 + q];
            }
            for (; s < ell; ++s) m.b4ell_col_id[r + (size_t)s * hal] = last;
        }
    } else {
        const int nslice = nbr / sh + (nbr % sh != 0);
        mat.ifusesbell = true;
        mat.sbell_meth_num = info.sbell_meth_num[wh];
        clspmv_sbell_matrix& m = mat.sbell;
        m.matinfo = {A.width, A.height, taken};
        m.sbell_bwidth = bw;
        m.sbell_bheight = bh;
        m.sbell_slice_height = sh;
        m.sbell_slice_num = nslice;
        m.sbell_row_num = nbr;
        std::vector<int> ptr(nslice + 1, 0);
        for (int i = 0; i < nslice; ++i) ptr[i + 1] = ptr[i] + limit[i] * sh;
        const int total = ptr[nslice];
        m.sbell_slice_ptr = host_copy(ptr);
        m.sbell_col_id = host_alloc_zero<int>(total);
        m.sbell_data = host_alloc_zero<float>((size_t)total * bs);
        const size_t slot = (size_t)sh * bw4 * bh * 4, hs = (size_t)sh * 4, w4s = (size_t)sh * bh * 4;
        for (int r = 0; r < nbr; ++r) {
            const int sl = r / sh, i = r % sh;
            float* data = m.sbell_data + (size_t)ptr[sl] * bs;
            int last = 0, s = 0;
            for (int j = brp[r]; j < brp[r + 1]; ++j, ++s) {
                last = bcol[j] / 4;
                m.sbell_col_id[ptr[sl] + i + (size_t)s * sh] = last;
                for (int h = 0; h < bh; ++h)
                    for (int w4 = 0; w4 < bw4; ++w4)
                        for (int q = 0; q < 4; ++q)
                            data[s * slot + h * hs + w4 * w4s + (size_t)i * 4 + q] =
                                tiles[(size_t)j * bs + h * bw + w4 * 4 + q];
            }
            for (; s < limit[sl]; ++s) m.sbell_col_id[ptr[sl] + i + (size_t)s * sh] = last;
        }
    }
}

// ---- flat formats --------------------------------------------------------------

double Partitioner::estimate_coo_time(int nnz, const FlopInfo& f) const {
    if (nnz == 0) return 0.0f;
    const double over = M.overhead(f.coo_groupnum, CLSPMV_COO_GROUP_SIZE) / 1000.0;
    return ((double)nnz) / 1000000.0 / f.coo_max_flop + over;
}

// Entries handled by the busiest work-group of the CSR kernel under the
// reference's static row assignment (src/analyze.cpp:2887-2926).
static int csr_max_group_nnz(const std::vector<int>& rownnz, int meth, int groups) {
    std::vector<int> per(groups > 0 ? groups : 0, 0);
    const long long rows = (long long)rownnz.size();
    if (meth == 0) {
        const int hgt = CLSPMV_CSR_VEC_GROUP_SIZE / CLSPMV_WARPSIZE;
        const long long warps = (long long)groups * hgt;
        if (warps > 0) {
            const long long iters = rows / warps + (rows % warps != 0);
            for (long long i = 0; i < iters; ++i)
                for (int g = 0; g < groups; ++g)
                    for (int j = 0; j < hgt; ++j) {
                        long long id = i * warps + (long long)g * hgt + j;
                        if (id >= rows) break;
                        per[g] += rownnz[id];
                    }
        }
    } else if (meth == 1) {
        // Group g takes rows g, g + groups, ...  The reference forms the row
        // index as a 32-bit int (i * groups + g, src/analyze.cpp:2919); the
        // same wrap-around is reproduced so that very large products give the
        // reference's sums too.
        for (long long i = 0; i < rows; ++i) {
            for (int g = 0; g < groups; ++g) {
                const int32_t id = (int32_t)((uint32_t)i * (uint32_t)groups + (uint32_t)g);
                if (id < 0 || id >= rows) break;  // negative ids compare as huge
                per[g] += rownnz[id];
            }
        }
    }
    int mx = 0;
    for (int v : per) mx = std::max(mx, v);
    return mx;
}

double Partitioner::estimate_csr_time(const std::vector<int>& rownnz, int nnz, const FlopInfo& f) const {
    if (nnz == 0) return 0.0f;
    const int newnnz = csr_max_group_nnz(rownnz, f.csr_meth_num, f.csr_groupnum) * f.csr_groupnum;
    const double over = M.overhead(f.csr_groupnum, CLSPMV_CSR_VEC_GROUP_SIZE) / 1000.0;
    return ((double)newnnz) / 1000000.0 / f.csr_max_flop + over;
}

double Partitioner::estimate_csr_coo_time(const std::vector<int>& rownnz, int nnz, FlopInfo& f) const {
    if (nnz == 0) return 0.0f;
    const int w = nnz_per_row(nnz, (long long)rownnz.size());
    csr_flops(f, (int)rownnz.size(), w);
    coo_flops(f, (int)rownnz.size(), w);
    if (f.coo_max_flop >= f.csr_max_flop) return estimate_coo_time(nnz, f);
    const double coo = estimate_coo_time(nnz, f);
    const double csr = estimate_csr_time(rownnz, nnz, f);
    return coo < csr ? coo : csr;
}

// Shared by estimate_sell_ell_time (:3258) and extract_sell_ell (:3451):
// best ELL width from the row-length histogram and per-slice SELL widths.
struct SellEllPlan {
    double ell_time = 1000000.0;
    int ell_num = 0, ell_nnz = 0;
    int sh = 0;
    std::vector<int> w_full, w_part;
    int part_nnz = 0, full_nnz = 0;
    double sell_full_time = 0, sell_part_time = 0;
};

static SellEllPlan plan_sell_ell(const Model& M, const std::vector<int>& rownnz, int nnz,
                                 const FlopInfo& f) {
    SellEllPlan p;
    double remain_flop;
    int remain_group, remain_thread;
    if (f.csr_max_flop > f.coo_max_flop) {
        remain_flop = f.csr_max_flop;
        remain_group = f.csr_groupnum;
        remain_thread = CLSPMV_CSR_VEC_GROUP_SIZE;
    } else {
        remain_flop = f.coo_max_flop;
        remain_group = f.coo_groupnum;
        remain_thread = CLSPMV_COO_GROUP_SIZE;
    }
    const double remain_over = M.overhead(remain_group, remain_thread) / 1000.0;
    int maxr = 0;
    for (int v : rownnz) maxr = std::max(maxr, v);
    std::vector<int> hist(maxr + 1, 0);
    for (int v : rownnz) hist[v]++;
    const int nrows = (int)rownnz.size();
    int rownum = nrows, ellnnz = 0;
    double ell_over = 0.0f;
    if (f.ell_meth_num == 0)
        ell_over = M.overhead(nrows / CLSPMV_WORK_GROUP_SIZE, CLSPMV_WORK_GROUP_SIZE);
    else if (f.ell_meth_num == 1)
        ell_over = M.overhead(nrows / (4 * CLSPMV_WORK_GROUP_SIZE), CLSPMV_WORK_GROUP_SIZE);
    ell_over /= 1000.0;
    const int newlength = round_up(nrows, CLSPMV_GPU_ALIGNMENT);
    for (size_t i = 0; i < hist.size(); ++i) {
        rownum -= hist[i];
        double t = 0.0f;
        if (i > 0) t += ((double)i * (double)newlength) / 1000000.0 / f.ell_max_flop + ell_over;
        if (ellnnz < nnz) t += ((double)(nnz - ellnnz)) / 1000000.0 / remain_flop + remain_over;
        if (t < p.ell_time) {
            p.ell_time = t;
            p.ell_num = (int)i;
            p.ell_nnz = ellnnz;
        }
        ellnnz += rownum;
    }
    p.sh = f.sell_meth_num == 0 ? CLSPMV_WARPSIZE : f.sell_meth_num == 1 ? CLSPMV_SELL_GROUP_SIZE : 0;
    const int sh = p.sh;
    const int nslice = nrows / sh + (nrows % sh != 0);
    p.w_full.assign(nslice, 0);
    p.w_part.assign(nslice, 0);
    p.full_nnz = nnz;
    for (int i = 0; i < nslice; ++i) {
        std::vector<int> sw;
        for (int j = i * sh; j < (i + 1) * sh && j < nrows; ++j) sw.push_back(rownnz[j]);
        std::sort(sw.begin(), sw.end());
        p.w_full[i] = sw.back();
        int partial = (int)(sw[0] * (long long)sw.size());
        int current = partial;
        p.w_part[i] = sw[0];
        for (size_t j = 1; j < sw.size(); ++j) {
            partial += (int)((long long)(sw[j] - sw[j - 1]) * (long long)(sw.size() - j));
            const double ratio = remain_flop / f.sell_max_flop;
            if (partial > (int)(((double)(sw.size() * (size_t)sw[j])) * ratio)) {
                p.w_part[i] = sw[j];
                current = partial;
            } else {
                break;
            }
        }
        p.part_nnz += current;
    }
    const double sell_over = M.overhead(nslice, CLSPMV_SELL_GROUP_SIZE) / 1000.0;
    int sfull = 0, spart = 0;
    for (int i = 0; i < nslice; ++i) {
        sfull += sh * p.w_full[i];
        spart += sh * p.w_part[i];
    }
    p.sell_full_time = ((double)sfull) / 1000000.0 / f.sell_max_flop + sell_over;
    p.sell_part_time = ((double)spart) / 1000000.0 / f.sell_max_flop + sell_over +
                       ((double)(nnz - p.part_nnz)) / 1000000.0 / remain_flop + remain_over;
    return p;
}

double Partitioner::estimate_sell_ell_time(const std::vector<int>& rownnz, int nnz, const FlopInfo& f,
                                           std::vector<int>& remain, int& nnzremain) const {
    if (nnz == 0) {
        nnzremain = 0;
        std::fill(remain.begin(), remain.end(), 0);
        return 0.0f;
    }
    SellEllPlan p = plan_sell_ell(M, rownnz, nnz, f);
    nnzremain = 0;
    if (p.ell_time < p.sell_full_time && p.ell_time < p.sell_part_time) {
        for (size_t i = 0; i < rownnz.size(); ++i) {
            remain[i] = rownnz[i] <= p.ell_num ? 0 : rownnz[i] - p.ell_num;
            nnzremain += remain[i];
        }
        return p.ell_time;
    }
    if (p.sell_part_time < p.sell_full_time) {
        for (size_t i = 0; i < rownnz.size(); ++i) {
            const int w = p.w_part[i / p.sh];
            remain[i] = rownnz[i] <= w ? 0 : rownnz[i] - w;
            nnzremain += remain[i];
        }
        return p.sell_part_time;
    }
    std::fill(remain.begin(), remain.end(), 0);
    return p.sell_full_time;
}

// estimate_flat_time (:3629)
double Partitioner::estimate_flat_time(const std::vector<int>& rownnz, int nnz) const {
    if (nnz == 0) return 0.0f;
    FlopInfo f;
    flat_flops(f, (int)rownnz.size(), nnz_per_row(nnz, (long long)rownnz.size()));
    if (f.max_flat_base_flop == f.coo_max_flop) return estimate_coo_time(nnz, f);
    if (f.max_flat_base_flop == f.csr_max_flop) return estimate_csr_time(rownnz, nnz, f);
    std::vector<int> remain(rownnz.size(), 0);
    int nnzremain = 0;
    const double t = estimate_sell_ell_time(rownnz, nnz, f, remain, nnzremain);
    return t + estimate_csr_coo_time(remain, nnzremain, f);
}

void Partitioner::extract_coo(const Coo& A) {
    if (A.nnz() == 0) return;
    coo_flops(info, A.height, nnz_per_row(A.nnz(), A.height));
    mat.ifusecoo = true;
    mat.coo_meth_num = info.coo_meth_num;
    mat.coo_group_num = info.coo_groupnum;
    mat.coo.matinfo = {A.width, A.height, A.nnz()};
    mat.coo.coo_row_id = host_copy(A.row);
    mat.coo.coo_col_id = host_copy(A.col);
    mat.coo.coo_data = host_copy(A.val);
    logf("flat: COO with %d nonzeros", A.nnz());
}

void Partitioner::extract_csr(const Coo& A) {
    if (A.nnz() == 0) return;
    csr_flops(info, A.height, nnz_per_row(A.nnz(), A.height));
    mat.ifusecsr = true;
    mat.csr_meth_num = info.csr_meth_num;
    mat.csr_group_num = info.csr_groupnum;
    std::vector<int> rp;
    rowptr_of(A, rp);
    mat.csr.matinfo = {A.width, A.height, A.nnz()};
    mat.csr.csr_row_ptr = host_copy(rp);
    mat.csr.csr_col_id = host_copy(A.col);
    mat.csr.csr_data = host_copy(A.val);
    logf("flat: CSR with %d nonzeros", A.nnz());
}

// extract_csr_coo (:2980): COO when its table wins outright, otherwise the
// faster of COO and CSR with CSR's static load imbalance.
void Partitioner::extract_csr_coo(const Coo& A) {
    if (A.nnz() == 0) return;
    const int w = nnz_per_row(A.nnz(), A.height);
    csr_flops(info, A.height, w);
    coo_flops(info, A.height, w);
    if (info.coo_max_flop >= info.csr_max_flop) {
        extract_coo(A);
        return;
    }
    std::vector<int> cnt(A.height, 0);
    for (int r : A.row) cnt[r]++;
    const double coo = ((double)A.nnz()) / 1000000.0 / info.coo_max_flop +
                       M.overhead(info.coo_groupnum, CLSPMV_COO_GROUP_SIZE) / 1000.0;
    const int newnnz = csr_max_group_nnz(cnt, info.csr_meth_num, info.csr_groupnum) * info.csr_groupnum;
    const double csr = ((double)newnnz) / 1000000.0 / info.csr_max_flop +
                       M.overhead(info.csr_groupnum, CLSPMV_CSR_VEC_GROUP_SIZE) / 1000.0;
    if (coo <= csr)
        extract_coo(A);
    else
        extract_csr(A);
}

// extract_sell (:3078): the first width[i] entries of every row of slice i.
void Partitioner::extract_sell(const Coo& A, Coo& rem, const std::vector<int>& count, int sh,
                               const std::vector<int>& width, int slice_nnz) {
    if (A.nnz() == 0 || slice_nnz == 0) return;
    mat.ifusesell = true;
    mat.sell_meth_num = info.sell_meth_num;
    std::vector<int> rp(count.size() + 1, 0);
    for (size_t i = 1; i < rp.size(); ++i) rp[i] = rp[i - 1] + count[i - 1];
    rem = Coo();
    rem.width = A.width;
    rem.height = A.height;
    rem.allocated = true;
    const int nslice = (int)width.size();
    clspmv_sell_matrix& s = mat.sell;
    s.matinfo = {A.width, A.height, slice_nnz};
    s.sell_slice_height = sh;
    s.sell_slice_num = nslice;
    std::vector<int> ptr(nslice + 1, 0);
    for (int i = 0; i < nslice; ++i) ptr[i + 1] = ptr[i] + width[i] * sh;
    s.sell_slice_ptr = host_copy(ptr);
    s.sell_col_id = host_alloc_zero<int>(ptr[nslice]);
    s.sell_data = host_alloc_zero<float>(ptr[nslice]);
    for (int i = 0; i < nslice; ++i) {
        int* col = s.sell_col_id + ptr[i];
        float* val = s.sell_data + ptr[i];
        for (int j = 0; j < sh; ++j) {
            const int r = i * sh + j;
            const int len = r < A.height ? rp[r + 1] - rp[r] : 0;
            for (int k = 0; k < width[i]; ++k) {
                const int at = k * sh + j;
                if (k < len) {
                    col[at] = A.col[rp[r] + k];
                    val[at] = A.val[rp[r] + k];
                } else {
                    col[at] = at >= sh ? col[at - sh] : col[0];
                }
            }
            for (int k = width[i]; k < len; ++k) {
                rem.row.push_back(r);
                rem.col.push_back(A.col[rp[r] + k]);
                rem.val.push_back(A.val[rp[r] + k]);
            }
        }
    }
    logf("flat: SELL (slice %d) with %d nonzeros", sh, slice_nnz);
}

// extract_ell (:3185): the first ellnum entries of every row.
void Partitioner::extract_ell(const Coo& A, Coo& rem, const std::vector<int>& count, int ellnum,
                              int ellnnz) {
    if (A.nnz() == 0 || ellnum == 0) return;
    ell_flops(info, A.height, ellnum);
    mat.ifuseell = true;
    mat.ell_meth_num = info.ell_meth_num;
    std::vector<int> rp(count.size() + 1, 0);
    for (size_t i = 1; i < rp.size(); ++i) rp[i] = rp[i - 1] + count[i - 1];
    rem = Coo();
    rem.width = A.width;
    rem.height = A.height;
    rem.allocated = true;
    clspmv_ell_matrix& e = mat.ell;
    e.matinfo = {A.width, A.height, ellnnz};
    e.ell_num = ellnum;
    const int hal = round_up(A.height, CLSPMV_GPU_ALIGNMENT);
    e.ell_height_aligned = hal;
    e.ell_col_id = host_alloc_zero<int>((size_t)hal * ellnum);
    e.ell_data = host_alloc_zero<float>((size_t)hal * ellnum);
    for (int r = 0; r < A.height; ++r) {
        const int len = rp[r + 1] - rp[r];
        int last = 0;
        for (int k = 0; k < ellnum; ++k) {
            const size_t at = (size_t)k * hal + r;
            if (k < len) {
                e.ell_col_id[at] = A.col[rp[r] + k];
                e.ell_data[at] = A.val[rp[r] + k];
                last = A.col[rp[r] + k];
            } else {
                e.ell_col_id[at] = last;
            }
        }
        for (int k = ellnum; k < len; ++k) {
            rem.row.push_back(r);
            rem.col.push_back(A.col[rp[r] + k]);
            rem.val.push_back(A.val[rp[r] + k]);
        }
    }
    logf("flat: ELL width %d with %d nonzeros", ellnum, ellnnz);
}

void Partitioner::extract_sell_ell(const Coo& A, Coo& rem) {
    if (A.nnz() == 0) return;
    std::vector<int> cnt(A.height, 0);
    for (int r : A.row) cnt[r]++;
    SellEllPlan p = plan_sell_ell(M, cnt, A.nnz(), info);
    if (p.ell_time < p.sell_full_time && p.ell_time < p.sell_part_time) {
        extract_ell(A, rem, cnt, p.ell_num, p.ell_nnz);
    } else if (p.sell_part_time < p.sell_full_time) {
        extract_sell(A, rem, cnt, p.sh, p.w_part, p.part_nnz);
    } else {
        extract_sell(A, rem, cnt, p.sh, p.w_full, p.full_nnz);
    }
}

// extract_flat (:3675)
void Partitioner::extract_flat(const Coo& A) {
    if (A.nnz() == 0) return;
    update_flat(A);
    if (info.max_flat_base_flop == info.coo_max_flop) {
        extract_coo(A);
        return;
    }
    if (info.max_flat_base_flop == info.csr_max_flop) {
        extract_csr_coo(A);
        return;
    }
    Coo rem;
    extract_sell_ell(A, rem);
    if (!rem.allocated)
        extract_csr_coo(A);
    else
        extract_csr_coo(rem);
}

// analyze_matrix (:3714): order the three families by their base GFLOP/s.
void Partitioner::run(const Coo& A, float dia_mult, bool flat_partial, bool flat_full) {
    mat.mat_width = A.width;
    mat.mat_height = A.height;
    mat.mat_nnz = A.nnz();
    update_dia(A);
    update_block(A);
    update_flat(A);
    double dia_base = info.max_dia_base_flop;
    dia_base *= dia_mult;
    const double block_base = info.max_block_base_flop;
    const double flat_base = info.max_flat_base_flop;
    logf("base GFLOP/s: diagonal %.4f block %.4f flat %.4f", dia_base, block_base, flat_base);
    const bool square = A.width == A.height;
    Coo r1, r2;
    auto block_then_flat = [&](const Coo& in) {
        if (extract_block(in, r2, flat_partial, flat_full))
            extract_flat(r2);
        else
            extract_flat(in);
    };
    if (dia_base > flat_base) {
        if (block_base > flat_base) {
            if (dia_base > block_base) {
                if (square) {
                    if (extract_dia(A, r1))
                        block_then_flat(r1);
                    else
                        block_then_flat(A);
                } else {
                    if (extract_block(A, r1, flat_partial, flat_full))
                        extract_flat(r1);
                    else
                        extract_flat(A);
                }
            } else {
                if (square) {
                    if (extract_block(A, r1, flat_partial, flat_full)) {
                        if (extract_dia(r1, r2))
                            extract_flat(r2);
                        else
                            extract_flat(r1);
                    } else {
                        if (extract_dia(A, r2))
                            extract_flat(r2);
                        else
                            extract_flat(A);
                    }
                } else {
                    if (extract_block(A, r1, flat_partial, flat_full))
                        extract_flat(r1);
                    else
                        extract_flat(A);
                }
            }
        } else {
            if (square) {
                if (extract_dia(A, r1))
                    extract_flat(r1);
                else
                    extract_flat(A);
            } else {
                extract_flat(A);
            }
        }
    } else {
        if (block_base > flat_base) {
            if (extract_block(A, r1, flat_partial, flat_full))
                extract_flat(r1);
            else
                extract_flat(A);
        } else {
            extract_flat(A);
        }
    }
}

}  // namespace
}  // namespace ckt
}  // namespace clspmv

using namespace clspmv;

void clspmv_set_analyze_log(void (*log)(const char* line)) { ckt::g_log = log; }

void clspmv_init_cocktail(clspmv_cocktail* mat) {
    if (mat) std::memset(mat, 0, sizeof(*mat));
}

void clspmv_free_cocktail(clspmv_cocktail* mat) {
    if (!mat) return;
    mat->ifusebdia = mat->ifusedia = mat->ifusesbell = mat->ifusebell = mat->ifusebcsr = false;
    mat->ifusesell = mat->ifuseell = mat->ifusecsr = mat->ifusecoo = false;
    clspmv_free_bdia_matrix(&mat->bdia);
    clspmv_free_dia_matrix(&mat->dia);
    clspmv_free_sbell_matrix(&mat->sbell);
    clspmv_free_b4ell_matrix(&mat->bell);
    clspmv_free_b4csr_matrix(&mat->bcsr);
    clspmv_free_sell_matrix(&mat->sell);
    clspmv_free_ell_matrix(&mat->ell);
    clspmv_free_csr_matrix(&mat->csr);
    clspmv_free_coo_matrix(&mat->coo);
}

int clspmv_analyze_matrix(clspmv_coo_matrix* coomat, clspmv_cocktail* mat,
                          float dia_priority_multiplier, int estimate_flat_partial,
                          int estimate_flat_full, const char* ben_dir) {
    int st = clspmv_validate_coo(coomat);
    if (st) return st;
    if (!mat) return fail(CLSPMV_ERR_INVALID, "null cocktail");
    if (!coo_sorted(*coomat))
        return fail(CLSPMV_ERR_INVALID, "analyze_matrix: the COO matrix must be sorted by (row, col)");
    for (int i = 0; i + 1 < coomat->matinfo.nnz; ++i)
        if (coomat->coo_row_id[i] == coomat->coo_row_id[i + 1] &&
            coomat->coo_col_id[i] == coomat->coo_col_id[i + 1])
            return fail(CLSPMV_ERR_INVALID, "analyze_matrix: duplicate entry (%d,%d)",
                        coomat->coo_row_id[i], coomat->coo_col_id[i]);
    std::string dir;
    if (ben_dir) {
        dir = ben_dir;
    } else {
        const char* env = std::getenv("CLSPMVPATH");
        if (!env) return fail(CLSPMV_ERR_INVALID, "no table directory: pass ben_dir or set CLSPMVPATH");
        dir = env;
    }
    ckt::Tables tables;
    st = ckt::read_tables(dir, tables);
    if (st) return st;
    if (coomat->matinfo.height == 0) return fail(CLSPMV_ERR_INVALID, "analyze_matrix: empty matrix");
    clspmv_init_cocktail(mat);
    ckt::Partitioner p(tables, *mat);
    p.run(ckt::from_c(*coomat), dia_priority_multiplier, estimate_flat_partial != 0,
          estimate_flat_full != 0);
    return CLSPMV_OK;
}
