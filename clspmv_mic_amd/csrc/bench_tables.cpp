// Reading and writing the per-format performance tables (.ben files) that
// drive the Cocktail partitioner.  Grammar per file (src/analyze.cpp:16-175,
// writers in src/bench_*.cpp):
//   overhead.ben           groups threads microseconds
//   bdia/dia/sell/ell.ben  dimension width gflops[IMP_NUM]
//   sbell/bell/bcsr.ben    bheight bwidth dimension blocks gflops[IMP_NUM]
//   csr/coo.ben            dimension width groups gflops[IMP_NUM]
// The partitioner's lookups assume an ordering the reference never checks
// (it reads out of bounds on an empty table, src/analyze.cpp:298); here it is
// checked when the tables are read.

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <sys/stat.h>

#include "cocktail.hpp"
#include "common.hpp"

namespace clspmv {
namespace ckt {
namespace {

// Reads whitespace-separated rows of `ni` unsigned ints then `nd` doubles.
int read_rows(const std::string& path, int ni, int nd,
              std::vector<std::vector<unsigned>>& ints,
              std::vector<std::vector<double>>& dbls) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return fail(CLSPMV_ERR_IO, "cannot open %s", path.c_str());
    ints.clear();
    dbls.clear();
    while (true) {
        std::vector<unsigned> iv(ni);
        std::vector<double> dv(nd);
        int got = 0;
        for (int k = 0; k < ni; ++k) {
            int v;
            int r = std::fscanf(f, "%d", &v);
            if (r == EOF && k == 0) {
                std::fclose(f);
                return CLSPMV_OK;
            }
            if (r != 1) {
                std::fclose(f);
                return fail(CLSPMV_ERR_IO, "%s: malformed row %zu", path.c_str(), ints.size() + 1);
            }
            iv[k] = (unsigned)v;
            ++got;
        }
        for (int k = 0; k < nd; ++k) {
            if (std::fscanf(f, "%lf", &dv[k]) != 1) {
                std::fclose(f);
                return fail(CLSPMV_ERR_IO, "%s: malformed row %zu", path.c_str(), ints.size() + 1);
            }
        }
        ints.push_back(iv);
        dbls.push_back(dv);
    }
}

template <class Row, int N>
int load_width(const std::string& dir, const char* name, std::vector<Row>& out) {
    std::vector<std::vector<unsigned>> I;
    std::vector<std::vector<double>> D;
    int st = read_rows(dir + "/benchmark/" + name, 2, N, I, D);
    if (st) return st;
    out.resize(I.size());
    for (size_t i = 0; i < I.size(); ++i) {
        out[i].dim = I[i][0];
        out[i].width = I[i][1];
        for (int k = 0; k < N; ++k) out[i].gf[k] = D[i][k];
    }
    return CLSPMV_OK;
}

template <class Row, int N>
int load_block(const std::string& dir, const char* name, std::vector<Row>& out) {
    std::vector<std::vector<unsigned>> I;
    std::vector<std::vector<double>> D;
    int st = read_rows(dir + "/benchmark/" + name, 4, N, I, D);
    if (st) return st;
    out.resize(I.size());
    for (size_t i = 0; i < I.size(); ++i) {
        out[i].bh = I[i][0];
        out[i].bw = I[i][1];
        out[i].dim = I[i][2];
        out[i].blocks = I[i][3];
        for (int k = 0; k < N; ++k) out[i].gf[k] = D[i][k];
    }
    return CLSPMV_OK;
}

template <class Row, int N>
int load_group(const std::string& dir, const char* name, std::vector<Row>& out) {
    std::vector<std::vector<unsigned>> I;
    std::vector<std::vector<double>> D;
    int st = read_rows(dir + "/benchmark/" + name, 3, N, I, D);
    if (st) return st;
    out.resize(I.size());
    for (size_t i = 0; i < I.size(); ++i) {
        out[i].dim = I[i][0];
        out[i].width = I[i][1];
        out[i].groups = I[i][2];
        for (int k = 0; k < N; ++k) out[i].gf[k] = D[i][k];
    }
    return CLSPMV_OK;
}

// dims non-decreasing, widths strictly increasing inside each dim run.
template <class Row>
int check_grid(const std::vector<Row>& t, const char* name) {
    if (t.empty()) return fail(CLSPMV_ERR_IO, "%s.ben is empty", name);
    for (size_t i = 1; i < t.size(); ++i) {
        if (t[i].dim < t[i - 1].dim)
            return fail(CLSPMV_ERR_IO, "%s.ben: dimensions not sorted at row %zu", name, i + 1);
        if (t[i].dim == t[i - 1].dim && t[i].width <= t[i - 1].width)
            return fail(CLSPMV_ERR_IO, "%s.ben: widths not increasing at row %zu", name, i + 1);
    }
    return CLSPMV_OK;
}

const unsigned kBH[8] = {1, 2, 4, 8, 1, 2, 4, 8};
const unsigned kBW[8] = {4, 4, 4, 4, 8, 8, 8, 8};

// Rows grouped by block shape in the order (1,4),(2,4),(4,4),(8,4),(1,8),
// (2,8),(4,8),(8,8) (src/analyze.cpp:545-560); inside a shape, dims sorted
// and block counts increasing per dim.
template <class Row>
int check_blocks(const std::vector<Row>& t, const char* name) {
    if (t.empty()) return fail(CLSPMV_ERR_IO, "%s.ben is empty", name);
    int shape = 0;
    for (size_t i = 0; i < t.size(); ++i) {
        while (shape < 8 && !(t[i].bh == kBH[shape] && t[i].bw == kBW[shape])) ++shape;
        if (shape == 8)
            return fail(CLSPMV_ERR_IO, "%s.ben: block shape %ux%u out of order at row %zu", name,
                        t[i].bh, t[i].bw, i + 1);
        if (i > 0 && t[i].bh == t[i - 1].bh && t[i].bw == t[i - 1].bw) {
            if (t[i].dim < t[i - 1].dim ||
                (t[i].dim == t[i - 1].dim && t[i].blocks <= t[i - 1].blocks))
                return fail(CLSPMV_ERR_IO, "%s.ben: rows not sorted at row %zu", name, i + 1);
        }
    }
    // Every shape present.
    for (int s = 0; s < 8; ++s) {
        bool found = false;
        for (const Row& r : t) found |= (r.bh == kBH[s] && r.bw == kBW[s]);
        if (!found) return fail(CLSPMV_ERR_IO, "%s.ben: block shape %ux%u missing", name, kBH[s], kBW[s]);
    }
    return CLSPMV_OK;
}

// (dim, width) groups of equal length; dims and widths sorted.
template <class Row>
int check_groups(const std::vector<Row>& t, const char* name) {
    if (t.empty()) return fail(CLSPMV_ERR_IO, "%s.ben is empty", name);
    size_t first_len = 0, run = 1;
    for (size_t i = 1; i <= t.size(); ++i) {
        bool same = i < t.size() && t[i].dim == t[i - 1].dim && t[i].width == t[i - 1].width;
        if (same) {
            ++run;
            continue;
        }
        if (first_len == 0) first_len = run;
        if (run != first_len)
            return fail(CLSPMV_ERR_IO, "%s.ben: group sizes differ (%zu vs %zu)", name, run, first_len);
        if (i < t.size() && (t[i].dim < t[i - 1].dim ||
                             (t[i].dim == t[i - 1].dim && t[i].width < t[i - 1].width)))
            return fail(CLSPMV_ERR_IO, "%s.ben: rows not sorted at row %zu", name, i + 1);
        run = 1;
    }
    return CLSPMV_OK;
}

template <class Row, int N>
int write_width(FILE* f, const std::vector<Row>& t) {
    for (const Row& r : t) {
        std::fprintf(f, "%d %d", (int)r.dim, (int)r.width);
        for (int k = 0; k < N; ++k) std::fprintf(f, " %f", r.gf[k]);
        std::fprintf(f, "\n");
    }
    return CLSPMV_OK;
}

}  // namespace

int read_tables(const std::string& dir, Tables& t) {
    std::vector<std::vector<unsigned>> I;
    std::vector<std::vector<double>> D;
    int st = read_rows(dir + "/benchmark/overhead.ben", 2, 1, I, D);
    if (st) return st;
    t.overhead.resize(I.size());
    for (size_t i = 0; i < I.size(); ++i) t.overhead[i] = {I[i][0], I[i][1], D[i][0]};
    if (t.overhead.empty()) return fail(CLSPMV_ERR_IO, "overhead.ben is empty");
    for (size_t i = 1; i < t.overhead.size(); ++i) {
        const OverheadRow &a = t.overhead[i - 1], &b = t.overhead[i];
        if (b.groups < a.groups || (b.groups == a.groups && b.threads <= a.threads))
            return fail(CLSPMV_ERR_IO, "overhead.ben: rows not sorted at row %zu", i + 1);
    }
    if ((st = load_width<WidthRow<CLSPMV_BDIA_IMP_NUM>, CLSPMV_BDIA_IMP_NUM>(dir, "bdia.ben", t.bdia))) return st;
    if ((st = load_width<WidthRow<CLSPMV_DIA_IMP_NUM>, CLSPMV_DIA_IMP_NUM>(dir, "dia.ben", t.dia))) return st;
    if ((st = load_block<BlockRow<CLSPMV_SBELL_IMP_NUM>, CLSPMV_SBELL_IMP_NUM>(dir, "sbell.ben", t.sbell))) return st;
    if ((st = load_block<BlockRow<CLSPMV_BELL_IMP_NUM>, CLSPMV_BELL_IMP_NUM>(dir, "bell.ben", t.bell))) return st;
    if ((st = load_block<BlockRow<CLSPMV_BCSR_IMP_NUM>, CLSPMV_BCSR_IMP_NUM>(dir, "bcsr.ben", t.bcsr))) return st;
    if ((st = load_width<WidthRow<CLSPMV_SELL_IMP_NUM>, CLSPMV_SELL_IMP_NUM>(dir, "sell.ben", t.sell))) return st;
    if ((st = load_width<WidthRow<CLSPMV_ELL_IMP_NUM>, CLSPMV_ELL_IMP_NUM>(dir, "ell.ben", t.ell))) return st;
    if ((st = load_group<GroupRow<CLSPMV_CSR_IMP_NUM>, CLSPMV_CSR_IMP_NUM>(dir, "csr.ben", t.csr))) return st;
    if ((st = load_group<GroupRow<CLSPMV_COO_IMP_NUM>, CLSPMV_COO_IMP_NUM>(dir, "coo.ben", t.coo))) return st;
    if ((st = check_grid(t.bdia, "bdia"))) return st;
    if ((st = check_grid(t.dia, "dia"))) return st;
    if ((st = check_grid(t.sell, "sell"))) return st;
    if ((st = check_grid(t.ell, "ell"))) return st;
    if ((st = check_blocks(t.sbell, "sbell"))) return st;
    if ((st = check_blocks(t.bell, "bell"))) return st;
    if ((st = check_blocks(t.bcsr, "bcsr"))) return st;
    if ((st = check_groups(t.csr, "csr"))) return st;
    if ((st = check_groups(t.coo, "coo"))) return st;
    // find_min_bdia_width indexes dia.ben with bdia.ben's row numbers
    // (src/analyze.cpp:893): the two grids must coincide.
    if (t.dia.size() != t.bdia.size())
        return fail(CLSPMV_ERR_IO, "bdia.ben and dia.ben have different grids");
    for (size_t i = 0; i < t.dia.size(); ++i)
        if (t.dia[i].dim != t.bdia[i].dim || t.dia[i].width != t.bdia[i].width)
            return fail(CLSPMV_ERR_IO, "bdia.ben and dia.ben differ at row %zu", i + 1);
    return CLSPMV_OK;
}

int write_tables(const std::string& dir, const Tables& t) {
    std::string bdir = dir + "/benchmark";
    mkdir(dir.c_str(), 0755);
    mkdir(bdir.c_str(), 0755);
    auto open = [&](const char* name) { return std::fopen((bdir + "/" + name).c_str(), "w"); };
    FILE* f = open("overhead.ben");
    if (!f) return fail(CLSPMV_ERR_IO, "cannot write %s/overhead.ben", bdir.c_str());
    for (const OverheadRow& r : t.overhead)
        std::fprintf(f, "%d %d %f\n", (int)r.groups, (int)r.threads, r.us);
    std::fclose(f);
#define WRITE_WIDTH(name, vec, N)                                             \
    if (!(f = open(name))) return fail(CLSPMV_ERR_IO, "cannot write %s", name); \
    write_width<std::decay_t<decltype(vec[0])>, N>(f, vec);                   \
    std::fclose(f);
    WRITE_WIDTH("bdia.ben", t.bdia, CLSPMV_BDIA_IMP_NUM)
    WRITE_WIDTH("dia.ben", t.dia, CLSPMV_DIA_IMP_NUM)
    WRITE_WIDTH("sell.ben", t.sell, CLSPMV_SELL_IMP_NUM)
    WRITE_WIDTH("ell.ben", t.ell, CLSPMV_ELL_IMP_NUM)
#undef WRITE_WIDTH
    auto wblock = [&](const char* name, const auto& vec, int n) {
        FILE* g = open(name);
        if (!g) return fail(CLSPMV_ERR_IO, "cannot write %s", name);
        for (const auto& r : vec) {
            std::fprintf(g, "%d %d %d %d", (int)r.bh, (int)r.bw, (int)r.dim, (int)r.blocks);
            for (int k = 0; k < n; ++k) std::fprintf(g, " %f", r.gf[k]);
            std::fprintf(g, "\n");
        }
        std::fclose(g);
        return CLSPMV_OK;
    };
    auto wgroup = [&](const char* name, const auto& vec, int n) {
        FILE* g = open(name);
        if (!g) return fail(CLSPMV_ERR_IO, "cannot write %s", name);
        for (const auto& r : vec) {
            std::fprintf(g, "%d %d %d", (int)r.dim, (int)r.width, (int)r.groups);
            for (int k = 0; k < n; ++k) std::fprintf(g, " %f", r.gf[k]);
            std::fprintf(g, "\n");
        }
        std::fclose(g);
        return CLSPMV_OK;
    };
    int st;
    if ((st = wblock("sbell.ben", t.sbell, CLSPMV_SBELL_IMP_NUM))) return st;
    if ((st = wblock("bell.ben", t.bell, CLSPMV_BELL_IMP_NUM))) return st;
    if ((st = wblock("bcsr.ben", t.bcsr, CLSPMV_BCSR_IMP_NUM))) return st;
    if ((st = wgroup("csr.ben", t.csr, CLSPMV_CSR_IMP_NUM))) return st;
    if ((st = wgroup("coo.ben", t.coo, CLSPMV_COO_IMP_NUM))) return st;
    return CLSPMV_OK;
}

}  // namespace ckt
}  // namespace clspmv
