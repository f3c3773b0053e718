// Internal interface of the Cocktail partitioner: the per-format GFLOP/s
// tables (.ben files) and their interpolation model.
#pragma once

#include <string>
#include <vector>

#include "clspmv_cocktail.h"

namespace clspmv {
namespace ckt {

// One row of each .ben grammar (src/analyze.cpp:16-175).
struct OverheadRow { unsigned groups, threads; double us; };            // overhead.ben
template <int N> struct WidthRow { unsigned dim, width; double gf[N]; };  // bdia/dia/sell/ell
template <int N> struct BlockRow { unsigned bh, bw, dim, blocks; double gf[N]; };  // sbell/bell/bcsr
template <int N> struct GroupRow { unsigned dim, width, groups; double gf[N]; };   // csr/coo

struct Tables {
    std::vector<OverheadRow> overhead;
    std::vector<WidthRow<CLSPMV_BDIA_IMP_NUM>> bdia;
    std::vector<WidthRow<CLSPMV_DIA_IMP_NUM>> dia;
    std::vector<BlockRow<CLSPMV_SBELL_IMP_NUM>> sbell;
    std::vector<BlockRow<CLSPMV_BELL_IMP_NUM>> bell;
    std::vector<BlockRow<CLSPMV_BCSR_IMP_NUM>> bcsr;
    std::vector<WidthRow<CLSPMV_SELL_IMP_NUM>> sell;
    std::vector<WidthRow<CLSPMV_ELL_IMP_NUM>> ell;
    std::vector<GroupRow<CLSPMV_CSR_IMP_NUM>> csr;
    std::vector<GroupRow<CLSPMV_COO_IMP_NUM>> coo;
};

// Reads <dir>/benchmark/<fmt>.ben for the ten tables.  Checks the ordering
// the model relies on; returns a status code.
int read_tables(const std::string& dir, Tables& t);
// Writes the ten tables in the same grammar (used by the autotuner).
int write_tables(const std::string& dir, const Tables& t);

}  // namespace ckt
}  // namespace clspmv
