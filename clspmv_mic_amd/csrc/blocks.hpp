// Block-row gathering shared by every block format (bcsr, b4csr, bell,
// b4ell, sbell) and by the Cocktail block extractors.
//
// For one block row (bheight consecutive rows) the entries are taken in
// (column, row) order and grouped by block column `col - col % bwidth`, left
// to right.  Each block carries a dense bheight x bwidth tile stored row-major
// (index h*bwidth + w).  This is the grouping the reference builds in
// coo2bcsr (include/matrix_storage.h:1654-1712) and in the extract_*_part
// routines of src/analyze.cpp; the entry order inside a tile does not matter
// because keys are unique.
#pragma once

#include <algorithm>
#include <vector>

namespace clspmv {

struct BlockEntry {
    int row;
    int col;
    float val;
};

struct BlockRowScratch {
    std::vector<BlockEntry> entries;
};

// Calls `emit(bcol, tile_entries_begin, tile_entries_end)` for every block of
// block row `brow`, in increasing block-column order.  `rp` is the row
// pointer of the (row-sorted) COO arrays.
template <class Emit>
void for_each_block(const int* rows, const int* cols, const float* vals,
                    const std::vector<int>& rp, int height, int brow,
                    int bheight, int bwidth, BlockRowScratch& scratch,
                    Emit&& emit) {
    int r0 = brow * bheight;
    int r1 = std::min(r0 + bheight, height);
    int start = rp[r0];
    int end = rp[r1];
    if (end <= start) return;
    auto& e = scratch.entries;
    e.resize(end - start);
    for (int i = start; i < end; ++i) e[i - start] = {rows[i], cols[i], vals[i]};
    std::sort(e.begin(), e.end(), [](const BlockEntry& a, const BlockEntry& b) {
        return a.col != b.col ? a.col < b.col : a.row < b.row;
    });
    size_t i = 0;
    while (i < e.size()) {
        int bcol = e[i].col - e[i].col % bwidth;
        size_t j = i + 1;
        while (j < e.size() && e[j].col - e[j].col % bwidth == bcol) ++j;
        emit(bcol, e.data() + i, e.data() + j);
        i = j;
    }
}

// Dense row-major tile of one block.
inline void fill_tile(const BlockEntry* b, const BlockEntry* e, int bheight,
                      int bwidth, float* tile) {
    std::fill(tile, tile + bheight * bwidth, 0.0f);
    for (const BlockEntry* p = b; p != e; ++p)
        tile[(p->row % bheight) * bwidth + (p->col % bwidth)] = p->val;
}

// Block-row CSR of tiles: the common intermediate of all block formats.
struct BlockCSR {
    int bheight = 0, bwidth = 0;
    std::vector<int> rowptr;   // size nbrows+1
    std::vector<int> bcol;     // element column of each block's left edge
    std::vector<float> tiles;  // block i: tiles[i*bh*bw + h*bw + w]
    int nbrows() const { return (int)rowptr.size() - 1; }
    int nblocks() const { return (int)bcol.size(); }
};

void build_block_csr(const int* rows, const int* cols, const float* vals,
                     int nnz, int height, int bheight, int bwidth,
                     BlockCSR& out);

}  // namespace clspmv
