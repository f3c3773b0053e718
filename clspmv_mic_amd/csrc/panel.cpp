// Host side of the x-panel encoding (k_panel.hip): regroups row-sorted
// entries by column panel, marks segment starts, numbers the segments and
// builds the per-row segment lists and the work-group table.

#include <algorithm>
#include <climits>

#include "kernels.hpp"
#include "runtime.hpp"

namespace clspmv {

int auto_panel_width(int cols, long long nnz) {
    if (cols <= 0 || nnz < 65536) return 0;
    const int pw = std::min(launch::kPanelMax, round_up(cols, 8));
    const long long panels = (cols + pw - 1) / pw;
    // Each panel's x is staged once per work-group; below about two
    // nonzeros per staged float the staging costs more than it saves.
    if (nnz < 2 * panels * pw) return 0;
    return pw;
}

int build_panel(clspmv_dmat* d, int rows, int cols, long long nnz, const int* row,
                const int* col, const float* val, int pw) {
    constexpr int kSub = launch::kPanelSub;
    if (pw <= 0 || pw % 8 || pw > launch::kPanelMax)
        return fail(CLSPMV_ERR_INVALID, "panel width %d: need a multiple of 8 in [8, %d]", pw,
                    launch::kPanelMax);
    if (cols <= 0 || rows <= 0 || nnz <= 0) return CLSPMV_OK;
    const int P = (cols + pw - 1) / pw;
    std::vector<long long> cnt(P, 0);
    for (long long i = 0; i < nnz; ++i) cnt[col[i] / pw]++;
    std::vector<long long> start(P + 1, 0);
    for (int p = 0; p < P; ++p) start[p + 1] = start[p] + (cnt[p] + kSub - 1) / kSub * kSub;
    const long long np = start[P];
    if (np >= INT_MAX)
        return fail(CLSPMV_ERR_TOO_LARGE, "panel encoding of %lld entries exceeds 32-bit indices", np);

    // Stable regrouping by panel keeps every panel row-sorted.  Padding
    // entries (row -1) fill each panel to whole sub-chunks.
    std::vector<unsigned short> pc(np, 0);
    std::vector<float> pv(np, 0.0f);
    std::vector<int> prow(np, -1);
    std::vector<long long> cur(start.begin(), start.end() - 1);
    for (long long i = 0; i < nnz; ++i) {
        const int p = col[i] / pw;
        const long long k = cur[p]++;
        pc[k] = (unsigned short)(col[i] - p * pw);
        pv[k] = val[i];
        prow[k] = row[i];
    }

    // Segments: a new one at every sub-chunk start and at every row change.
    const long long nsc = np / kSub;
    std::vector<int> sc_seg(nsc);
    std::vector<int> seg_row;
    seg_row.reserve(nsc + (size_t)rows);
    for (long long k = 0; k < np; ++k) {
        bool flag;
        if (k % kSub == 0) {
            sc_seg[k / kSub] = (int)seg_row.size();
            flag = true;
        } else {
            flag = prow[k] >= 0 && prow[k] != prow[k - 1];
        }
        if (flag) {
            seg_row.push_back(prow[k]);
            pc[k] |= 0x8000u;
        }
    }
    const long long nseg = (long long)seg_row.size();

    // Per-row segment lists, rows ascending, segments in panel order.
    std::vector<int> per_row(rows, 0);
    for (int r : seg_row)
        if (r >= 0) per_row[r]++;
    std::vector<int> rrow, rs_ptr(1, 0), slot(rows, -1);
    for (int r = 0; r < rows; ++r)
        if (per_row[r]) {
            slot[r] = rs_ptr.back();
            rrow.push_back(r);
            rs_ptr.push_back(rs_ptr.back() + per_row[r]);
        }
    std::vector<int> rs_idx(rs_ptr.back());
    for (long long s = 0; s < nseg; ++s)
        if (seg_row[s] >= 0) rs_idx[slot[seg_row[s]]++] = (int)s;

    // Work-groups: runs of at most G sub-chunks of one panel.
    const int G = std::max(32, std::min(128, pw / kSub));
    std::vector<int> wg;
    for (int p = 0; p < P; ++p) {
        const int b0 = (int)(start[p] / kSub), b1 = (int)(start[p + 1] / kSub);
        const int real_end = (int)(start[p] + cnt[p]);
        for (int b = b0; b < b1; b += G) {
            wg.push_back(p);
            wg.push_back(b);
            wg.push_back(std::min(b + G, b1));
            wg.push_back(real_end);
        }
    }

    d->pw = pw;
    d->pnwg = (int)(wg.size() / 4);
    d->prows = (int)rrow.size();
    int st = upload_array(d, wg.data(), wg.size(), &d->pwg);
    if (!st) st = upload_array(d, pc.data(), pc.size(), &d->pcol);
    if (!st) st = upload_array(d, pv.data(), pv.size(), &d->pval);
    if (!st) st = upload_array(d, sc_seg.data(), sc_seg.size(), &d->psc_seg);
    if (!st) st = upload_array<float>(d, nullptr, (size_t)nseg, &d->pseg);
    if (!st) st = upload_array(d, rrow.data(), rrow.size(), &d->prrow);
    if (!st) st = upload_array(d, rs_ptr.data(), rs_ptr.size(), &d->prs_ptr);
    if (!st) st = upload_array(d, rs_idx.data(), rs_idx.size(), &d->prs_idx);
    if (st) return st;
    d->use_panel = true;
    return CLSPMV_OK;
}

void drop_encoding(clspmv_dmat* d, bool panel) {
    if (panel) {
        if (!d->plain) return;  // the only encoding stays
        release_array(d, &d->pwg);
        release_array(d, &d->pcol);
        release_array(d, &d->pval);
        release_array(d, &d->psc_seg);
        release_array(d, &d->pseg);
        release_array(d, &d->prrow);
        release_array(d, &d->prs_ptr);
        release_array(d, &d->prs_idx);
        d->pw = 0;
        d->use_panel = false;
        return;
    }
    if (!d->pw) return;
    release_array(d, &d->row);
    release_array(d, &d->rp);
    release_array(d, &d->col);
    release_array(d, &d->val);
    release_array(d, &d->carry_row);
    release_array(d, &d->carry_val);
    release_array(d, &d->blocks);
    d->plain = false;
    d->use_panel = true;
}

}  // namespace clspmv
