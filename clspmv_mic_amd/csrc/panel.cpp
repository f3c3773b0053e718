// Host side of the x-panel encoding (k_panel.hip): regroups row-sorted
// entries by column panel, marks segment starts, numbers the segments and
// builds the row-block tables and the work-group table.

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

int panel_group_size(const std::vector<int>& sc, int slots, int pw) {
    // Staging a panel costs about as much as this many sub-chunks (pw floats
    // from L2 against 6 bytes of encoded entry per nonzero from HBM).
    const long long stage = std::max(4, pw / (4 * launch::kPanelSub));
    long long best = -1;
    int bw = launch::kPanelScPerWg;
    for (int W = 8; W <= launch::kPanelScPerWg; W += 8) {
        long long n = 0;
        for (int c : sc) n += (c + W - 1) / W;
        const long long rounds = (n + slots - 1) / std::max(slots, 1);
        const long long cost = rounds * (W + stage);
        if (best < 0 || cost < best) {
            best = cost;
            bw = W;
        }
    }
    return bw;
}

int build_panel(clspmv_dmat* d, int rows, int cols, long long nnz, const int* row,
                const int* col, const float* val, int pw, int dtype, int group) {
    constexpr int kSub = launch::kPanelSub;
    if (pw <= 0 || pw % 8 || pw > launch::kPanelMax)
        return fail(CLSPMV_ERR_INVALID, "panel width %d: need a multiple of 8 in [8, %d]", pw,
                    launch::kPanelMax);
    if (cols <= 0 || rows <= 0 || nnz <= 0) return CLSPMV_OK;
    const int P = (cols + pw - 1) / pw;
    if (P > launch::kPanelsMax)
        return fail(CLSPMV_ERR_UNSUPPORTED, "%d panels of width %d: at most %d panels", P, pw,
                    launch::kPanelsMax);

    // Compute units of the device: both passes size their work-groups to
    // fill them in whole rounds.
    int ncu = 256;
    {
        hipDeviceProp_t prop{};
        if (hipGetDeviceProperties(&prop, d->device) == hipSuccess && prop.multiProcessorCount > 0)
            ncu = prop.multiProcessorCount;
    }

    // Row blocks for pass 2: at most kBlockRowsMax rows and, bounding each
    // row's segments by its per-panel runs (a run of L > 1 entries becomes
    // at most L/kSub + 2 segments), at most `cap` segments (a longer row
    // alone, up to kBlockSegsMax).  Two pass-2 work-groups run per CU; cap
    // splits the segments evenly over whole rounds of them (for one round,
    // the smallest cap whose blocks fit it).
    if (rows >= (1 << 30)) return fail(CLSPMV_ERR_UNSUPPORTED, "panel encoding needs rows < 2^30");
    std::vector<int> rbound(rows);
    long long total_bound = 0;
    {
        long long i = 0;
        for (int r = 0; r < rows; ++r) {
            long long bound = 0;
            while (i < nnz && row[i] == r) {
                const int p = col[i] / pw;
                long long run = 0;
                while (i < nnz && row[i] == r && col[i] / pw == p) {
                    ++run;
                    ++i;
                }
                bound += run == 1 ? 1 : run / kSub + 2;
            }
            if (bound > launch::kBlockSegsMax)
                return fail(CLSPMV_ERR_UNSUPPORTED, "row %d is too long for the panel encoding", r);
            rbound[r] = (int)bound;
            total_bound += bound;
        }
        if (i != nnz) return fail(CLSPMV_ERR_INVALID, "panel encoding needs row-sorted entries");
    }
    // Greedy cut of the rows into blocks under a segment cap.
    auto cut_blocks = [&](long long cap, std::vector<int>* out) {
        long long bound_sum = 0, nb = 1;
        int r0 = 0;
        if (out) out->assign(1, 0);
        for (int r = 0; r < rows; ++r) {
            if (r > r0 && (r - r0 == launch::kBlockRowsMax || bound_sum + rbound[r] > cap)) {
                if (out) out->push_back(r);
                ++nb;
                r0 = r;
                bound_sum = 0;
            }
            bound_sum += rbound[r];
        }
        if (out) out->push_back(rows);
        return nb;
    };
    long long cap = launch::kBlockSegsMax;
    {
        const long long slots = 2ll * ncu;
        const long long rounds = std::max(1ll, (total_bound + slots * cap - 1) / (slots * cap));
        if (rounds == 1 && cut_blocks(cap, nullptr) <= slots) {
            // One round: the smallest cap (not under 1024) whose blocks still
            // fit it, so the blocks -- and the two per CU -- come out even.
            long long lo = 1024, hi = cap;
            while (lo < hi) {
                const long long mid = (lo + hi) / 2;
                if (cut_blocks(mid, nullptr) <= slots) hi = mid;
                else lo = mid + 1;
            }
            cap = std::min<long long>(cap, lo);
        } else {
            // 10 % headroom: greedy packing leaves blocks short of the cap.
            cap = std::min<long long>(cap, std::max(1024ll, total_bound * 11 / (10 * rounds * slots) + 1));
        }
    }
    std::vector<int> blk_row;
    cut_blocks(cap, &blk_row);
    const int nb = (int)blk_row.size() - 1;
    if ((long long)P * (nb + 1) > (1ll << 26))
        return fail(CLSPMV_ERR_UNSUPPORTED, "%d panels x %d row blocks: too many groups", P, nb);
    std::vector<int> blk_of(rows);
    for (int b = 0; b < nb; ++b)
        std::fill(blk_of.begin() + blk_row[b], blk_of.begin() + blk_row[b + 1], b);

    // Groups (panel, block) in pass-1 order, each padded to a multiple of
    // kPer entries (one lane's load) and cut into sub-chunks of kSub entries,
    // the last one shorter.
    const size_t G = (size_t)P * nb;
    std::vector<long long> cnt(G, 0);
    for (long long i = 0; i < nnz; ++i) cnt[(size_t)(col[i] / pw) * nb + blk_of[row[i]]]++;
    std::vector<long long> gstart(G + 1, 0);
    constexpr int kPer = kSub / 64;  // entries per pass-1 lane
    for (size_t g = 0; g < G; ++g) gstart[g + 1] = gstart[g] + (cnt[g] + kPer - 1) / kPer * kPer;
    const long long np = gstart[G];
    if (np + kPer > INT_MAX)
        return fail(CLSPMV_ERR_TOO_LARGE, "panel encoding of %lld entries exceeds 32-bit indices", np);
    std::vector<int> sc_start;   // first entry of each sub-chunk (+ end)
    std::vector<int> g_sc(G + 1);  // first sub-chunk of each group
    for (size_t g = 0; g < G; ++g) {
        g_sc[g] = (int)sc_start.size();
        for (long long k = gstart[g]; k < gstart[g + 1]; k += kSub) sc_start.push_back((int)k);
    }
    g_sc[G] = (int)sc_start.size();
    sc_start.push_back((int)np);
    const long long nsc = (long long)sc_start.size() - 1;

    // Stable scatter keeps each group row-sorted; padding: column code
    // kPanelPad, value 0, row -1.
    std::vector<unsigned short> pc(np, (unsigned short)launch::kPanelPad);
    std::vector<float> pv(np, 0.0f);
    std::vector<int> prow(np, -1);
    {
        std::vector<long long> cur(gstart.begin(), gstart.end() - 1);
        for (long long i = 0; i < nnz; ++i) {
            const int p = col[i] / pw;
            const long long k = cur[(size_t)p * nb + blk_of[row[i]]]++;
            pc[k] = (unsigned short)(col[i] - p * pw);
            pv[k] = val[i];
            prow[k] = row[i];
        }
    }

    // Segments: one at every sub-chunk start and at every row change.
    // Padding (under kPer entries, at a group's end) never starts a sub-chunk,
    // so every segment has a row.  Count per group first.
    std::vector<int> gseg(G, 0);
    for (size_t g = 0; g < G; ++g)
        for (long long k = gstart[g]; k < gstart[g + 1]; ++k) {
            const bool flag = ((k - gstart[g]) % kSub == 0) || (prow[k] >= 0 && prow[k] != prow[k - 1]);
            if (flag) {
                pc[k] |= 0x8000u;
                gseg[g]++;
            }
        }
    // Numbering in (block, panel) order: the first id of each group.
    std::vector<int> blk_tab((size_t)nb * (P + 1));
    std::vector<long long> gbase(G);
    long long nseg = 0;
    for (int b = 0; b < nb; ++b) {
        for (int p = 0; p < P; ++p) {
            blk_tab[(size_t)b * (P + 1) + p] = (int)nseg;
            gbase[(size_t)p * nb + b] = nseg;
            nseg += gseg[(size_t)p * nb + b];
        }
        blk_tab[(size_t)b * (P + 1) + P] = (int)nseg;
    }
    if (nseg >= INT_MAX) return fail(CLSPMV_ERR_TOO_LARGE, "too many segments");
    std::vector<int> sc_seg(nsc);
    std::vector<int> seg_row(nseg);
    for (size_t g = 0; g < G; ++g) {
        long long id = gbase[g] - 1;
        int sc = g_sc[g];
        for (long long k = gstart[g]; k < gstart[g + 1]; ++k) {
            if (sc < g_sc[g + 1] && k == sc_start[sc]) sc_seg[sc++] = (int)(id + 1);
            if (pc[k] & 0x8000u) {
                ++id;
                // Bit 31: the segment continues its predecessor's row (a run
                // cut at a sub-chunk boundary); bit 30 on the predecessor:
                // a continuation follows.
                const bool cont = k > gstart[g] && prow[k] == prow[k - 1];
                seg_row[id] = prow[k] | (cont ? (int)0x80000000u : 0);
                if (cont) seg_row[id - 1] |= 0x40000000;
            }
        }
    }

    // Pass 2 gathers each block's segments row by row: segment id -> slot
    // in the block's row-major layout (rows ascending, each row's segments
    // in id order, i.e. panel order with continuations after their head),
    // and per block the first slot of each row.
    std::vector<unsigned short> seg_dst(nseg);
    std::vector<int> blk_seg(nb + 1);
    std::vector<int> row_slot((size_t)rows + nb);
    for (int b = 0; b < nb; ++b) {
        const int j0 = blk_tab[(size_t)b * (P + 1)], j1 = blk_tab[(size_t)b * (P + 1) + P];
        const int r0 = blk_row[b], nr = blk_row[b + 1] - r0;
        blk_seg[b] = j0;
        int* slot = &row_slot[(size_t)r0 + b];  // nr + 1 entries
        std::fill(slot, slot + nr + 1, 0);
        for (int j = j0; j < j1; ++j) slot[(seg_row[j] & 0x3FFFFFFF) - r0 + 1]++;
        for (int i = 0; i < nr; ++i) slot[i + 1] += slot[i];
        std::vector<int> cur(slot, slot + nr);
        for (int j = j0; j < j1; ++j) seg_dst[j] = (unsigned short)cur[(seg_row[j] & 0x3FFFFFFF) - r0]++;
    }
    blk_seg[nb] = (int)nseg;
    // Rows with more than kLongSlots slots, per block: pass 2 spreads them
    // over its threads and sums them in FP64 ({row in block, first slot, end
    // slot, 0}).
    std::vector<int> blk_long(nb + 1, 0);
    std::vector<int> longs;
    for (int b = 0; b < nb; ++b) {
        const int r0 = blk_row[b], nr = blk_row[b + 1] - r0;
        const int* slot = &row_slot[(size_t)r0 + b];
        blk_long[b] = (int)(longs.size() / 4);
        for (int i = 0; i < nr; ++i)
            if (slot[i + 1] - slot[i] > launch::kLongSlots) {
                longs.push_back(i);
                longs.push_back(slot[i]);
                longs.push_back(slot[i + 1]);
                longs.push_back(0);
            }
    }
    blk_long[nb] = (int)(longs.size() / 4);

    // What the kernels rely on: every sub-chunk opens a segment, its
    // segments' slots exist, and its lanes read whole kPer-entry groups; each
    // block's segments fit one pass-2 work-group and belong to its rows.
    for (long long c = 0; c < nsc; ++c) {
        const int k0 = sc_start[c], k1 = sc_start[c + 1];
        if (!(pc[k0] & 0x8000u) || k1 - k0 > kSub || (k1 - k0) % kPer)
            return fail(CLSPMV_ERR_INVALID, "panel encoding: malformed sub-chunk %lld", c);
        int nf = 0;
        for (int k = k0; k < k1; ++k) nf += (pc[k] >> 15) & 1;
        if ((long long)sc_seg[c] + nf > nseg)
            return fail(CLSPMV_ERR_INVALID, "panel encoding: sub-chunk %lld writes past the segments", c);
    }
    for (int b = 0; b < nb; ++b) {
        const int j0 = blk_tab[(size_t)b * (P + 1)], j1 = blk_tab[(size_t)b * (P + 1) + P];
        if (j1 - j0 > launch::kBlockSegsMax || blk_row[b + 1] - blk_row[b] > launch::kBlockRowsMax)
            return fail(CLSPMV_ERR_INVALID, "panel encoding: block %d over its limits", b);
        for (int j = j0; j < j1; ++j) {
            const int r = seg_row[j] & 0x3FFFFFFF;
            if (r < blk_row[b] || r >= blk_row[b + 1])
                return fail(CLSPMV_ERR_INVALID, "panel encoding: segment %d outside block %d", j, b);
        }
    }

    // Pass-1 work-groups: runs of at most W sub-chunks of one panel.  Each
    // holds its panel in LDS, so few fit on a CU at once: W is chosen so
    // that the work-groups fill the CUs in whole rounds.
    int W = group;
    if (W <= 0) {
        const long long lds = 4ll * pw + 8ll * (launch::kPanelScPerWg + 1);
        const int per_cu = (int)std::max(1ll, std::min(2ll, (160ll << 10) / lds));
        std::vector<int> scp(P);
        for (int p = 0; p < P; ++p) scp[p] = g_sc[(size_t)(p + 1) * nb] - g_sc[(size_t)p * nb];
        W = panel_group_size(scp, ncu * per_cu, pw);
    }
    std::vector<int> wg;
    for (int p = 0; p < P; ++p) {
        const int b0 = g_sc[(size_t)p * nb], b1 = g_sc[(size_t)(p + 1) * nb];
        for (int b = b0; b < b1; b += W) {
            wg.push_back(p);
            wg.push_back(b);
            wg.push_back(std::min(b + W, b1));
            wg.push_back(0);
        }
    }

    {
        const hipError_t e = launch::panel_prepare();
        if (e != hipSuccess)
            return fail(CLSPMV_ERR_HIP, "panel kernel LDS limit: %s", hipGetErrorString(e));
    }
    d->pw = pw;
    d->pnwg = (int)(wg.size() / 4);
    d->pnsc = g_sc[G];
    d->pnp = P;
    d->pnb = nb;
    // One lane's worth of padding after the last group: pass 1 points the
    // loads of lanes with no entry at it.
    pc.insert(pc.end(), kPer, (unsigned short)launch::kPanelPad);
    pv.insert(pv.end(), kPer, 0.0f);
    d->ppad = (int)np;
    int st = upload_array(d, wg.data(), wg.size(), &d->pwg);
    if (!st) st = upload_array(d, pc.data(), pc.size(), &d->pcol);
    if (!st && dtype == CLSPMV_DT_F32) {
        st = upload_array(d, pv.data(), pv.size(), &d->pval);
    } else if (!st) {
        std::vector<unsigned short> lp(pv.size());
        st = clspmv_convert_from_f32(pv.data(), lp.data(), lp.size(), dtype);
        if (!st) st = upload_array(d, lp.data(), lp.size(), &d->plpval);
    }
    if (!st) st = upload_array(d, sc_seg.data(), sc_seg.size(), &d->psc_seg);
    if (!st) st = upload_array(d, sc_start.data(), sc_start.size(), &d->psc_start);
    if (!st) st = upload_array<float>(d, nullptr, (size_t)nseg, &d->pseg);
    if (!st) st = upload_array(d, seg_dst.data(), seg_dst.size(), &d->pseg_dst);
    if (!st) st = upload_array(d, blk_row.data(), blk_row.size(), &d->pblk_row);
    if (!st) st = upload_array(d, blk_seg.data(), blk_seg.size(), &d->pblk_seg);
    if (!st) st = upload_array(d, row_slot.data(), row_slot.size(), &d->prow_slot);
    if (!st) st = upload_array(d, blk_long.data(), blk_long.size(), &d->pblk_long);
    if (!st) st = upload_array(d, longs.data(), longs.size(), &d->plong);
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
        release_array(d, &d->plpval);
        release_array(d, &d->psc_seg);
        release_array(d, &d->psc_start);
        release_array(d, &d->pseg);
        release_array(d, &d->pseg_dst);
        release_array(d, &d->pblk_row);
        release_array(d, &d->pblk_seg);
        release_array(d, &d->prow_slot);
        release_array(d, &d->pblk_long);
        release_array(d, &d->plong);
        d->pw = 0;
        d->use_panel = false;
        return;
    }
    if (!d->pw) return;
    release_array(d, &d->row);
    release_array(d, &d->rp);
    release_array(d, &d->col);
    release_array(d, &d->val);
    release_array(d, &d->lpval);
    release_array(d, &d->carry_row);
    release_array(d, &d->carry_val);
    release_array(d, &d->blocks);
    d->plain = false;
    d->use_panel = true;
}

}  // namespace clspmv
