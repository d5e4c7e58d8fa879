// Level schedule builder (host).
//
// The reference walks cblks in order and lets every source lock its target
// while it adds an update (kernels/core_zgemmsp.c:455-463).  Here the update
// of a level is reorganised by target: every 64x64 tile of a target panel
// becomes one work item that owns the tile and lists, in source order, the
// (source rows, source columns) ranges that land in it.
#include "plan.h"

#include <stdexcept>

#include <algorithm>
#include <cstdio>
#include <functional>
#include <map>

namespace pxa {

namespace {
struct Range {
    int32_t t, a, b;  // tile index, [a, b) source-local rows
};
struct KeyEnt {
    uint64_t key;
    UpdEntry e;
};
}  // namespace

Plan build_plan(const Symbol &s, const Layout &L, int64_t schur_cblk, const PlanKnobs &knobs) {
    const int64_t cb = s.cblknbr();
    Plan P;
    const int64_t sc = std::min<int64_t>(std::max<int64_t>(schur_cblk, 0), cb);
    P.schur_cblk = (int32_t)sc;
    auto factored = [&](int64_t c) { return c < sc; };
    // panel row maps
    P.rowbase.assign(cb + 1, 0);
    for (int64_t c = 0; c < cb; c++) P.rowbase[c + 1] = P.rowbase[c] + L.cblk_stride[c];
    P.panel_grow.assign(P.rowbase[cb], 0);
    P.panel_blk.assign(P.rowbase[cb], 0);
    for (int64_t c = 0; c < cb; c++)
        for (int64_t b = s.cblk_bloknum[c]; b < s.cblk_bloknum[c + 1]; b++) {
            int64_t nr = s.blok_lrow[b] - s.blok_frow[b] + 1;
            for (int64_t t = 0; t < nr; t++) {
                P.panel_grow[P.rowbase[c] + L.blok_roff[b] + t] = (int32_t)(s.blok_frow[b] + t);
                P.panel_blk[P.rowbase[c] + L.blok_roff[b] + t] = (int32_t)b;
            }
        }
    P.winv_off.assign(cb + 1, 0);
    for (int64_t c = 0; c < cb; c++) {
        int64_t w = s.cblk_fcol[c + 1] - s.cblk_fcol[c];
        P.winv_off[c + 1] = P.winv_off[c] + w * w;
    }
    P.winvnbr = P.winv_off[cb];

    // flop accounting
    for (int64_t c = 0; c < cb; c++) {
        const double K = (double)(s.cblk_fcol[c + 1] - s.cblk_fcol[c]);
        const double M = (double)(L.cblk_stride[c]) - K;
        P.flops_diag += K * ((K / 6.0 + 0.5) * K + 1.0 / 3.0) + K * ((K / 6.0) * K - 1.0 / 6.0);
        P.flops_trsm += M * K * (K + 1.0);
        for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++) {
            const double nrow = (double)(s.blok_lrow[b] - s.blok_frow[b] + 1);
            const double m = (double)(L.cblk_stride[c] - L.blok_roff[b]);
            P.flops_upd_potrf += 2.0 * m * nrow * K;
            P.flops_upd_lu += 2.0 * m * nrow * K + ((b + 1 < s.cblk_bloknum[c + 1]) ? 2.0 * (m - nrow) * nrow * K : 0.0);
        }
    }
    P.flops_upd = P.flops_upd_potrf;

    // levels
    const int32_t nl = L.nlevels;
    P.nlevels = nl;
    P.lv_cblk_ptr.assign(nl + 1, 0);
    for (int64_t c = 0; c < cb; c++) P.lv_cblk_ptr[L.cblk_level[c] + 1]++;
    for (int32_t l = 0; l < nl; l++) P.lv_cblk_ptr[l + 1] += P.lv_cblk_ptr[l];
    P.lv_cblks.assign(cb, 0);
    P.lv_fact_cnt.assign(nl, 0);
    {
        // factored cblks first within each level, then the Schur ones
        std::vector<int32_t> pos(P.lv_cblk_ptr.begin(), P.lv_cblk_ptr.end() - 1);
        for (int64_t c = 0; c < sc; c++) {
            P.lv_cblks[pos[L.cblk_level[c]]++] = (int32_t)c;
            P.lv_fact_cnt[L.cblk_level[c]]++;
        }
        for (int64_t c = sc; c < cb; c++) P.lv_cblks[pos[L.cblk_level[c]]++] = (int32_t)c;
        for (int64_t c = 0; c < sc; c++) P.fact_cblks.push_back((int32_t)c);
    }
    auto width = [&](int64_t c) { return s.cblk_fcol[c + 1] - s.cblk_fcol[c]; };
    auto rows = [&](int64_t b) { return s.blok_lrow[b] - s.blok_frow[b] + 1; };

    P.lv_upd_ptr.assign(nl + 1, 0);
    P.lv_upd_split.assign(nl, 0);
    P.lv_upd_dnext.assign(nl, 0);
    P.lv_upd_mnext.assign(nl, 0);
    P.lv_wt_next.assign(nl, 0);
    P.lv_wt_ptr.assign(nl + 1, 0);
    P.lv_wsize.assign(nl, 0);
    P.lv_single_pass.assign(nl, 0);
    P.lv_panel_right.assign(nl, 0);
    std::vector<KeyEnt> ke;
    std::vector<Range> rr, cr;
    const int64_t WT_P = 128;     // source rows (p) of one product tile; 64 columns (q)

    // Update tiles of target panels from the sources `srcs` (increasing cblk
    // index), for the targets `keep` accepts: every run of a source's blocks
    // facing one target contributes entries to the 64x64 target tiles it
    // lands in; a tile's entries stay in source order.  Shifts and entries are
    // appended to P.shift and P.ents, the tiles to `out`.
    auto build_items = [&](const std::vector<int32_t> &srcs, const std::function<bool(int64_t)> &keep,
                           std::vector<UpdItem> &out) {
        ke.clear();
        for (const int32_t src : srcs) {
            const int64_t last = s.cblk_bloknum[src + 1];
            int64_t b = s.cblk_bloknum[src] + 1;
            while (b < last) {
                const int64_t f = s.blok_fcblk[b];
                const int64_t i0 = b;
                while (b < last && s.blok_fcblk[b] == f) b++;
                const int64_t i1 = b;
                if (!keep(f)) continue;
                // shift and entry positions are 32-bit in the device lists
                if (P.shift.size() + (size_t)(last - i0) > (size_t)INT32_MAX)
                    throw std::runtime_error("update tile lists exceed 32-bit indices");
                const int32_t sb = (int32_t)P.shift.size();
                int64_t fb = s.cblk_bloknum[f];
                for (int64_t i = i0; i < last; i++) {
                    while (!(s.blok_frow[fb] <= s.blok_frow[i] && s.blok_lrow[i] <= s.blok_lrow[fb])) fb++;
                    P.shift.push_back((int32_t)(L.blok_roff[fb] - s.blok_frow[fb]));
                }
                // row ranges per target row tile
                rr.clear();
                for (int64_t i = i0; i < last; i++) {
                    int64_t pb = L.blok_roff[i], pe = pb + rows(i);
                    int64_t flr0 = s.blok_frow[i] + P.shift[sb + (i - i0)];
                    int64_t p = pb;
                    while (p < pe) {
                        int64_t tr = (flr0 + (p - pb)) / 64;
                        int64_t pn = std::min(pe, pb + ((tr + 1) * 64 - flr0));
                        if (!rr.empty() && rr.back().t == tr && rr.back().b == p) rr.back().b = (int32_t)pn;
                        else rr.push_back(Range{(int32_t)tr, (int32_t)p, (int32_t)pn});
                        p = pn;
                    }
                }
                // column ranges per target column tile
                cr.clear();
                for (int64_t i = i0; i < i1; i++) {
                    int64_t qb = L.blok_roff[i], qe = qb + rows(i);
                    int64_t fcl0 = s.blok_frow[i] - s.cblk_fcol[f];
                    int64_t q = qb;
                    while (q < qe) {
                        int64_t tc = (fcl0 + (q - qb)) / 64;
                        int64_t qn = std::min(qe, qb + ((tc + 1) * 64 - fcl0));
                        if (!cr.empty() && cr.back().t == tc && cr.back().b == q) cr.back().b = (int32_t)qn;
                        else cr.push_back(Range{(int32_t)tc, (int32_t)q, (int32_t)qn});
                        q = qn;
                    }
                }
                const int64_t rb = P.rowbase[src];
                for (const Range &r : rr)
                    for (const Range &cc : cr) {
                        // keep only tiles with at least one pair block(p) >= block(q)
                        if (P.panel_blk[rb + r.b - 1] < P.panel_blk[rb + cc.a]) continue;
                        uint64_t key = ((uint64_t)f << 40) | ((uint64_t)r.t << 16) | (uint64_t)cc.t;
                        ke.push_back(KeyEnt{key, UpdEntry{(int32_t)src, r.a, r.b, cc.a, cc.b, sb, (int32_t)i0, -1}});
                    }
            }
        }
        std::stable_sort(ke.begin(), ke.end(), [](const KeyEnt &a, const KeyEnt &b) { return a.key < b.key; });
        for (size_t k = 0; k < ke.size();) {
            size_t e = k;
            while (e < ke.size() && ke[e].key == ke[k].key) e++;
            UpdItem it;
            it.f = (int32_t)(ke[k].key >> 40);
            it.r0 = (int32_t)((ke[k].key >> 16) & 0xFFFFFF) * 64;
            it.c0 = (int32_t)(ke[k].key & 0xFFFF) * 64;
            if (P.ents.size() + (e - k) > (size_t)INT32_MAX)
                throw std::runtime_error("update tile lists exceed 32-bit indices");
            it.e0 = (int32_t)P.ents.size();
            for (size_t t = k; t < e; t++) P.ents.push_back(ke[t].e);
            it.e1 = (int32_t)P.ents.size();
            it.pad0 = it.pad1 = it.pad2 = 0;
            out.push_back(it);
            k = e;
        }
    };
    // exact flops of the kept (p, q) pairs of one entry: [0] block(p) >=
    // block(q) (LL^t, L side of LU), [1] block(p) > block(q) (U side)
    auto entry_flops = [&](const UpdEntry &en, double &f0, double &f1) {
        const int64_t rb = P.rowbase[en.s];
        const double K2 = 2.0 * (double)width(en.s);
        int64_t qa = en.q0, qb = en.q0;  // first q with blk(q) >= / > blk(p)
        double c0 = 0, c1 = 0;
        for (int64_t p = en.p0; p < en.p1; p++) {
            const int32_t bp = P.panel_blk[rb + p];
            while (qa < en.q1 && P.panel_blk[rb + qa] < bp) qa++;
            while (qb < en.q1 && P.panel_blk[rb + qb] <= bp) qb++;
            c0 += (double)(qb - en.q0);
            c1 += (double)(qa - en.q0);
        }
        f0 = K2 * c0;
        f1 = K2 * c1;
    };
    // leaf region: the longest run of levels from 0 whose cblks are all
    // factored and at most LEAF_MAX_WIDTH wide (downward closed: a cblk's
    // sources sit on lower levels)
    {
        int32_t lr = 0;
        for (; lr < nl; lr++) {
            bool ok = P.lv_cblk_ptr[lr + 1] > P.lv_cblk_ptr[lr];
            for (int32_t k = P.lv_cblk_ptr[lr]; k < P.lv_cblk_ptr[lr + 1] && ok; k++) {
                const int32_t c = P.lv_cblks[k];
                ok = factored(c) && width(c) <= LEAF_MAX_WIDTH;
            }
            if (!ok) break;
        }
        P.leaf_levels = lr;
    }

    for (int64_t c = 0; c < sc; c++)
        if (width(c) > 64) P.wide_cblks.push_back((int32_t)c);
    P.lv_step_ptr.assign(nl + 1, 0);
    P.step_cb_ptr.push_back(0);
    P.step_trsm_ptr.push_back(0);
    P.step_upd_ptr.push_back(0);

    for (int32_t lv = 0; lv < nl; lv++) {
        const int32_t l0 = P.lv_cblk_ptr[lv];
        // the factored cblks of the level: [l0, lf)
        const int32_t lf = l0 + P.lv_fact_cnt[lv];
        // blocked panel steps of this level
        int64_t maxk = 0;
        for (int32_t k = l0; k < lf; k++) maxk = std::max<int64_t>(maxk, width(P.lv_cblks[k]));
        const int64_t nsteps = (maxk + 63) / 64;
        // in-cblk updates: left-looking (the step's block takes all earlier
        // blocks' products, K up to the cblk width, one tile per row tile)
        // where the level has rows enough to fill the device; right-looking
        // (every later block takes the step's products, 64 deep) near the top
        // of the tree, where few row tiles would leave the device idle
        int64_t rowtiles = 0;
        for (int32_t k = l0; k < lf; k++) rowtiles += (L.cblk_stride[P.lv_cblks[k]] + 63) / 64;
        const bool right = knobs.pr(lv) >= 0 ? knobs.pr(lv) != 0 : rowtiles < PANEL_RIGHT_ROWTILES;
        P.lv_panel_right[lv] = right;
        for (int64_t kb = 0; kb < nsteps; kb++) {
            const int64_t k0 = kb * 64;
            for (int32_t k = l0; k < lf; k++) {
                const int32_t c = P.lv_cblks[k];
                const int64_t K = width(c), st = L.cblk_stride[c];
                if (K <= k0) continue;
                const int64_t nb = std::min<int64_t>(64, K - k0);
                P.step_cblks.push_back(c);
                for (int64_t r0 = k0 + nb; r0 < st; r0 += 64) P.ptrsm.push_back(PanelItem{c, (int32_t)r0, 0, 0});
                if (right) {
                    // every later column block by this one, from each
                    // block's diagonal tile down
                    for (int64_t c0 = k0 + 64; c0 < K; c0 += 64)
                        for (int64_t r0 = c0; r0 < st; r0 += 64)
                            P.pupd.push_back(PanelItem{c, (int32_t)r0, (int32_t)c0, 0});
                } else if (kb > 0) {
                    // this column block by the blocks before it, from the
                    // diagonal tile down
                    for (int64_t r0 = k0; r0 < st; r0 += 64) P.pupd.push_back(PanelItem{c, (int32_t)r0, (int32_t)k0, 0});
                }
            }
            P.step_cb_ptr.push_back((int32_t)P.step_cblks.size());
            P.step_trsm_ptr.push_back((int32_t)P.ptrsm.size());
            P.step_upd_ptr.push_back((int32_t)P.pupd.size());
        }
        P.lv_step_ptr[lv + 1] = P.lv_step_ptr[lv] + (int32_t)nsteps;

        // update tiles from the sources of this level
        {
            std::vector<int32_t> srcs(P.lv_cblks.begin() + l0, P.lv_cblks.begin() + lf);
            build_items(srcs, [](int64_t) { return true; }, P.upd);
        }
        // single-entry tiles first (register-direct kernel), then the others
        std::stable_partition(P.upd.begin() + P.lv_upd_ptr[lv], P.upd.end(),
                              [](const UpdItem &it) { return it.e1 - it.e0 == 1; });
        int32_t split = P.lv_upd_ptr[lv];
        while (split < (int32_t)P.upd.size() && P.upd[split].e1 - P.upd[split].e0 == 1) split++;
        P.lv_upd_split[lv] = split;
        P.lv_upd_ptr[lv + 1] = (int32_t)P.upd.size();
        // tiles of the next level's cblks first within each range
        auto next_level = [&](const UpdItem &it) { return L.cblk_level[it.f] == lv + 1; };
        P.lv_upd_dnext[lv] = (int32_t)(std::stable_partition(P.upd.begin() + P.lv_upd_ptr[lv],
                                                             P.upd.begin() + split, next_level) - P.upd.begin());
        P.lv_upd_mnext[lv] =
            (int32_t)(std::stable_partition(P.upd.begin() + split, P.upd.end(), next_level) - P.upd.begin());
        // levels whose multi-source tiles hold about two entries on average
        // are updated in one pass: the product workspace would write and
        // re-read every product for little reuse of the target tile
        {
            int64_t nm = 0, nme = 0;
            for (int32_t t = split; t < P.lv_upd_ptr[lv + 1]; t++) {
                nm++;
                nme += P.upd[t].e1 - P.upd[t].e0;
            }
            P.lv_single_pass[lv] = nm > 0 && (knobs.sp(lv) >= 0 ? knobs.sp(lv) != 0
                                                                : (double)nme <= SINGLE_PASS_ENTS * (double)nm);
        }
        const bool single = P.lv_single_pass[lv] != 0;
        // exact flops of the kept (p, q) pairs of every entry
        for (int32_t t = P.lv_upd_ptr[lv]; t < P.lv_upd_ptr[lv + 1]; t++) {
            const int kind = (t < split) ? 0 : (single ? 2 : 1);
            for (int32_t e = P.upd[t].e0; e < P.upd[t].e1; e++) {
                double f0, f1;
                entry_flops(P.ents[e], f0, f1);
                P.flops_kind[0][kind] += f0;
                P.flops_kind[1][kind] += f1;
                if (lv < P.leaf_levels) {
                    P.flops_kind_leaf[0][kind] += f0;
                    P.flops_kind_leaf[1][kind] += f1;
                }
            }
        }

        // product groups of the multi-source tiles and their 64x64 tiles
        if (single) {
            P.lv_wt_next[lv] = P.lv_wt_ptr[lv];
            P.lv_wt_ptr[lv + 1] = P.lv_wt_ptr[lv];
        } else {
            struct TileKey {
                int32_t g;
                int64_t qp;  // pt << 32 | qt: consecutive tiles share their p rows
                bool operator<(const TileKey &o) const { return g != o.g ? g < o.g : qp < o.qp; }
                bool operator==(const TileKey &o) const { return g == o.g && qp == o.qp; }
            };
            std::map<std::pair<int32_t, int32_t>, int32_t> gid;
            std::vector<TileKey> tl;
            int64_t wsz = 0;
            for (int32_t t = split; t < (int32_t)P.upd.size(); t++)
                for (int32_t e = P.upd[t].e0; e < P.upd[t].e1; e++) {
                    UpdEntry &en = P.ents[e];
                    const auto key = std::make_pair(en.s, en.ib0);
                    auto gi = gid.find(key);
                    int32_t g;
                    if (gi == gid.end()) {
                        const int64_t f = s.blok_fcblk[en.ib0], lastb = s.cblk_bloknum[en.s + 1];
                        int64_t i1 = en.ib0;
                        while (i1 < lastb && s.blok_fcblk[i1] == f) i1++;
                        const int64_t base = L.blok_roff[en.ib0];
                        const int64_t np = L.cblk_stride[en.s] - base;
                        const int64_t nq = (i1 < lastb ? L.blok_roff[i1] : L.cblk_stride[en.s]) - base;
                        g = (int32_t)P.wgrp.size();
                        P.wgrp.push_back(WGroup{wsz, (int32_t)np, (int32_t)nq});
                        wsz += (np * nq + 31) / 32 * 32;
                        gid.emplace(key, g);
                    } else {
                        g = gi->second;
                    }
                    en.grp = g;
                    const int64_t base = L.blok_roff[en.ib0];
                    for (int64_t pt = (en.p0 - base) / WT_P; pt <= (en.p1 - 1 - base) / WT_P; pt++)
                        for (int64_t qt = (en.q0 - base) / 64; qt <= (en.q1 - 1 - base) / 64; qt++)
                            tl.push_back(TileKey{g, (pt << 32) | qt});
                }
            std::sort(tl.begin(), tl.end());
            tl.erase(std::unique(tl.begin(), tl.end()), tl.end());
            P.lv_wsize[lv] = wsz;
            P.wmax = std::max(P.wmax, wsz);
            const int32_t g0 = (int32_t)(P.wgrp.size() - gid.size());
            std::vector<std::pair<int32_t, int32_t>> gsrc(gid.size());
            for (const auto &kv : gid) gsrc[kv.second - g0] = kv.first;
            // tiles of groups facing cblks of level lv+1 first
            std::stable_partition(tl.begin(), tl.end(), [&](const TileKey &t) {
                return L.cblk_level[s.blok_fcblk[gsrc[t.g - g0].second]] == lv + 1;
            });
            int64_t nnext = 0;
            for (const auto &t : tl) nnext += L.cblk_level[s.blok_fcblk[gsrc[t.g - g0].second]] == lv + 1;
            P.lv_wt_next[lv] = P.lv_wt_ptr[lv] + (int32_t)nnext;
            for (const auto &t : tl) {
                const WGroup &G = P.wgrp[t.g];
                const auto &sk = gsrc[t.g - g0];
                const int64_t ts = 64;
                const int64_t qt = t.qp & 0xFFFFFFFF, pt = t.qp >> 32;
                const int64_t base = L.blok_roff[sk.second];
                WTile w;
                w.s = sk.first;
                w.p0 = (int32_t)(base + pt * WT_P);
                w.q0 = (int32_t)(base + qt * ts);
                w.mp = (int32_t)std::min<int64_t>(WT_P, G.ldw - pt * WT_P);
                w.mq = (int32_t)std::min<int64_t>(ts, G.nq - qt * ts);
                w.ldw = G.ldw;
                w.wofs = G.wofs + pt * WT_P + qt * ts * (int64_t)G.ldw;
                P.wtiles.push_back(w);
                P.flops_wtile += 2.0 * w.mp * w.mq * (double)width(w.s);
                if (lv < P.leaf_levels) P.flops_wtile_leaf += 2.0 * w.mp * w.mq * (double)width(w.s);
            }
            P.lv_wt_ptr[lv + 1] = (int32_t)P.wtiles.size();
        }

    }
    // ---- solve nodes, solve levels and their items ----
    {
        const int32_t nmax = knobs.node_max < 0 ? SOLVE_NODE_MAX : knobs.node_max;
        // no merged runs beside a Schur part (its cblks are not solved)
        const bool merge = nmax > 1 && sc == cb;
        P.node_cblk.assign(1, 0);
        P.cblk_node.assign(cb, 0);
        P.cblk_intra.assign(cb, 0);
        for (int64_t c = 0; c < cb;) {
            int64_t e = c + 1, K = width(c);  // members [c, e), K columns
            while (merge && e < cb) {
                // e is the parent of the last member (its first off-diagonal
                // block faces e) ...
                const int64_t fb = s.cblk_bloknum[e - 1] + 1;
                if (fb >= s.cblk_bloknum[e] || s.blok_fcblk[fb] != e || K + width(e) > nmax) break;
                // ... and every member's off-diagonal rows go on with e's columns
                bool ok = true;
                for (int64_t k = c; k < e && ok; k++) {
                    const int64_t r0 = width(k) + P.cblk_intra[k];
                    if (r0 + width(e) > L.cblk_stride[k]) {
                        ok = false;
                        break;
                    }
                    const int32_t *g = &P.panel_grow[P.rowbase[k] + r0];
                    for (int64_t t = 0; t < width(e) && ok; t++) ok = g[t] == s.cblk_fcol[e] + t;
                }
                if (!ok) break;
                for (int64_t k = c; k < e; k++) P.cblk_intra[k] += (int32_t)width(e);
                K += width(e);
                e++;
            }
            const int32_t nd = (int32_t)P.node_cblk.size() - 1;
            for (int64_t k = c; k < e; k++) P.cblk_node[k] = nd;
            P.node_cblk.push_back((int32_t)e);
            c = e;
        }
        const int32_t nn = (int32_t)P.node_cblk.size() - 1;
        auto nfcol = [&](int32_t nd) { return s.cblk_fcol[P.node_cblk[nd]]; };
        auto nwidth = [&](int32_t nd) { return s.cblk_fcol[P.node_cblk[nd + 1]] - nfcol(nd); };
        auto nfact = [&](int32_t nd) { return factored(P.node_cblk[nd]); };
        P.node_winv.assign(nn, 0);
        for (int32_t nd = 0; nd < nn; nd++) {
            const int32_t c0 = P.node_cblk[nd];
            if (P.node_cblk[nd + 1] - c0 == 1) {
                P.node_winv[nd] = P.winv_off[c0];
            } else {
                const int64_t K = nwidth(nd);
                P.node_winv[nd] = P.winvnbr;
                P.winvnbr += K * K;
            }
        }
        // solve levels over the nodes (sources: factored cblks; every source
        // of a node precedes the node's first cblk)
        std::vector<int32_t> nlev(nn, 0);
        for (int64_t c = 0; c < sc; c++) {
            const int32_t nd = P.cblk_node[c];
            for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++) {
                const int32_t t = P.cblk_node[s.blok_fcblk[b]];
                if (t != nd) nlev[t] = std::max(nlev[t], nlev[nd] + 1);
            }
        }
        int32_t nsl = 0;
        for (int32_t nd = 0; nd < nn; nd++) nsl = std::max(nsl, nlev[nd] + 1);
        P.nslevels = nsl;
        std::vector<int32_t> lptr(nsl + 1, 0), lnodes(nn);
        for (int32_t nd = 0; nd < nn; nd++) lptr[nlev[nd] + 1]++;
        for (int32_t l = 0; l < nsl; l++) lptr[l + 1] += lptr[l];
        {
            std::vector<int32_t> pos(lptr.begin(), lptr.end() - 1);
            for (int32_t nd = 0; nd < nn; nd++) lnodes[pos[nlev[nd]]++] = nd;
        }
        // update buffer: one entry per off-diagonal panel row outside the
        // cblk's node, sliced by 64 permuted rows (slice = the rows' j-th
        // entries side by side, padded to the slice's longest row)
        std::vector<int64_t> yoff(cb + 1, 0);
        for (int64_t c = 0; c < cb; c++) yoff[c + 1] = yoff[c] + (L.cblk_stride[c] - width(c));
        P.ycnt.assign(s.n, 0);
        for (int64_t c = 0; c < cb; c++)
            for (int64_t r = width(c) + P.cblk_intra[c]; r < L.cblk_stride[c]; r++)
                P.ycnt[P.panel_grow[P.rowbase[c] + r]]++;
        const int64_t nblk = (s.n + 63) / 64;
        P.ybase.assign(nblk, 0);
        int64_t ytot = 0;
        for (int64_t k = 0; k < nblk; k++) {
            int32_t mx = 0;
            for (int64_t g = 64 * k; g < std::min<int64_t>(s.n, 64 * k + 64); g++) mx = std::max(mx, P.ycnt[g]);
            P.ybase[k] = ytot;
            ytot += 64 * (int64_t)mx;
        }
        P.nyrows = ytot;
        if (ytot > (int64_t)INT32_MAX) throw std::runtime_error("update buffer exceeds 32-bit entries");
        P.ydst.assign(yoff[cb], 0);
        {
            std::vector<int32_t> pos(s.n, 0);
            for (int64_t c = 0; c < cb; c++)
                for (int64_t r = width(c) + P.cblk_intra[c]; r < L.cblk_stride[c]; r++) {
                    const int64_t g = P.panel_grow[P.rowbase[c] + r];
                    P.ydst[yoff[c] + r - width(c)] = (int32_t)(P.ybase[g >> 6] + 64 * (int64_t)pos[g]++ + (g & 63));
                }
        }
        const int32_t FWD_ROWS = 64;  // panel rows of one forward product tile
        // panel rows of one backward product strip: 64 on the levels of
        // narrow cblks, BWD_ROWS_WIDE where a cblk is wider than 64 columns
        // (its diagonal step then sums a quarter of the partials, while the
        // level still has a few hundred strips)
        const int32_t BWD_ROWS = 64, BWD_ROWS_WIDE = 256;
        P.slv_fwd_ptr.assign(nsl + 1, 0);
        P.slv_bwd_ptr.assign(nsl + 1, 0);
        P.slv_diag_ptr.assign(nsl + 1, 0);
        P.slv_maxw.assign(nsl, 1);
        P.slv_maxcw.assign(nsl, 1);
        P.cblk_btile.assign(2 * cb, 0);
        int64_t bslot = 0;  // backward partial slots (numbered over all levels)
        for (int32_t lv = 0; lv < nsl; lv++) {
            bool lv_wide = false;
            for (int32_t k = lptr[lv]; k < lptr[lv + 1]; k++) {
                const int32_t nd = lnodes[k];
                P.slv_maxw[lv] = std::max<int32_t>(P.slv_maxw[lv], (int32_t)nwidth(nd));
                for (int32_t c = P.node_cblk[nd]; c < P.node_cblk[nd + 1]; c++) {
                    P.slv_maxcw[lv] = std::max<int32_t>(P.slv_maxcw[lv], (int32_t)width(c));
                    lv_wide = lv_wide || width(c) > 64;
                }
            }
            for (int32_t k = lptr[lv]; k < lptr[lv + 1]; k++) {
                const int32_t nd = lnodes[k];
                if (!nfact(nd)) continue;  // the Schur part has no solve items
                for (int32_t c = P.node_cblk[nd]; c < P.node_cblk[nd + 1]; c++) {
                    const int64_t K = width(c), st = L.cblk_stride[c];
                    for (int64_t r0 = K + P.cblk_intra[c]; r0 < st; r0 += FWD_ROWS) {
                        const int64_t r1 = std::min<int64_t>(st, r0 + FWD_ROWS);
                        P.fwd.push_back(FwdItem{c, (int32_t)r0, (int32_t)r1, (int32_t)(yoff[c] + r0 - K)});
                    }
                }
            }
            P.slv_fwd_ptr[lv + 1] = (int32_t)P.fwd.size();
            const int64_t brows = lv_wide ? BWD_ROWS_WIDE : BWD_ROWS;
            for (int32_t k = lptr[lv]; k < lptr[lv + 1]; k++) {
                const int32_t nd = lnodes[k];
                if (!nfact(nd)) continue;
                for (int32_t c = P.node_cblk[nd]; c < P.node_cblk[nd + 1]; c++) {
                    const int64_t K = width(c), st = L.cblk_stride[c];
                    P.cblk_btile[2 * c] = (int32_t)P.bwd.size();
                    for (int64_t r0 = K + P.cblk_intra[c]; r0 < st; r0 += brows) {
                        const int64_t r1 = std::min<int64_t>(st, r0 + brows);
                        P.bwd.push_back(BwdItem{c, (int32_t)r0, (int32_t)r1, (int32_t)bslot});
                        bslot += K;
                    }
                    P.cblk_btile[2 * c + 1] = (int32_t)P.bwd.size();
                }
                for (int64_t ib = 0; ib * 64 < nwidth(nd); ib++) P.diag.push_back(DiagItem{nd, (int32_t)ib, 0, 0});
            }
            P.slv_diag_ptr[lv + 1] = (int32_t)P.diag.size();
            P.slv_bwd_ptr[lv + 1] = (int32_t)P.bwd.size();
        }
        P.bwd_part = bslot;
        if (bslot > (int64_t)INT32_MAX) throw std::runtime_error("backward partial slots exceed 32 bits");

        // merge rounds of the nodes' inverses: leaves are the members'
        // 64-column blocks; each round merges adjacent pairs of runs.  The
        // nodes are batched by the factorisation level of their last cblk
        // (after which their panels and leaf inverses are final): a batch
        // gathers the nodes of consecutive levels until it holds
        // WINV_BATCH_TILES tiles (fewer, larger launches) or a node of
        // several cblks (whose merges are long: they start as soon as the
        // node is complete), rounds numbered within each batch.
        const size_t WINV_BATCH_TILES = 1024;
        std::vector<std::vector<int32_t>> by_level(nl);
        for (int32_t nd = 0; nd < nn; nd++)
            if (nfact(nd)) by_level[L.cblk_level[P.node_cblk[nd + 1] - 1]].push_back(nd);
        P.winv_round_ptr.assign(1, 0);
        P.winv_batch_leaf.assign(1, 0);
        P.winv_batch_round.assign(1, 0);
        std::vector<std::vector<WinvTile>> rounds;
        std::vector<int64_t> rsize;
        size_t btiles = 0, leaves0 = 0;
        for (int32_t lv = 0; lv < nl; lv++) {
            bool bgroup = false;
            for (const int32_t nd : by_level[lv]) {
                const bool group = P.node_cblk[nd + 1] - P.node_cblk[nd] > 1;
                bgroup = bgroup || group;
                std::vector<std::pair<int32_t, int32_t>> runs;
                for (int32_t c = P.node_cblk[nd]; c < P.node_cblk[nd + 1]; c++) {
                    const int32_t o = (int32_t)(s.cblk_fcol[c] - nfcol(nd)), w = (int32_t)width(c);
                    for (int32_t ib = 0; ib * 64 < w; ib++) {
                        runs.push_back({o + 64 * ib, o + std::min(w, 64 * ib + 64)});
                        if (group) P.winv_leaves.push_back(WinvLeaf{nd, c, ib, 0});
                    }
                }
                for (size_t r = 0; runs.size() > 1; r++) {
                    if (rounds.size() <= r) {
                        rounds.emplace_back();
                        rsize.push_back(0);
                    }
                    std::vector<std::pair<int32_t, int32_t>> next;
                    for (size_t i = 0; i < runs.size(); i += 2) {
                        if (i + 1 == runs.size()) {
                            next.push_back(runs[i]);
                            break;
                        }
                        const int32_t a = runs[i].first, m = runs[i + 1].first, b = runs[i + 1].second;
                        for (int32_t i0 = 0; i0 < b - m; i0 += 64)
                            for (int32_t j0 = 0; j0 < m - a; j0 += 64) {
                                rounds[r].push_back(WinvTile{nd, a, m, b, i0, j0, rsize[r]});
                                btiles++;
                            }
                        rsize[r] += (int64_t)(b - m) * (m - a);
                        P.flops_winv += (double)(b - m) * (m - a) * ((m - a) + (b - m));
                        next.push_back({a, b});
                    }
                    runs.swap(next);
                }
            }
            if (rounds.empty() && P.winv_leaves.size() == leaves0) continue;
            if (btiles < WINV_BATCH_TILES && !bgroup && lv + 1 < nl) continue;
            for (size_t r = 0; r < rounds.size(); r++) {
                P.winv_tiles.insert(P.winv_tiles.end(), rounds[r].begin(), rounds[r].end());
                P.winv_round_ptr.push_back((int32_t)P.winv_tiles.size());
                P.winv_tmax = std::max(P.winv_tmax, rsize[r]);
            }
            P.winv_batch_lv.push_back(lv);
            P.winv_batch_leaf.push_back((int32_t)P.winv_leaves.size());
            P.winv_batch_round.push_back((int32_t)P.winv_round_ptr.size() - 1);
            rounds.clear();
            rsize.clear();
            btiles = 0;
            leaves0 = P.winv_leaves.size();
        }
    }
    // leaf region task lists
    if (P.leaf_levels > 0) {
        const int32_t nreg = P.lv_cblk_ptr[P.leaf_levels];
        std::vector<char> in_region(cb, 0);
        std::vector<int32_t> srcs(P.lv_cblks.begin(), P.lv_cblks.begin() + nreg);
        std::sort(srcs.begin(), srcs.end());
        for (const int32_t c : srcs) in_region[c] = 1;
        P.leaf_n = nreg;
        P.leaf_ctr0.assign(cb, 0);
        P.leaf_tgt_ptr.assign(cb + 1, 0);
        for (int64_t c = 0; c < cb; c++) {
            if (in_region[c]) {
                int64_t prev = -1;
                for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++) {
                    const int64_t f = s.blok_fcblk[b];
                    if (f == prev) continue;  // blocks facing one cblk are consecutive
                    prev = f;
                    if (!in_region[f]) continue;
                    P.leaf_tgt.push_back((int32_t)f);
                    P.leaf_ctr0[f]++;
                }
            }
            P.leaf_tgt_ptr[c + 1] = (int32_t)P.leaf_tgt.size();
        }
        for (const int32_t c : srcs)
            if (P.leaf_ctr0[c] == 0) P.leaf_init.push_back(c);
        std::vector<UpdItem> in;
        build_items(srcs, [&](int64_t f) { return in_region[f] != 0; }, in);
        // grouped by target (the tile key orders by target first)
        P.leaf_in_ptr.assign(cb + 1, 0);
        for (const UpdItem &it : in) P.leaf_in_ptr[it.f + 1]++;
        for (int64_t c = 0; c < cb; c++) P.leaf_in_ptr[c + 1] += P.leaf_in_ptr[c];
        P.leaf_in = in;
        for (const UpdItem &it : in)
            for (int32_t e = it.e0; e < it.e1; e++) {
                double f0, f1;
                entry_flops(P.ents[e], f0, f1);
                P.flops_leaf_upd += f0;
            }
        // replay the task protocol on the host: every region cblk must become
        // ready exactly once, after every source its tiles read is done; a
        // list that fails leaves the region to the level schedule
        {
            std::vector<int32_t> ctr(P.leaf_ctr0), order(P.leaf_init);
            std::vector<char> done(cb, 0);
            bool ok = true;
            for (size_t i = 0; i < order.size() && ok; i++) {
                const int32_t c = order[i];
                for (int32_t k = P.leaf_in_ptr[c]; k < P.leaf_in_ptr[c + 1] && ok; k++)
                    for (int32_t e = P.leaf_in[k].e0; e < P.leaf_in[k].e1; e++)
                        if (!done[P.ents[e].s] || P.leaf_in[k].f != c) ok = false;
                done[c] = 1;
                for (int32_t k = P.leaf_tgt_ptr[c]; k < P.leaf_tgt_ptr[c + 1]; k++) {
                    const int32_t f = P.leaf_tgt[k];
                    if (--ctr[f] == 0) order.push_back(f);
                    if (ctr[f] < 0) ok = false;
                }
            }
            if (!ok || (int32_t)order.size() != nreg) {
                fprintf(stderr, "pastix_amd: leaf-region task lists failed their replay; using the level schedule\n");
                P.leaf_levels = 0;
                P.leaf_n = 0;
                P.leaf_ctr0.clear();
                P.leaf_init.clear();
                P.leaf_in_ptr.clear();
                P.leaf_in.clear();
                P.leaf_tgt_ptr.clear();
                P.leaf_tgt.clear();
                P.flops_leaf_upd = 0;
                for (int a = 0; a < 2; a++)
                    for (int k = 0; k < 3; k++) P.flops_kind_leaf[a][k] = 0;
                P.flops_wtile_leaf = 0;
                return P;
            }
        }
        build_items(srcs, [&](int64_t f) { return in_region[f] == 0; }, P.leaf_out);
        P.leaf_out_split = (int32_t)(std::stable_partition(P.leaf_out.begin(), P.leaf_out.end(),
                                                           [](const UpdItem &it) { return it.e1 - it.e0 == 1; }) -
                                     P.leaf_out.begin());
        for (size_t t = 0; t < P.leaf_out.size(); t++)
            for (int32_t e = P.leaf_out[t].e0; e < P.leaf_out[t].e1; e++) {
                double f0, f1;
                entry_flops(P.ents[e], f0, f1);
                P.flops_leaf_out[(int32_t)t < P.leaf_out_split ? 0 : 1] += f0;
            }
    }
    return P;
}

}  // namespace pxa
