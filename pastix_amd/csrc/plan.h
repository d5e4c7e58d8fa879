// Level schedule of the device engine: per dependency level, the diagonal
// factorisation list, the panel-TRSM tiles, the owner-computes update tiles
// and the solve gather items.  Built once per analysis on the host.
#pragma once
#include <cstdint>
#include <vector>
#include "analysis.h"
#include "dev_types.h"

namespace pxa {

// a level whose multi-source tiles average at most this many entries skips
// the product workspace (the one-pass update kernel sums its entries)
constexpr double SINGLE_PASS_ENTS = 2.05;
// levels with fewer 64-row panel tiles than this update their panels
// right-looking (64-deep products over all later blocks)
constexpr int64_t PANEL_RIGHT_ROWTILES = 200;
// widest cblk of the leaf region (one diagonal sub-block per cblk)
constexpr int64_t LEAF_MAX_WIDTH = 64;
// widest solve node by default (columns of a merged run of cblks)
constexpr int32_t SOLVE_NODE_MAX = 1024;

struct Plan {
    int32_t nlevels = 0;
    // cblks of each level (increasing cblk index inside a level)
    std::vector<int32_t> lv_cblk_ptr, lv_cblks;
    // cblks at and after schur_cblk hold the Schur complement: they receive
    // updates but are neither factored nor sources; within each level they
    // come after the factored cblks, of which there are lv_fact_cnt[lv]
    int32_t schur_cblk = 0;
    std::vector<int32_t> lv_fact_cnt;
    std::vector<int32_t> fact_cblks;  // every factored cblk
    // blocked panel steps: level L owns steps [lv_step_ptr[L], lv_step_ptr[L+1]),
    // step s is column block kb = s - lv_step_ptr[L] of the level's cblks
    std::vector<int32_t> lv_step_ptr;
    std::vector<int32_t> step_cb_ptr, step_cblks;
    // per step: panel-TRSM row tiles below the sub-block, and the tiles of
    // the in-cblk update: on right-looking levels (lv_panel_right), after the
    // step, of the later column blocks by the step's block (for each later
    // block, its row tiles from the diagonal tile down); on the others,
    // before the step, of the step's column block by the blocks before it
    // (from the diagonal tile down)
    std::vector<int32_t> step_trsm_ptr, step_upd_ptr;
    std::vector<PanelItem> ptrsm, pupd;
    std::vector<uint8_t> lv_panel_right;
    std::vector<int32_t> wide_cblks;  // cblks wider than 64 columns
    // update tiles: items of level L come from the sources factored at level L;
    // [lv_upd_ptr[L], lv_upd_split[L]) have a single entry, the rest several
    std::vector<int32_t> lv_upd_ptr, lv_upd_split;
    // within each of the two ranges, the tiles of cblks of level L+1 come
    // first: [lv_upd_ptr[L], lv_upd_dnext[L]) and [lv_upd_split[L],
    // lv_upd_mnext[L]), so that the updates the next panels wait for can be
    // issued apart from the others
    std::vector<int32_t> lv_upd_dnext, lv_upd_mnext;
    std::vector<UpdItem> upd;
    std::vector<UpdEntry> ents;
    std::vector<int32_t> shift;
    // multi-source tiles are updated in two phases: the products of their
    // entries are computed as dense tiles in the sources' own row space
    // (wtiles of level L: [lv_wt_ptr[L], lv_wt_ptr[L+1])) into a level
    // workspace of lv_wsize[L] entries, then summed per target tile
    std::vector<WGroup> wgrp;
    std::vector<WTile> wtiles;
    std::vector<int32_t> lv_wt_ptr;
    std::vector<int32_t> lv_wt_next;  // [lv_wt_ptr[L], lv_wt_next[L]): groups facing level L+1
    std::vector<int64_t> lv_wsize;
    // levels whose multi-source tiles are summed in one pass (no product
    // tiles, lv_wsize 0): those averaging at most SINGLE_PASS_ENTS entries
    std::vector<uint8_t> lv_single_pass;
    int64_t wmax = 0;
    // ---- solve ----
    // Solve nodes: runs of consecutive factored cblks solved as one diagonal
    // block.  A run is a chain: each member's parent is the next member and
    // each member's first off-diagonal panel rows are exactly the columns of
    // the members after it, so the run's diagonal block is dense lower
    // triangular; the solve applies its whole inverse at once (one level
    // instead of one per member).  A node of one cblk is the cblk itself.
    std::vector<int32_t> node_cblk;      // nnodes + 1: first cblk of every node
    std::vector<int32_t> cblk_node;      // cblknbr
    std::vector<int64_t> node_winv;      // nnodes: offset of the node's K x K inverse in the W arrays
    std::vector<int32_t> cblk_intra;     // cblknbr: off-diagonal panel rows of a cblk inside its node
    std::vector<int32_t> cblk_btile;     // 2 cblknbr: backward strips [tb, te) of every cblk
    // solve levels: a node's level is one more than the highest level of
    // the nodes that update it
    int32_t nslevels = 0;
    std::vector<int32_t> slv_maxw;       // widest node of each solve level (diagonal kernels)
    std::vector<int32_t> slv_maxcw;      // widest cblk of each solve level (product kernels)
    // forward products of the nodes of solve level L (their members' panel
    // rows outside the node) into the update buffer.  Its entries are sliced
    // by 64 permuted rows: entry j of row g (its j-th source in increasing
    // order) sits at ybase[g / 64] + 64 j + g % 64, so that consecutive rows'
    // j-th entries are contiguous; row g has ycnt[g] entries.
    std::vector<int32_t> slv_fwd_ptr;
    std::vector<FwdItem> fwd;
    int64_t nyrows = 0;                  // update-buffer entries (slices padded to their longest row)
    std::vector<int64_t> ybase;          // (n + 63) / 64
    std::vector<int32_t> ycnt;           // n
    std::vector<int32_t> ydst;           // update-buffer entry of every off-diagonal panel row
                                         // outside its node, at yoff[c] + r - width(c) (rows
                                         // inside the node have none)
    // backward strips of solve level L
    std::vector<int32_t> slv_bwd_ptr;
    std::vector<BwdItem> bwd;
    // diagonal solve steps: every 64-row block of the level's nodes
    std::vector<int32_t> slv_diag_ptr;
    std::vector<DiagItem> diag;
    int64_t bwd_part = 0;  // partial buffer size per right-hand side (slots of all levels)
    // Inverses of the nodes' diagonal blocks, by recursive doubling over
    // 64-column leaves whose inverses the diagonal-block kernels leave in
    // place: round r merges pairs of adjacent runs [a, m), [m, b) whose
    // inverses are known, tiles [winv_round_ptr[r], winv_round_ptr[r+1])
    // (each tile in both phases: T = M21 X11, then X21 = -X22 T).  Batch k
    // holds the nodes whose last cblk is factored at level winv_batch_lv[k]:
    // leaves [winv_batch_leaf[k], winv_batch_leaf[k+1]) and rounds
    // [winv_batch_round[k], winv_batch_round[k+1]).
    std::vector<WinvTile> winv_tiles;
    std::vector<int32_t> winv_round_ptr;
    std::vector<int32_t> winv_batch_lv, winv_batch_leaf, winv_batch_round;
    int64_t winv_tmax = 0;               // T scratch entries of the largest round
    // leaves of the nodes of several cblks (their inverses are copied from
    // the members' into the node's)
    std::vector<WinvLeaf> winv_leaves;
    double flops_winv = 0;               // products of the merge rounds (one side)
    // panel row maps and inverse offsets
    std::vector<int32_t> panel_grow, panel_blk;
    std::vector<int64_t> rowbase, winv_off;
    int64_t winvnbr = 0;
    // Leaf region: the cblks of levels [0, leaf_levels), all factored and at
    // most LEAF_MAX_WIDTH wide, factored by one persistent launch in which a
    // cblk is a task that becomes ready when its dependency counter (the
    // number of region cblks updating it, reference ctrbcnt) reaches zero:
    // it pulls its incoming updates, factors its diagonal block and solves
    // its rows below, then decrements the counters of the region cblks it
    // faces.  leaf_levels = 0: no region.
    int32_t leaf_levels = 0;
    int32_t leaf_n = 0;                  // region cblks
    std::vector<int32_t> leaf_ctr0;      // cblknbr: initial counter (region cblks)
    std::vector<int32_t> leaf_init;      // region cblks with no incoming update, increasing
    // incoming update tiles of region cblk f: leaf_in[leaf_in_ptr[f] ..
    // leaf_in_ptr[f+1]), entries of every region source in source order
    std::vector<int32_t> leaf_in_ptr;    // cblknbr + 1
    std::vector<UpdItem> leaf_in;
    // region cblks each region cblk faces (its counters to decrement)
    std::vector<int32_t> leaf_tgt_ptr;   // cblknbr + 1
    std::vector<int32_t> leaf_tgt;
    // updates of the region's cblks into the cblks above it, applied after
    // the persistent launch: [0, leaf_out_split) single-entry tiles, the
    // rest several entries
    std::vector<UpdItem> leaf_out;
    int32_t leaf_out_split = 0;
    // reference flops of the region: its update pairs (LL^t) and the share
    // of flops_kind that levels [0, leaf_levels) hold
    double flops_leaf_upd = 0;
    double flops_kind_leaf[2][3] = {{0, 0, 0}, {0, 0, 0}};
    double flops_leaf_out[2] = {0, 0};  // leaf_out single-entry and multi-entry tiles (LL^t)
    double flops_wtile_leaf = 0;        // flops_wtile of levels [0, leaf_levels)
    // algorithmic flops per kernel family (reference flop conventions):
    // diag = FLOPS_DPOTRF/GETRF(K), trsm = FLOPS_DTRSM(Right, M, K) per cblk,
    // update = sum over (cblk, blok) of FLOPS_DGEMM(m, n, k) with the m, n, k
    // of kernels/core_zgemmsp.c:1352-1357 (x2 sides for LU)
    double flops_diag = 0, flops_trsm = 0, flops_upd = 0;
    double flops_upd_potrf = 0, flops_upd_lu = 0;
    // the same update flops split by kernel: [side][0] single-source tiles,
    // [side][1] multi-source tiles of two-phase levels, [side][2]
    // multi-source tiles of single-pass levels (side 0: L, block(p) >=
    // block(q); side 1: U of LU, block(p) > block(q)); flops_wtile: product
    // tiles as computed (128 x 64 tiles of the multi-source products, one side)
    double flops_kind[2][3] = {{0, 0, 0}, {0, 0, 0}};
    double flops_wtile = 0;
};

// Per-level schedule choices that override the plan's rules (a tuning table
// supplies them): -1 keeps the rule; single_pass 1 sums the level's
// multi-source tiles in one pass, 0 through the product workspace;
// panel_right 1 updates the level's panels right-looking, 0 left-looking.
struct PlanKnobs {
    std::vector<int8_t> single_pass, panel_right;
    // widest solve node (columns): -1 the default SOLVE_NODE_MAX, 0 or 1
    // every cblk its own node
    int32_t node_max = -1;
    int8_t sp(int64_t lv) const { return lv < (int64_t)single_pass.size() ? single_pass[lv] : -1; }
    int8_t pr(int64_t lv) const { return lv < (int64_t)panel_right.size() ? panel_right[lv] : -1; }
};

// schur_cblk: first cblk of the Schur complement (cblknbr when there is none)
Plan build_plan(const Symbol &s, const Layout &L, int64_t schur_cblk, const PlanKnobs &knobs = PlanKnobs());

}  // namespace pxa
