// Level schedule of the device engine: per dependency level, the diagonal
// factorisation list, the panel-TRSM tiles, the owner-computes update tiles
// and the solve gather items.  Built once per analysis on the host.
#pragma once
#include <cstdint>
#include <vector>
#include "analysis.h"
#include "dev_types.h"

namespace pxa {

struct Plan {
    int32_t nlevels = 0;
    // cblks of each level (increasing cblk index inside a level)
    std::vector<int32_t> lv_cblk_ptr, lv_cblks;
    // panel TRSM tiles
    std::vector<int32_t> lv_trsm_ptr;
    std::vector<TrsmItem> trsm;
    // update tiles: items of level L come from the sources factored at level L
    std::vector<int32_t> lv_upd_ptr;
    std::vector<UpdItem> upd;
    std::vector<UpdEntry> ents;
    std::vector<int32_t> shift;
    // forward gather items (targets of level L) and per-cblk item ranges
    std::vector<int32_t> lv_fwd_ptr;
    std::vector<FwdItem> fwd;
    std::vector<int32_t> fwd_itemptr;  // segment of level L starts at lv_cblk_ptr[L] + L
    // backward gather items
    std::vector<int32_t> lv_bwd_ptr;
    std::vector<BwdItem> bwd;
    std::vector<int32_t> bwd_itemptr;
    int64_t fwd_part = 0, bwd_part = 0;  // partial buffer sizes per right-hand side
    // panel row maps and inverse offsets
    std::vector<int32_t> panel_grow, panel_blk;
    std::vector<int64_t> rowbase, winv_off;
    int64_t winvnbr = 0;
    // algorithmic flops per kernel family (reference flop conventions):
    // diag = FLOPS_DPOTRF/GETRF(K), trsm = FLOPS_DTRSM(Right, M, K) per cblk,
    // update = sum over (cblk, blok) of FLOPS_DGEMM(m, n, k) with the m, n, k
    // of kernels/core_zgemmsp.c:1352-1357 (x2 sides for LU)
    double flops_diag = 0, flops_trsm = 0, flops_upd = 0;
    double flops_upd_potrf = 0, flops_upd_lu = 0;
};

Plan build_plan(const Symbol &s, const Layout &L);

}  // namespace pxa
