// Device-side views of the solver structure and the per-level work lists.
// Shared by the kernels (device) and the engine (host) translation units.
#pragma once
#include <cstdint>

namespace pxa {

// Flattened SolverMatrix (reference blend/solver.h:83-213) in device memory.
// Panels use the 1D layout: cblk c owns coef[cblk_coef[c] .. +stride*width),
// column-major with ld = stride; block b is rows [blok_roff[b], +nrows).
struct DevSymbol {
    const int32_t *cblk_fcol;     // cblknbr+1
    const int32_t *cblk_bloknum;  // cblknbr+1
    const int32_t *cblk_stride;   // cblknbr
    const int64_t *cblk_coef;     // cblknbr+1
    const int64_t *cblk_rowbase;  // cblknbr+1: prefix sum of strides
    const int64_t *cblk_winv;     // cblknbr+1: offsets of the width x width diagonal inverses
    const int32_t *blok_frow;
    const int32_t *blok_lrow;
    const int32_t *blok_fcblk;
    const int32_t *blok_lcblk;
    const int32_t *blok_roff;
    const int32_t *panel_grow;    // permuted global row of every panel row
    const int32_t *panel_blk;     // block index of every panel row
    const int32_t *cblk_brow;     // cblknbr+1
    const int32_t *browtab;
    // solve nodes (runs of cblks solved as one diagonal block)
    const int32_t *node_cblk;     // nnodes+1: first cblk of every node
    const int64_t *node_winv;     // nnodes: offset of the node's K x K inverse
    const int32_t *cblk_btile;    // 2 cblknbr: backward strips [tb, te) of every cblk
};

// One target tile of a Schur-complement update: rows [r0,r0+64) x columns
// [c0,c0+64) of cblk f's panel, fed by entries [e0,e1).
struct UpdItem {
    int32_t f, r0, c0, e0, e1, pad0, pad1, pad2;
};

// One source contribution to a tile: rows [p0,p1) and columns [q0,q1) are
// row ranges of source cblk s's panel; row p lands on target row
// panel_grow[p] + shift[sb + panel_blk[p] - ib0].  grp: product group of
// the entry (multi-source tiles only, -1 otherwise).
struct UpdEntry {
    int32_t s, p0, p1, q0, q1, sb, ib0, grp;
};

// Product group of a source s and the run of its blocks facing one target
// (first block ib0): the dense product P(p, q) = A(p, :) . B(q, :)^T over the
// source panel rows p in [roff(ib0), stride) and q in [roff(ib0), roff(ib0)+nq),
// stored column-major at wofs with leading dimension ldw = stride - roff(ib0).
struct WGroup {
    int64_t wofs;
    int32_t ldw, nq;
};

// One tile of a product group (up to 128 p rows x 64 q rows): source panel rows [p0, p0+mp) x
// [q0, q0+mq), written at W[wofs + i + j*ldw].
struct WTile {
    int32_t s, p0, q0, mp, mq, ldw;
    int64_t wofs;
};

// Forward-solve product: off-diagonal panel rows [r0,r1) of cblk c times x_c
// into the update buffer at row offset `slot` (ybuf row of panel row r0).
struct FwdItem {
    int32_t c, r0, r1, slot;
};

// Backward-solve gather: off-diagonal panel rows [r0,r1) of cblk c.
struct BwdItem {
    int32_t c, r0, r1, slot;
};

// Update buffer of the forward solve, sliced by 64 permuted rows: entry j
// of row g at base[g / 64] + 64 j + g % 64 (times the right-hand side's
// leading dimension), cnt[g] entries.
struct YSlices {
    const int64_t *base;
    const int32_t *cnt;
};

// Diagonal step of the solve: row block [64*ib, 64*ib+64) of solve node nd.
struct DiagItem {
    int32_t nd, ib, pad0, pad1;
};

// One 64x64 output tile of a merge round of the diagonal-block inverses:
// node-local column runs [a, m) and [m, b) of node nd with known inverses
// X11 and X22; rows [i0, i0+64) of the run [m, b), columns [j0, j0+64) of
// [a, m).  Phase 1: T = M21 X11 (stored transposed at toff, leading
// dimension m - a); phase 2: X21 = -X22 T.
struct WinvTile {
    int32_t nd, a, m, b, i0, j0;
    int64_t toff;
};

// Leaf ib (64-column block) of member cblk c of a node of several cblks.
struct WinvLeaf {
    int32_t nd, c, ib, pad;
};

// Leaf-region task lists on the device (see Plan::leaf_*): incoming update
// tiles per region cblk, the region cblks each one faces, the initial
// counters and the initially ready cblks.
struct LeafDev {
    const int32_t *in_ptr;   // cblknbr + 1
    const UpdItem *in;
    const int32_t *tgt_ptr;  // cblknbr + 1
    const int32_t *tgt;
    const int32_t *ctr0;     // cblknbr
    const int32_t *init;     // ninit
    int32_t ninit, nregion, cblknbr, pad;
};

// Blocked panel step tile: rows [r0, r0+64) (and columns [c0, c0+64) for the
// trailing update) of cblk c's panel.
struct PanelItem {
    int32_t c, r0, c0, pad;
};

}  // namespace pxa
