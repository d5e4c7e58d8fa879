// Device engine implementation: memory, uploads, and the level-by-level
// drivers of numfact and solve.  Replaces the reference schedulers
// (sopalin/sequential_zpotrf.c:34-286, sequential_zgetrf.c,
// sequential_ztrsm.c:34-322) for one GPU.
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "kernels.h"

namespace pxa {

#define HIPCHK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (side_) (void)hipStreamSynchronize(side_);
    if (winv_) (void)hipStreamSynchronize(winv_);
    if (fact_exec_) (void)hipGraphExecDestroy(fact_exec_);
    for (SolveGraph &g : solve_graphs_) (void)hipGraphExecDestroy(g.exec);
    for (const Owned &o : owned_) (void)hipFree(o.p);
    if (h_pin_) (void)hipHostFree(h_pin_);
    for (hipEvent_t e : pev_) (void)hipEventDestroy(e);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (ev_panel_) (void)hipEventDestroy(ev_panel_);
    if (ev_side_) (void)hipEventDestroy(ev_side_);
    if (side_) (void)hipStreamDestroy(side_);
    if (ev_winv_) (void)hipEventDestroy(ev_winv_);
    if (winv_) (void)hipStreamDestroy(winv_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::fail(const std::string &m) {
    err_ = m;
    fprintf(stderr, "pastix_amd: %s\n", m.c_str());
    return -1;
}

void *Engine::alloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 8;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        fail("hipMalloc of " + std::to_string(bytes) + " bytes failed");
        return nullptr;
    }
    owned_.push_back(Owned{p, bytes});
    bytes_ += (int64_t)bytes;
    return p;
}

void Engine::release_ptr(void *p) {
    if (!p) return;
    for (size_t k = 0; k < owned_.size(); k++)
        if (owned_[k].p == p) {
            (void)hipStreamSynchronize(stream_);
            (void)hipFree(p);
            bytes_ -= (int64_t)owned_[k].bytes;
            owned_.erase(owned_.begin() + (long)k);
            return;
        }
}

template <typename T> static T *upload(Engine &E, const std::vector<T> &v) {
    T *d = (T *)E.alloc(v.size() * sizeof(T));
    if (d && !v.empty() && hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return nullptr;
    return d;
}
template <typename T, typename S> static T *upload_as(Engine &E, const std::vector<S> &v) {
    std::vector<T> t(v.begin(), v.end());
    return upload(E, t);
}

int Engine::init(int device, const Symbol &s, const Layout &L, const Plan &P, int factotype, int prec) {
    dev_ = device;
    HIPCHK(hipSetDevice(dev_));
    // the no-allocation entry points' scratch: its address on this device
    if (noalloc_scratch_init(dev_)) return fail("cannot resolve the device scratch array");
    HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&ev0_));
    HIPCHK(hipEventCreate(&ev1_));
    HIPCHK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev_panel_, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_side_, hipEventDisableTiming));
    HIPCHK(hipStreamCreateWithFlags(&winv_, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev_winv_, hipEventDisableTiming));
    facto_ = factotype;
    prec_ = prec;
    n_ = s.n;
    cblknbr_ = s.cblknbr();
    coefnbr_ = L.coefnbr;
    winvnbr_ = P.winvnbr;
    maxwidth_ = (int)L.maxwidth;
    // widest node and cblk of each solve level: size the solve kernels'
    // workgroups and staging (the leaf levels hold thousands of narrow cblks)
    slv_maxw_.assign(P.slv_maxw.begin(), P.slv_maxw.end());
    slv_maxcw_.assign(P.slv_maxcw.begin(), P.slv_maxcw.end());
    if (n_ >= (int64_t)INT32_MAX) return fail("matrix order exceeds the 32-bit device index range");

    S_.cblk_fcol = upload_as<int32_t>(*this, s.cblk_fcol);
    S_.cblk_bloknum = upload_as<int32_t>(*this, s.cblk_bloknum);
    S_.cblk_stride = upload_as<int32_t>(*this, L.cblk_stride);
    S_.cblk_coef = upload(*this, L.cblk_coef);
    S_.cblk_rowbase = upload(*this, P.rowbase);
    S_.cblk_winv = upload(*this, P.winv_off);
    S_.blok_frow = upload_as<int32_t>(*this, s.blok_frow);
    S_.blok_lrow = upload_as<int32_t>(*this, s.blok_lrow);
    S_.blok_fcblk = upload_as<int32_t>(*this, s.blok_fcblk);
    S_.blok_lcblk = upload_as<int32_t>(*this, L.blok_lcblk);
    S_.blok_roff = upload_as<int32_t>(*this, L.blok_roff);
    S_.panel_grow = upload(*this, P.panel_grow);
    S_.panel_blk = upload(*this, P.panel_blk);
    S_.cblk_brow = upload_as<int32_t>(*this, L.cblk_brow);
    S_.browtab = upload_as<int32_t>(*this, L.browtab);
    S_.node_cblk = upload(*this, P.node_cblk);
    S_.node_winv = upload(*this, P.node_winv);
    S_.cblk_btile = upload(*this, P.cblk_btile);

    lv_cblk_ptr_ = P.lv_cblk_ptr;
    lv_fact_cnt_ = P.lv_fact_cnt;
    schurfcol_ = (P.schur_cblk < cblknbr_) ? s.cblk_fcol[P.schur_cblk] : n_;
    d_fact_cblks_ = upload(*this, P.fact_cblks);
    n_fact_ = (int)P.fact_cblks.size();
    // leaf region of LL^t factorisations
    leaf_ = LeafDev{};
    leaf_levels_ = 0;
    if (factotype != 2 && P.leaf_levels > 0 && P.leaf_n > 0) {
        leaf_levels_ = P.leaf_levels;
        int rate_khz = 0;
        if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev_) != hipSuccess || rate_khz <= 0)
            rate_khz = 100000;
        leaf_maxwait_ = 2000LL * rate_khz;
        leaf_.in_ptr = upload(*this, P.leaf_in_ptr);
        leaf_.in = upload(*this, P.leaf_in);
        leaf_.tgt_ptr = upload(*this, P.leaf_tgt_ptr);
        leaf_.tgt = upload(*this, P.leaf_tgt);
        leaf_.ctr0 = upload(*this, P.leaf_ctr0);
        leaf_.init = upload(*this, P.leaf_init);
        leaf_.ninit = (int32_t)P.leaf_init.size();
        leaf_.nregion = P.leaf_n;
        leaf_.cblknbr = (int32_t)cblknbr_;
        d_leaf_ctr_ = (int32_t *)alloc(std::max<int64_t>(cblknbr_, 1) * sizeof(int32_t));
        d_leaf_queue_ = (int32_t *)alloc(std::max<int32_t>(P.leaf_n, 1) * sizeof(int32_t));
        d_leaf_sync_ = (int32_t *)alloc(4 * sizeof(int32_t));
        d_leaf_out_ = upload(*this, P.leaf_out);
        n_leaf_out_ = (int32_t)P.leaf_out.size();
        leaf_out_split_ = P.leaf_out_split;
        if (!leaf_.in_ptr || !leaf_.tgt_ptr || !leaf_.ctr0 || !leaf_.init || !d_leaf_ctr_ || !d_leaf_queue_ ||
            !d_leaf_sync_)
            return fail("leaf region lists");
    }
    lv_upd_ptr_ = P.lv_upd_ptr;
    lv_upd_split_ = P.lv_upd_split;
    lv_upd_dnext_ = P.lv_upd_dnext;
    lv_upd_mnext_ = P.lv_upd_mnext;
    lv_wt_next_ = P.lv_wt_next;
    slv_fwd_ptr_ = P.slv_fwd_ptr;
    slv_bwd_ptr_ = P.slv_bwd_ptr;
    d_lv_cblks_ = upload(*this, P.lv_cblks);
    lv_step_ptr_ = P.lv_step_ptr;
    step_cb_ptr_ = P.step_cb_ptr;
    step_trsm_ptr_ = P.step_trsm_ptr;
    step_upd_ptr_ = P.step_upd_ptr;
    lv_panel_right_ = P.lv_panel_right;
    d_step_cblks_ = upload(*this, P.step_cblks);
    d_ptrsm_ = upload(*this, P.ptrsm);
    d_pupd_ = upload(*this, P.pupd);
    d_winv_leaves_ = upload(*this, P.winv_leaves);
    n_winv_leaves_ = (int)P.winv_leaves.size();
    d_winv_tiles_ = upload(*this, P.winv_tiles);
    winv_round_ptr_ = P.winv_round_ptr;
    winv_batch_lv_ = P.winv_batch_lv;
    winv_batch_leaf_ = P.winv_batch_leaf;
    winv_batch_round_ = P.winv_batch_round;
    ys_.base = upload(*this, P.ybase);
    ys_.cnt = upload(*this, P.ycnt);
    d_ydst_ = upload(*this, P.ydst);
    nyrows_ = P.nyrows;
    d_shift_ = upload(*this, P.shift);
    d_upd_ = upload(*this, P.upd);
    d_ents_ = upload(*this, P.ents);
    d_fwd_ = upload(*this, P.fwd);
    d_diag_ = upload(*this, P.diag);
    slv_diag_ptr_ = P.slv_diag_ptr;
    d_bwd_ = upload(*this, P.bwd);
    bwd_part_ = P.bwd_part;
    {
        std::vector<int64_t> dp(n_);
        for (int64_t c = 0; c < cblknbr_; c++) {
            const int64_t f = s.cblk_fcol[c], K = s.cblk_fcol[c + 1] - f, st = L.cblk_stride[c];
            for (int64_t l = 0; l < K; l++) dp[f + l] = L.cblk_coef[c] + l + l * st;
        }
        d_diagpos_ = upload(*this, dp);
    }
    d_wgrp_ = upload(*this, P.wgrp);
    d_wtiles_ = upload(*this, P.wtiles);
    lv_wt_ptr_ = P.lv_wt_ptr;
    lv_wsize_ = P.lv_wsize;
    lv_single_pass_ = P.lv_single_pass;

    const size_t fsz = (prec_ == 2) ? sizeof(float) : sizeof(double);
    // level workspace of the two-phase update: the largest level's products,
    // capped at a quarter of the free device memory (levels above the cap
    // keep the single-pass LDS-accumulating kernel)
    {
        size_t freeb = 0, totb = 0;
        HIPCHK(hipMemGetInfo(&freeb, &totb));
        const int64_t cap = (int64_t)(freeb / 4 / fsz);
        wcap_ = std::min<int64_t>(P.wmax, cap);
        if (wcap_ > 0) {
            d_wspace_ = alloc((size_t)wcap_ * fsz);
            if (!d_wspace_) return -1;
        }
    }
    d_L_ = alloc(coefnbr_ * fsz);
    d_W1_ = alloc(winvnbr_ * fsz);
    if (!d_L_ || !d_W1_) return -1;
    HIPCHK(hipMemsetAsync(d_W1_, 0, winvnbr_ * fsz, stream_));
    if (facto_ == 2) {
        d_U_ = alloc(coefnbr_ * fsz);
        d_W2_ = alloc(winvnbr_ * fsz);
        if (!d_U_ || !d_W2_) return -1;
        HIPCHK(hipMemsetAsync(d_W2_, 0, winvnbr_ * fsz, stream_));
    } else {
        d_W2_ = d_W1_;
    }
    d_stat_ = (int *)alloc(4 * sizeof(int));
    // merge rounds of the solve nodes' inverses: the largest round's products
    if (P.winv_tmax > 0 && !(d_tbuf_ = alloc((size_t)P.winv_tmax * fsz))) return -1;
    d_WB_ = alloc(winvnbr_ * fsz);
    if (d_WB_) HIPCHK(hipMemsetAsync(d_WB_, 0, winvnbr_ * fsz, stream_));
    // right-hand sides of one sweep: up to 32, the kernels taking them in
    // groups of 2 or 8 on the grid (the factor is read once per sweep); the
    // dependency-driven solve runs 8 at a time
    nrhs_chunk_ = 32;
    d_part_ = alloc((size_t)std::max<int64_t>(bwd_part_, 1) * nrhs_chunk_ * fsz);
    d_ybuf_ = alloc((size_t)std::max<int64_t>(nyrows_, 1) * nrhs_chunk_ * fsz);
    d_z_ = alloc((size_t)std::max<int64_t>(n_, 1) * nrhs_chunk_ * fsz);
    if (!d_z_) return -1;
    if (!d_WB_ || !d_ybuf_) return -1;
    // The update-buffer entries of the Schur cblks' own rows are never
    // written (Schur cblks are not sources) and the Interface forward sweep
    // subtracts every entry of a Schur row: they must read as zero.
    if (schurfcol_ < n_)
        HIPCHK(hipMemsetAsync(d_ybuf_, 0, (size_t)std::max<int64_t>(nyrows_, 1) * nrhs_chunk_ * fsz, stream_));
    d_red_ = (double *)alloc((size_t)(512 * 32 + 64) * sizeof(double));
    if (!d_part_ || !d_red_ || !d_stat_) return -1;
    HIPCHK(hipStreamSynchronize(stream_));
    inited_ = true;
    return 0;
}

double *Engine::host_block(int64_t count) {
    if (h_pin_cap_ < count) {
        if (h_pin_) (void)hipHostFree(h_pin_);
        h_pin_ = nullptr;
        h_pin_cap_ = 0;
        if (hipHostMalloc((void **)&h_pin_, (size_t)count * sizeof(double), hipHostMallocDefault) != hipSuccess) {
            h_pin_ = nullptr;
            fail("hipHostMalloc of " + std::to_string(count * sizeof(double)) + " bytes failed");
            return nullptr;
        }
        h_pin_cap_ = count;
    }
    return h_pin_;
}

int Engine::download_diag(double *h) {
    if (!d_diagout_) {
        d_diagout_ = (double *)alloc(n_ * sizeof(double));
        if (!d_diagout_) return -1;
    }
    if (prec_ == 2) launch_gather_diag<float>(n_, d_diagpos_, (const float *)d_L_, d_diagout_, stream_);
    else launch_gather_diag<double>(n_, d_diagpos_, (const double *)d_L_, d_diagout_, stream_);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, d_diagout_, n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

int Engine::download_coef_range(int64_t off, int64_t count, void *hL, void *hU) {
    if (off < 0 || count < 0 || off + count > coefnbr_) return fail("coefficient range outside the table");
    const size_t fsz = (prec_ == 2) ? sizeof(float) : sizeof(double);
    if (hL) HIPCHK(hipMemcpyAsync(hL, (char *)d_L_ + off * fsz, count * fsz, hipMemcpyDeviceToHost, stream_));
    if (hU && facto_ == 2)
        HIPCHK(hipMemcpyAsync(hU, (char *)d_U_ + off * fsz, count * fsz, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

int Engine::download_coef(void *hL, void *hU) {
    const size_t fsz = (prec_ == 2) ? sizeof(float) : sizeof(double);
    if (hL) HIPCHK(hipMemcpyAsync(hL, d_L_, coefnbr_ * fsz, hipMemcpyDeviceToHost, stream_));
    if (hU && facto_ == 2) HIPCHK(hipMemcpyAsync(hU, d_U_, coefnbr_ * fsz, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

// Level L: factor the diagonal blocks of its cblks, solve their panels, then
// apply every update these cblks send to later levels.
int Engine::prof_begin(hipStream_t st) {
    if (!profile) return -1;
    while ((int)pev_.size() < nev_ + 2) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        pev_.push_back(e);
    }
    if (hipEventRecord(pev_[nev_], st) != hipSuccess) return -1;
    return nev_++;
}

void Engine::prof_end(int cls, int k, hipStream_t st) {
    if (k < 0) return;
    if (hipEventRecord(pev_[nev_], st) != hipSuccess) return;
    spans_.push_back(Span{cls, k, nev_, cur_lv_});
    nev_++;
}

template <typename T> int Engine::factorize_t(double thr) {
    T *L = (T *)d_L_, *U = (T *)d_U_, *W1 = (T *)d_W1_, *W2 = (T *)d_W2_;
    const size_t nl = lv_cblk_ptr_.size() - 1;
    nev_ = 0;
    spans_.clear();
    // Updates of level L into cblks of level L+1 go on the main stream, right
    // after L's panels; the others go on the side stream, where they run
    // beside the panels of level L+1.  The main stream waits for them before
    // issuing level L+1's updates: targets and workspace regions of the two
    // streams are then disjoint at every moment, and every target tile still
    // receives its contributions level by level, as with one stream.
    bool side_open = false;
    size_t lv_first = 0;
    cur_lv_ = 0;
    // Whole inverses of the solve nodes' diagonal blocks (needed only by the
    // solve), batch by batch: a batch's nodes are complete once the level of
    // their last cblk is factored.  With the second stream they run on a
    // third one from that point on, beside the factorisation; otherwise all
    // at the end, in stream order.
    const int nbatch = (int)winv_batch_lv_.size();
    int kb_next = 0;
    bool winv_open = false;
    auto winv_batch = [&](int k, hipStream_t st) {
        const int l0 = winv_batch_leaf_[k], nlv = winv_batch_leaf_[k + 1] - l0;
        const int r0 = winv_batch_round_[k], nr = winv_batch_round_[k + 1] - r0;
        const int32_t *rp = winv_round_ptr_.data() + r0;
        T *tb = (T *)d_tbuf_;
        if (facto_ != 2) {
            launch_winv_nodes<T>(S_, d_winv_leaves_ + l0, nlv, d_winv_tiles_, rp, nr, L, W1, (T *)d_WB_, tb, st);
        } else {
            // inv(L) from the unit lower L, inv(U^T) (and its transpose
            // inv(U)) from U^T kept in the U panel's square and below
            launch_winv_nodes<T>(S_, d_winv_leaves_ + l0, nlv, d_winv_tiles_, rp, nr, L, W1, (T *)nullptr, tb, st);
            launch_winv_nodes<T>(S_, d_winv_leaves_ + l0, nlv, d_winv_tiles_, rp, nr, U, W2, (T *)d_WB_, tb, st);
        }
    };
    // batches complete before level lv (all factored): onto the third stream
    auto winv_upto = [&](size_t lv) -> int {
        if (!overlap || kb_next >= nbatch || (size_t)winv_batch_lv_[kb_next] >= lv) return 0;
        HIPCHK(hipEventRecord(ev_winv_, stream_));
        HIPCHK(hipStreamWaitEvent(winv_, ev_winv_, 0));
        while (kb_next < nbatch && (size_t)winv_batch_lv_[kb_next] < lv) winv_batch(kb_next++, winv_);
        winv_open = true;
        return 0;
    };
    if (leaf_active()) {
        // the leaf region in one persistent launch, then its updates into the
        // cblks above it; the level schedule takes over from its last level
        int pk = prof_begin(stream_);
        launch_leaf_init(leaf_, d_leaf_ctr_, d_leaf_queue_, d_leaf_sync_, stream_);
        launch_leaf_tasks<T>(S_, leaf_, d_ents_, d_shift_, L, W1, (T *)d_WB_, thr, d_stat_, d_leaf_ctr_,
                             d_leaf_queue_, d_leaf_sync_, leaf_grid > 0 ? leaf_grid : LEAF_GRID, leaf_maxwait_,
                             stream_);
        prof_end(KC_LEAF, pk, stream_);
        if (leaf_out_split_ > 0) {
            pk = prof_begin(stream_);
            launch_update<T>(S_, 0, true, d_leaf_out_, leaf_out_split_, d_ents_, d_shift_, L, L, L, stream_);
            prof_end(KC_DIRECT, pk, stream_);
        }
        if (n_leaf_out_ > leaf_out_split_) {
            pk = prof_begin(stream_);
            launch_update<T>(S_, 0, false, d_leaf_out_ + leaf_out_split_, n_leaf_out_ - leaf_out_split_, d_ents_,
                             d_shift_, L, L, L, stream_);
            prof_end(KC_ACC, pk, stream_);
        }
        lv_first = (size_t)leaf_levels_;
        if (const int rc = winv_upto(lv_first)) return rc;
    }
    for (size_t lv = lv_first; lv < nl; lv++) {
        cur_lv_ = (int)lv;
        int pk = prof_begin(stream_);
        if (facto_ == 2) {
            // the level's factored cblks (the Schur ones keep their sums)
            const int c0 = lv_cblk_ptr_[lv], nc = lv_fact_cnt_[lv];
            launch_lu_square_merge<T>(S_, d_lv_cblks_ + c0, nc, L, U, stream_);
        }
        // blocked panel: 64-column steps, each over every cblk of the level
        for (int s = lv_step_ptr_[lv]; s < lv_step_ptr_[lv + 1]; s++) {
            const int kb = s - lv_step_ptr_[lv];
            const int b0 = step_cb_ptr_[s], nb = step_cb_ptr_[s + 1] - b0;
            const int p0 = step_trsm_ptr_[s], np = step_trsm_ptr_[s + 1] - p0;
            const int q0 = step_upd_ptr_[s], nq = step_upd_ptr_[s + 1] - q0;
            // the step's diagonal sub-block is factored and the rows below
            // are solved against it; the in-cblk products come before (the
            // step's block takes those of the blocks before it) or, on
            // right-looking levels, after (the later blocks take the step's)
            const bool right = lv_panel_right_[lv] != 0;
            if (facto_ == 2) {
                // L -= L U^T and U^T -= U^T L^T on the tiles; rows below the
                // sub-block: L side times inv(U_kk), U side (stored
                // transposed) times inv(L_kk)^T
                // (each pair of sides in one launch)
                if (!right) launch_panel_upd_lu<T>(S_, d_pupd_ + q0, nq, kb, false, L, U, stream_);
                launch_getrf64<T>(S_, d_step_cblks_ + b0, nb, kb, L, U, W1, W2, (T *)d_WB_, thr, d_stat_, stream_);
                launch_panel_trsm2<T>(S_, d_ptrsm_ + p0, np, kb, L, W2, U, W1, stream_);
                if (right) launch_panel_upd_lu<T>(S_, d_pupd_ + q0, nq, kb, true, L, U, stream_);
            } else if (!right) {
                // (one launch for the update and the factorisation is slower
                // here: it reserves the factorisation's LDS for every update
                // workgroup, which the second stream's updates then lack)
                launch_panel_upd<T>(S_, d_pupd_ + q0, nq, kb, false, L, L, L, false, stream_);
                launch_potrf64<T>(S_, d_step_cblks_ + b0, nb, kb, L, W1, (T *)d_WB_, thr, d_stat_, stream_);
                launch_panel_trsm<T>(S_, d_ptrsm_ + p0, np, kb, L, W1, stream_);
            } else {
                // right-looking: the step's update and the next step's
                // diagonal factorisations share one launch
                if (kb == 0)
                    launch_potrf64<T>(S_, d_step_cblks_ + b0, nb, kb, L, W1, (T *)d_WB_, thr, d_stat_, stream_);
                launch_panel_trsm<T>(S_, d_ptrsm_ + p0, np, kb, L, W1, stream_);
                if (s + 1 < lv_step_ptr_[lv + 1]) {
                    const int b1 = step_cb_ptr_[s + 1], nb1 = step_cb_ptr_[s + 2] - b1;
                    launch_panel_upd_potrf<T>(S_, d_step_cblks_ + b1, nb1, d_pupd_ + q0, nq, kb, true, L, W1,
                                              (T *)d_WB_, thr, d_stat_, stream_);
                }
            }
        }
        prof_end(KC_PANEL, pk, stream_);
        const int u0 = lv_upd_ptr_[lv], us = lv_upd_split_[lv], u1 = lv_upd_ptr_[lv + 1];
        const int w0 = lv_wt_ptr_[lv], w1 = lv_wt_ptr_[lv + 1];
        const bool two_phase = u1 > us && !lv_single_pass_[lv] && lv_wsize_[lv] <= wcap_;
        T *Wk = (T *)d_wspace_;
        // one side of an update range: direct tiles [da, db), multi-source
        // tiles [ma, mb) with their product tiles [wa, wb); (A, B) are the
        // operand panels of the products, C the panels they are subtracted from
        auto part = [&](hipStream_t st, int mode, const T *A, const T *B, T *C, int da, int db, int ma, int mb,
                        int wa, int wb) {
            int k;
            if (db > da) {
                k = prof_begin(st);
                launch_update<T>(S_, mode, true, d_upd_ + da, db - da, d_ents_, d_shift_, A, B, C, st);
                prof_end(KC_DIRECT, k, st);
            }
            if (mb == ma) return;
            if (two_phase) {
                if (wb > wa) {
                    k = prof_begin(st);
                    launch_update_wgemm<T>(S_, d_wtiles_ + wa, wb - wa, A, B, Wk, st);
                    prof_end(KC_WGEMM, k, st);
                }
                k = prof_begin(st);
                launch_update_gather<T>(S_, mode, d_upd_ + ma, mb - ma, d_ents_, d_shift_, d_wgrp_, Wk, C, st);
                prof_end(KC_GATHER, k, st);
            } else {
                k = prof_begin(st);
                launch_update<T>(S_, mode, false, d_upd_ + ma, mb - ma, d_ents_, d_shift_, A, B, C, st);
                prof_end(KC_ACC, k, st);
            }
        };
        auto sides = [&](hipStream_t st, int da, int db, int ma, int mb, int wa, int wb) {
            if (facto_ == 2) {
                part(st, 0, L, U, L, da, db, ma, mb, wa, wb);
                part(st, 1, U, L, U, da, db, ma, mb, wa, wb);
            } else {
                part(st, 0, L, L, L, da, db, ma, mb, wa, wb);
            }
        };
        if (!overlap) {
            sides(stream_, u0, us, us, u1, w0, w1);
            continue;
        }
        const int dn = lv_upd_dnext_[lv], mn = lv_upd_mnext_[lv], wn = lv_wt_next_[lv];
        if (const int rc = winv_upto(lv + 1)) return rc;
        HIPCHK(hipEventRecord(ev_panel_, stream_));
        if (side_open) HIPCHK(hipStreamWaitEvent(stream_, ev_side_, 0));
        sides(stream_, u0, dn, us, mn, w0, wn);
        HIPCHK(hipStreamWaitEvent(side_, ev_panel_, 0));
        sides(side_, dn, us, mn, u1, wn, w1);
        HIPCHK(hipEventRecord(ev_side_, side_));
        side_open = true;
    }
    if (side_open) HIPCHK(hipStreamWaitEvent(stream_, ev_side_, 0));
    // the remaining inverse batches (all of them on one stream), profiled
    // with the panel steps, outside every level
    if (kb_next < nbatch) {
        cur_lv_ = (int)nl;
        const int pk = prof_begin(stream_);
        while (kb_next < nbatch) winv_batch(kb_next++, stream_);
        prof_end(KC_PANEL, pk, stream_);
    }
    if (winv_open) {
        HIPCHK(hipEventRecord(ev_winv_, winv_));
        HIPCHK(hipStreamWaitEvent(stream_, ev_winv_, 0));
    }
    // LU: the squares in the reference layout once every inverse is formed
    if (facto_ == 2) launch_lu_square_layout<T>(S_, d_fact_cblks_, n_fact_, L, U, stream_);
    HIPCHK(hipGetLastError());
    return 0;
}

int Engine::snapshot_coef() {
    const size_t fsz = (prec_ == 2) ? sizeof(float) : sizeof(double);
    if (!d_L0_) {
        d_L0_ = alloc(coefnbr_ * fsz);
        if (!d_L0_) return -1;
        if (facto_ == 2) {
            d_U0_ = alloc(coefnbr_ * fsz);
            if (!d_U0_) return -1;
        }
    }
    HIPCHK(hipMemcpyAsync(d_L0_, d_L_, coefnbr_ * fsz, hipMemcpyDeviceToDevice, stream_));
    if (facto_ == 2) HIPCHK(hipMemcpyAsync(d_U0_, d_U_, coefnbr_ * fsz, hipMemcpyDeviceToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

int Engine::restore_coef() {
    if (!d_L0_) return fail("no coefficient snapshot");
    const size_t fsz = (prec_ == 2) ? sizeof(float) : sizeof(double);
    HIPCHK(hipMemcpyAsync(d_L_, d_L0_, coefnbr_ * fsz, hipMemcpyDeviceToDevice, stream_));
    if (facto_ == 2) HIPCHK(hipMemcpyAsync(d_U_, d_U0_, coefnbr_ * fsz, hipMemcpyDeviceToDevice, stream_));
    return 0;
}

int Engine::upload_perm(const std::vector<int64_t> &permtab) {
    if (!d_perm_) {
        std::vector<int32_t> p(permtab.begin(), permtab.end());
        d_perm_ = upload(*this, p);
        if (!d_perm_) return -1;
        // the inverse (peritab), for kernels that read user-ordered vectors
        // in the permuted row order
        std::vector<int32_t> q(p.size());
        for (size_t i = 0; i < p.size(); i++) q[p[i]] = (int32_t)i;
        d_iperm_ = upload(*this, q);
        if (!d_iperm_) return -1;
    }
    return 0;
}

// Device right-hand sides in the user ordering: permute into the work block,
// solve, permute back (reference sopalin/pastix_task_solve.c:550-577 with
// bvec_zlapmr replaced by a gather/scatter kernel).
int Engine::solve_host_part(int part, int flttype, void *b, int64_t ldb, int nrhs, double *seconds) {
    // b is in the permuted ordering already; partial sweeps run eagerly
    std::vector<int64_t> id(n_);
    for (int64_t i = 0; i < n_; i++) id[i] = i;
    const bool g = use_graph;
    use_graph = false;
    part_ = part;
    const int rc = solve_host(id.data(), flttype, b, ldb, nrhs, seconds);
    part_ = 0;
    use_graph = g;
    return rc;
}

int Engine::enqueue_solve(void *dx, int64_t ldx, int nrhs, bool user) {
    if (!user) return (prec_ == 2) ? solve_t<float>((float *)dx, ldx, nrhs) : solve_t<double>((double *)dx, ldx, nrhs);
    if (prec_ == 2) launch_permute<float>(n_, nrhs, d_perm_, true, (const float *)dx, ldx, (float *)d_rhs_, n_, stream_);
    else launch_permute<double>(n_, nrhs, d_perm_, true, (const double *)dx, ldx, (double *)d_rhs_, n_, stream_);
    const int rc = (prec_ == 2) ? solve_t<float>((float *)d_rhs_, n_, nrhs) : solve_t<double>((double *)d_rhs_, n_, nrhs);
    if (rc) return rc;
    if (prec_ == 2) launch_permute<float>(n_, nrhs, d_perm_, false, (const float *)d_rhs_, n_, (float *)dx, ldx, stream_);
    else launch_permute<double>(n_, nrhs, d_perm_, false, (const double *)d_rhs_, n_, (double *)dx, ldx, stream_);
    HIPCHK(hipGetLastError());
    return 0;
}

hipGraphExec_t Engine::solve_graph(void *dx, int64_t ldx, int nrhs, bool user) {
    for (const SolveGraph &g : solve_graphs_)
        if (g.ptr == dx && g.ld == ldx && g.nrhs == nrhs && g.user == user && g.iface == schur_interface) return g.exec;
    if (solve_graphs_.size() >= 16) {
        (void)hipGraphExecDestroy(solve_graphs_.front().exec);
        solve_graphs_.erase(solve_graphs_.begin());
    }
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    const int rc = enqueue_solve(dx, ldx, nrhs, user);
    const hipError_t e = hipStreamEndCapture(stream_, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        fail("solve graph capture failed");
        return nullptr;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        fail(std::string("solve graph instantiation: ") + hipGetErrorString(ei));
        return nullptr;
    }
    solve_graphs_.push_back(SolveGraph{dx, ldx, nrhs, user, schur_interface, ex});
    return ex;
}

// Device solve in the user ordering: permute into the work block, solve,
// permute back (the timed region includes both permutations).
int Engine::solve_user_device(void *db, int64_t ldb, int nrhs, double *seconds) {
    if (!d_perm_) return fail("permutation not uploaded");
    const size_t fsz = (prec_ == 2) ? sizeof(float) : sizeof(double);
    if (rhs_cap_ < (int64_t)nrhs * n_) {
        release(d_rhs_);
        rhs_cap_ = (int64_t)nrhs * n_;
        d_rhs_ = alloc(rhs_cap_ * fsz);
        if (!d_rhs_) {
            rhs_cap_ = 0;
            return -1;
        }
    }
    HIPCHK(hipEventRecord(ev0_, stream_));
    if (use_graph) {
        hipGraphExec_t ex = solve_graph(db, ldb, nrhs, true);
        if (!ex) return -1;
        HIPCHK(hipGraphLaunch(ex, stream_));
    } else {
        const int rc = enqueue_solve(db, ldb, nrhs, true);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(ev1_, stream_));
    HIPCHK(hipEventSynchronize(ev1_));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev0_, ev1_));
    if (seconds) *seconds = ms * 1e-3;
    return 0;
}

int Engine::enqueue_factorize(double thr) {
    launch_zero_int(4, d_stat_, stream_);
    return (prec_ == 2) ? factorize_t<float>(thr) : factorize_t<double>(thr);
}

int Engine::factorize(double thr, int64_t *npiv, int *negdiag, double *seconds) {
    if (!inited_) return fail("engine not initialised");
    const bool graph = use_graph && !profile;
    const int leaf_key = leaf_active() ? 1 + leaf_grid : 0;
    if (graph && (!fact_exec_ || thr != fact_thr_ || overlap != fact_overlap_ || leaf_key != fact_leaf_)) {
        if (fact_exec_) {
            HIPCHK(hipGraphExecDestroy(fact_exec_));
            fact_exec_ = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_factorize(thr);
        const hipError_t e = hipStreamEndCapture(stream_, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : fail(std::string("numfact graph capture: ") + hipGetErrorString(e));
        }
        const hipError_t ei = hipGraphInstantiate(&fact_exec_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            fact_exec_ = nullptr;
            return fail(std::string("numfact graph instantiation: ") + hipGetErrorString(ei));
        }
        fact_thr_ = thr;
        fact_overlap_ = overlap;
        fact_leaf_ = leaf_key;
    }
    HIPCHK(hipEventRecord(ev0_, stream_));
    if (graph) {
        HIPCHK(hipGraphLaunch(fact_exec_, stream_));
    } else {
        const int rc = enqueue_factorize(thr);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(ev1_, stream_));
    HIPCHK(hipEventSynchronize(ev1_));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev0_, ev1_));
    if (profile) {
        for (int c = 0; c < KC_N; c++) {
            kt[c] = 0;
            kn[c] = 0;
        }
        lv_panel_s.assign(lv_cblk_ptr_.size() - 1, 0.0);
        lv_upd_s.assign(lv_cblk_ptr_.size() - 1, 0.0);
        for (const Span &sp : spans_) {
            float a = 0;
            HIPCHK(hipEventElapsedTime(&a, pev_[sp.e0], pev_[sp.e1]));
            kt[sp.cls] += a * 1e-3;
            kn[sp.cls]++;
            const int lv = std::max(sp.lv, 0);
            if (lv < (int)lv_panel_s.size()) {
                if (sp.cls == KC_PANEL || sp.cls == KC_LEAF) lv_panel_s[lv] += a * 1e-3;
                else lv_upd_s[lv] += a * 1e-3;
            }
        }
        t_diag = kt[KC_PANEL] + kt[KC_LEAF];
        n_diag = kn[KC_PANEL] + kn[KC_LEAF];
        t_upd = kt[KC_DIRECT] + kt[KC_WGEMM] + kt[KC_GATHER] + kt[KC_ACC];
        n_upd = kn[KC_DIRECT] + kn[KC_WGEMM] + kn[KC_GATHER] + kn[KC_ACC];
    }
    int st[4];
    HIPCHK(hipMemcpy(st, d_stat_, sizeof(st), hipMemcpyDeviceToHost));
    if (st[2]) return fail("leaf-region task queue did not drain (code " + std::to_string(st[2]) + ")");
    if (npiv) *npiv = st[0];
    if (negdiag) *negdiag = st[1];
    if (seconds) *seconds = ms * 1e-3;
    return 0;
}

// Forward sweep by increasing level: gather the update-buffer rows facing the
// level's cblks and apply inv(L_cc), then write their off-diagonal products to
// the update buffer.  Backward sweep by decreasing level: transposed products
// of the off-diagonal panels on row chunks, then inv(L_cc)^T (inv(U_cc)).
template <typename T> int Engine::solve_t(T *x, int64_t ldx, int nrhs) {
    const T *L = (const T *)d_L_;
    const T *U = (facto_ == 2) ? (const T *)d_U_ : L;
    const T *W1 = (const T *)d_W1_, *WB = (const T *)d_WB_;
    // LL^t: WB = inv(L)^T, the forward inverse transposed (read down its
    // columns by the wide levels' forward GEMV)
    const T *W1T = (facto_ != 2) ? WB : nullptr;
    T *part = (T *)d_part_;
    T *ybuf = (T *)d_ybuf_;
    T *z = (T *)d_z_;
    const int64_t ldy = std::max<int64_t>(nyrows_, 1), ldz = std::max<int64_t>(n_, 1);
    const int nl = (int)slv_maxw_.size();  // solve levels
    const int chunk = nrhs_chunk_;
    for (int r0 = 0; r0 < nrhs; r0 += chunk) {
        const int nr = std::min(chunk, nrhs - r0);
        T *xx = x + (int64_t)r0 * ldx;
        // partial sweeps (pastix_subtask_trsm / _diag): 1 forward only, 2
        // backward only, 3 backward after x *= D (LDL^t: L^{-t} = U^{-1} D),
        // 4 x /= diag(factor)
        if (part_ == 4) {
            launch_diag_scale<T>(schurfcol_, nr, d_diagpos_, L, true, xx, ldx, stream_);
            continue;
        }
        // Schur rows [sc, n): not solved; the forward sweep leaves them as
        // they are (Local) or subtracts the factored part's products
        // (Interface); the backward sweep ignores them (Local: read as zero,
        // then restored) or reads them as the Schur solution (Interface)
        const int64_t sc = schurfcol_;
        const bool schur = sc < n_;
        if (part_ == 0 || part_ == 1) {
            for (int lv = 0; lv < nl; lv++) {
                const int d0 = slv_diag_ptr_[lv], nd = slv_diag_ptr_[lv + 1] - d0;
                const int f0 = slv_fwd_ptr_[lv], nf = slv_fwd_ptr_[lv + 1] - f0;
                launch_fwd_diag<T>(S_, d_diag_ + d0, nd, ys_, ybuf, ldy, xx, ldx, z, ldz, nr, W1, W1T,
                                   slv_maxw_[lv], stream_);
                launch_fwd_prod<T>(S_, d_fwd_ + f0, nf, L, z, ldz, nr, d_ydst_, ybuf, ldy, slv_maxcw_[lv], stream_);
            }
            if (schur)
                launch_schur_rows<T>(sc, n_, nr, schur_interface ? 1 : 0, ys_, ybuf, ldy, xx, ldx, z, ldz,
                                     stream_);
        }
        if (part_ == 1) {
            launch_copy2d<T>(n_, nr, z, ldz, xx, ldx, stream_);
            continue;
        }
        if (part_ == 0 && schur) launch_schur_rows<T>(sc, n_, nr, 3, ys_, ybuf, ldy, xx, ldx, z, ldz, stream_);
        if (part_ == 3) launch_diag_scale<T>(sc, nr, d_diagpos_, L, false, xx, ldx, stream_);
        if (part_ == 2 || part_ == 3) launch_copy2d<T>(n_, nr, xx, ldx, z, ldz, stream_);
        if (schur && !schur_interface)
            launch_schur_rows<T>(sc, n_, nr, 2, ys_, ybuf, ldy, xx, ldx, z, ldz, stream_);
        for (int lv = nl - 1; lv >= 0; lv--) {
            const int d0 = slv_diag_ptr_[lv], nd = slv_diag_ptr_[lv + 1] - d0;
            const int b0 = slv_bwd_ptr_[lv], nb = slv_bwd_ptr_[lv + 1] - b0;
            launch_bwd_prod<T>(S_, d_bwd_ + b0, nb, U, xx, ldx, nr, part, slv_maxcw_[lv], stream_);
            launch_bwd_diag<T>(S_, d_diag_ + d0, nd, d_bwd_, z, ldz, xx, ldx, nr, part, WB, slv_maxw_[lv], stream_);
        }
        if (schur && !schur_interface)
            launch_schur_rows<T>(sc, n_, nr, 3, ys_, ybuf, ldy, xx, ldx, z, ldz, stream_);
    }
    HIPCHK(hipGetLastError());
    return 0;
}

int Engine::solve_device(void *dx, int64_t ldx, int nrhs, double *seconds) {
    if (!inited_) return fail("engine not initialised");
    HIPCHK(hipEventRecord(ev0_, stream_));
    if (use_graph) {
        hipGraphExec_t ex = solve_graph(dx, ldx, nrhs, false);
        if (!ex) return -1;
        HIPCHK(hipGraphLaunch(ex, stream_));
    } else {
        const int rc = enqueue_solve(dx, ldx, nrhs, false);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(ev1_, stream_));
    HIPCHK(hipEventSynchronize(ev1_));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev0_, ev1_));
    if (seconds) *seconds = ms * 1e-3;
    return 0;
}

int Engine::solve_host(const int64_t *perm, int flttype, void *b, int64_t ldb, int nrhs, double *seconds) {
    const size_t fsz = (prec_ == 2) ? sizeof(float) : sizeof(double);
    if (rhs_cap_ < (int64_t)nrhs * n_) {
        release(d_rhs_);
        rhs_cap_ = (int64_t)nrhs * n_;
        d_rhs_ = alloc(rhs_cap_ * fsz);
        if (!d_rhs_) {
            rhs_cap_ = 0;
            return -1;
        }
    }
    // permute on the host into the factor precision: xp[perm[i]] = b[i]
    std::vector<char> tmp((size_t)n_ * nrhs * fsz);
    for (int r = 0; r < nrhs; r++) {
        for (int64_t i = 0; i < n_; i++) {
            const int64_t p = perm[i] + (int64_t)r * n_;
            if (flttype == 3) {
                double v = ((const double *)b)[i + r * ldb];
                if (prec_ == 2) ((float *)tmp.data())[p] = (float)v;
                else ((double *)tmp.data())[p] = v;
            } else {
                float v = ((const float *)b)[i + r * ldb];
                if (prec_ == 2) ((float *)tmp.data())[p] = v;
                else ((double *)tmp.data())[p] = v;
            }
        }
    }
    HIPCHK(hipMemcpyAsync(d_rhs_, tmp.data(), tmp.size(), hipMemcpyHostToDevice, stream_));
    int rc = solve_device(d_rhs_, n_, nrhs, seconds);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(tmp.data(), d_rhs_, tmp.size(), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (int r = 0; r < nrhs; r++) {
        for (int64_t i = 0; i < n_; i++) {
            const int64_t p = perm[i] + (int64_t)r * n_;
            double v = (prec_ == 2) ? (double)((float *)tmp.data())[p] : ((double *)tmp.data())[p];
            if (flttype == 3) ((double *)b)[i + r * ldb] = v;
            else ((float *)b)[i + r * ldb] = (float)v;
        }
    }
    return 0;
}

int Engine::upload_matrix(const HostBcsc &bc) {
    if (!have_matrix_) {
        // a new pattern: the buffers of the previous one are freed first, so
        // that an instance whose input alternates between patterns keeps a
        // constant footprint
        release(d_rowptr_);
        release(d_cols_);
        release(d_vals_);
        release(d_csrdiag_);
        release(d_diag0_);
        release(d_cmap_);
        release(d_trowptr_);
        release(d_tcols_);
        release(d_tpos_);
        have_fillmap_ = false;
        have_transpose_ = false;
        csr_nnz_ = (int64_t)bc.tvals.size();
        d_rowptr_ = upload(*this, bc.rowptr);
        d_cols_ = upload(*this, bc.cols);
        d_vals_ = (double *)alloc(bc.tvals.size() * sizeof(double));
        if (!d_rowptr_ || !d_cols_ || !d_vals_) return -1;
        std::vector<int64_t> cd(n_, -1);
        for (int64_t i = 0; i < n_; i++)
            for (int64_t k = bc.rowptr[i]; k < bc.rowptr[i + 1]; k++)
                if (bc.cols[k] == i) cd[i] = k;
        diag_complete_ = std::find(cd.begin(), cd.end(), (int64_t)-1) == cd.end();
        d_csrdiag_ = upload(*this, cd);
        csrdiag_ = cd;
        d_diag0_ = (double *)alloc(n_ * sizeof(double));
        if (!d_csrdiag_ || !d_diag0_) return -1;
        have_matrix_ = true;
    }
    // the cached diagonal positions hold only for the uploaded pattern
    if ((int64_t)bc.tvals.size() != csr_nnz_) return fail("matrix pattern changed without drop_matrix");
    {
        std::vector<double> dg(n_, 0.0);
        diag_sum_ = 0;
        for (int64_t i = 0; i < n_; i++)
            if (csrdiag_[i] >= 0) dg[i] = bc.tvals[csrdiag_[i]];
        for (int64_t i = 0; i < n_; i++) diag_sum_ += dg[i];
        HIPCHK(hipMemcpy(d_diag0_, dg.data(), n_ * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpyAsync(d_vals_, bc.tvals.data(), bc.tvals.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

int Engine::upload_transpose(const HostBcsc &bc) {
    if (!have_matrix_ || (int64_t)bc.tvals.size() != csr_nnz_) return fail("transpose needs the uploaded matrix");
    release(d_trowptr_);
    release(d_tcols_);
    release(d_tpos_);
    std::vector<int64_t> tp(n_ + 1, 0), pos(csr_nnz_);
    std::vector<int32_t> tc(csr_nnz_);
    for (int64_t k = 0; k < csr_nnz_; k++) tp[bc.cols[k] + 1]++;
    for (int64_t j = 0; j < n_; j++) tp[j + 1] += tp[j];
    std::vector<int64_t> next(tp.begin(), tp.end() - 1);
    for (int64_t i = 0; i < n_; i++)  // rows in increasing order: columns of A^T sorted
        for (int64_t k = bc.rowptr[i]; k < bc.rowptr[i + 1]; k++) {
            const int64_t q = next[bc.cols[k]]++;
            tc[q] = (int32_t)i;
            pos[q] = k;
        }
    d_trowptr_ = upload(*this, tp);
    d_tcols_ = upload(*this, tc);
    d_tpos_ = upload(*this, pos);
    if (!d_trowptr_ || !d_tcols_ || !d_tpos_) return -1;
    have_transpose_ = true;
    return 0;
}

int Engine::upload_fillmap(const std::vector<int64_t> &cmap) {
    if (!have_matrix_ || (int64_t)cmap.size() != csr_nnz_) return fail("fill map does not match the uploaded matrix");
    release(d_cmap_);
    d_cmap_ = upload(*this, cmap);
    if (!d_cmap_) return -1;
    have_fillmap_ = true;
    return 0;
}

int Engine::fill_coef() {
    if (!have_matrix_ || !have_fillmap_) return fail("coefficient fill needs the matrix and its fill map");
    const size_t fsz = (prec_ == 2) ? sizeof(float) : sizeof(double);
    HIPCHK(hipMemsetAsync(d_L_, 0, coefnbr_ * fsz, stream_));
    if (facto_ == 2) HIPCHK(hipMemsetAsync(d_U_, 0, coefnbr_ * fsz, stream_));
    if (prec_ == 2)
        launch_fill_coef<float>(csr_nnz_, d_cmap_, d_vals_, (float *)d_L_, facto_ == 2 ? (float *)d_U_ : nullptr,
                                stream_);
    else
        launch_fill_coef<double>(csr_nnz_, d_cmap_, d_vals_, (double *)d_L_, facto_ == 2 ? (double *)d_U_ : nullptr,
                                 stream_);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

int Engine::shift_diagonal(double delta) {
    if (!have_matrix_) return fail("matrix not uploaded");
    if (!diag_complete_) return fail("diagonal shift needs every diagonal entry stored");
    if (restore_coef()) return -1;
    if (prec_ == 2) launch_add_diag<float>(n_, d_diagpos_, delta, (float *)d_L_, stream_);
    else launch_add_diag<double>(n_, d_diagpos_, delta, (double *)d_L_, stream_);
    launch_set_csr_diag(n_, d_csrdiag_, d_diag0_, delta, d_vals_, stream_);
    shift_ = delta;
    HIPCHK(hipGetLastError());
    return 0;
}

int Engine::spmv(double alpha, const double *x, double beta, double *y) {
    if (!have_matrix_) return fail("matrix not uploaded");
    launch_spmv<double>(n_, d_rowptr_, d_cols_, d_vals_, alpha, x, beta, y, stream_);
    return 0;
}

int Engine::spmv_t(double alpha, const double *x, double beta, double *y) {
    if (!have_matrix_ || !have_transpose_) return fail("transposed matrix not uploaded");
    launch_spmv_t(n_, d_trowptr_, d_tcols_, d_tpos_, d_vals_, alpha, x, beta, y, stream_);
    HIPCHK(hipGetLastError());
    return 0;
}

int Engine::vec_dot(int64_t n, const double *x, const double *y, double *out) {
    launch_dot(n, x, y, d_red_, d_red_ + 512 * 32, stream_);
    HIPCHK(hipMemcpyAsync(out, d_red_ + 512 * 32, sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}
int Engine::vec_nrm2(int64_t n, const double *x, double *out) {
    launch_nrm2(n, x, d_red_, d_red_ + 512 * 32, stream_);
    HIPCHK(hipMemcpyAsync(out, d_red_ + 512 * 32, sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}
int Engine::vec_axpy(int64_t n, double alpha, const double *x, double *y) {
    launch_axpy(n, alpha, x, y, stream_);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}
int Engine::vec_scal(int64_t n, double alpha, double *x) {
    launch_scal(n, alpha, x, stream_);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}
int Engine::vec_copy(int64_t n, const double *x, double *y) {
    launch_copy(n, x, y, stream_);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}
int Engine::vec_gemv(int64_t m, int64_t n, double alpha, const double *A, int64_t lda, const double *x, double beta,
                     double *y) {
    launch_gemv(m, n, alpha, A, lda, x, beta, y, stream_);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

double Engine::dot(const double *x, const double *y) {
    launch_dot(n_, x, y, d_red_, d_red_ + 512 * 32, stream_);
    double v = 0;
    if (hipMemcpyAsync(&v, d_red_ + 512 * 32, sizeof(double), hipMemcpyDeviceToHost, stream_) != hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess)
        return std::nan("");
    return v;
}

int Engine::precond(double *dv) {
    if (prec_ == 3) return solve_t<double>(dv, n_, 1);
    launch_d2s(n_, dv, d_svec_, stream_);
    int rc = solve_t<float>(d_svec_, n_, 1);
    if (rc) return rc;
    launch_s2d(n_, d_svec_, dv, stream_);
    return 0;
}

void Engine::release_work() {
    release(d_sr_);
    release(d_bvec_);
    release(d_bscal_);
    release(d_bsvec_);
    release(d_cpart_);
    release(d_rhs_);
    sr_cap_ = bvec_cap_ = bscal_cap_ = bsvec_cap_ = cpart_cap_ = rhs_cap_ = 0;
}

int64_t Engine::work_bytes(int R) const {
    // refine_user: permuted B and X (2 n R), batched CG vectors (4 n R), the
    // factor-precision copy of the preconditioned block (n R), per-column
    // scalars and segmented-dot partials (one segment)
    const int64_t nr = n_ * (int64_t)R;
    const int64_t fsz = (prec_ == 2) ? 4 : 8;
    const int64_t nchunk = (n_ + seg_dot_chunk() - 1) / seg_dot_chunk() + 1;
    return 8 * (2 * nr + 4 * nr) + (prec_ == 2 ? fsz * nr : 0) + 8 * 8 * (int64_t)R + 8 * nchunk * R;
}

int Engine::precond_block(double *zb, int R) {
    if (prec_ == 3) return solve_t<double>(zb, n_, R);
    const int64_t nr = n_ * (int64_t)R;
    if (bsvec_cap_ < nr) {
        release(d_bsvec_);
        bsvec_cap_ = nr;
        d_bsvec_ = (float *)alloc(nr * sizeof(float));
        if (!d_bsvec_) {
            bsvec_cap_ = 0;
            return -1;
        }
    }
    launch_d2s(nr, zb, d_bsvec_, stream_);
    const int rc = solve_t<float>(d_bsvec_, n_, R);
    if (rc) return rc;
    launch_s2d(nr, d_bsvec_, zb, stream_);
    return 0;
}

int Engine::refine_cg_batch(double *X, const double *B, int R, const std::vector<int64_t> &segoff, int itermax,
                            double eps, int *niter, double *resid) {
    const int nseg = (int)segoff.size() - 1;
    if (nseg < 1 || segoff.front() != 0 || segoff.back() != n_ || R < 1) return fail("bad batch segments");
    const int64_t nr = n_ * R, ns = (int64_t)nseg * R;
    if (bvec_cap_ < 4 * nr) {
        release(d_bvec_);
        bvec_cap_ = 4 * nr;
        d_bvec_ = (double *)alloc(bvec_cap_ * sizeof(double));
        if (!d_bvec_) {
            bvec_cap_ = 0;
            return -1;
        }
    }
    if (bscal_cap_ < 8 * ns) {
        release(d_bscal_);
        bscal_cap_ = 8 * ns;
        d_bscal_ = (double *)alloc(bscal_cap_ * sizeof(double));
        if (!d_bscal_) {
            bscal_cap_ = 0;
            return -1;
        }
    }
    if (segoff != segoff_host_) {
        if (!d_seg_) {
            d_seg_ = (int32_t *)alloc(n_ * sizeof(int32_t));
            if (!d_seg_) return -1;
        }
        std::vector<int32_t> seg(n_);
        for (int s = 0; s < nseg; s++)
            for (int64_t i = segoff[s]; i < segoff[s + 1]; i++) seg[i] = s;
        HIPCHK(hipMemcpy(d_seg_, seg.data(), n_ * sizeof(int32_t), hipMemcpyHostToDevice));
        if (segoff_dcap_ < (int64_t)segoff.size()) {
            release(d_segoff_);
            segoff_dcap_ = (int64_t)segoff.size();
            d_segoff_ = (int64_t *)alloc(segoff.size() * sizeof(int64_t));
            if (!d_segoff_) {
                segoff_dcap_ = 0;
                return -1;
            }
        }
        HIPCHK(hipMemcpy(d_segoff_, segoff.data(), segoff.size() * sizeof(int64_t), hipMemcpyHostToDevice));
        // chunks of the segmented dot products
        const int64_t ch = seg_dot_chunk();
        std::vector<int32_t> coff(nseg + 1, 0), cseg;
        for (int s = 0; s < nseg; s++) {
            const int64_t nc = std::max<int64_t>(1, (segoff[s + 1] - segoff[s] + ch - 1) / ch);
            coff[s + 1] = coff[s] + (int32_t)nc;
            for (int64_t k = 0; k < nc; k++) cseg.push_back(s);
        }
        nchunk_ = coff[nseg];
        if (chunk_cap_ < nchunk_ + nseg + 1) {
            release(d_coff_);
            chunk_cap_ = nchunk_ + nseg + 1;
            d_coff_ = (int32_t *)alloc(chunk_cap_ * sizeof(int32_t));
            if (!d_coff_) {
                chunk_cap_ = 0;
                return -1;
            }
        }
        d_cseg_ = d_coff_ + nseg + 1;
        HIPCHK(hipMemcpy(d_coff_, coff.data(), coff.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_cseg_, cseg.data(), cseg.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        segoff_host_ = segoff;
    }
    if (cpart_cap_ < (int64_t)nchunk_ * R) {
        release(d_cpart_);
        cpart_cap_ = (int64_t)nchunk_ * R;
        d_cpart_ = (double *)alloc(cpart_cap_ * sizeof(double));
        if (!d_cpart_) {
            cpart_cap_ = 0;
            return -1;
        }
    }
    double *Rv = d_bvec_, *P = d_bvec_ + nr, *Z = d_bvec_ + 2 * nr, *AP = d_bvec_ + 3 * nr;
    double *dots = d_bscal_, *coef = d_bscal_ + 4 * ns;
    std::vector<double> h(4 * ns), rb(ns), normb(ns), rz(ns), a(ns);
    auto segdot = [&](const double *x, const double *y, double *out_h) -> int {
        launch_seg_dot(d_segoff_, d_coff_, d_cseg_, nseg, nchunk_, R, n_, x, y, d_cpart_, dots, stream_);
        HIPCHK(hipMemcpyAsync(out_h, dots, ns * sizeof(double), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        return 0;
    };
    auto upload = [&](const std::vector<double> &v) -> int {
        HIPCHK(hipMemcpyAsync(coef, v.data(), ns * sizeof(double), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        return 0;
    };
    if (segdot(B, B, normb.data())) return -1;
    for (int64_t k = 0; k < ns; k++) normb[k] = std::sqrt(normb[k]);
    for (int c = 0; c < R; c++) {
        launch_copy(n_, B + c * n_, Rv + c * n_, stream_);
        spmv(-1.0, X + c * n_, 1.0, Rv + c * n_);
    }
    auto residuals = [&]() -> double {
        double worst = 0;
        for (int64_t k = 0; k < ns; k++) {
            const double nr_ = std::sqrt(h[k]);
            rb[k] = normb[k] > 0 ? nr_ / normb[k] : 0.0;
            worst = std::max(worst, rb[k]);
        }
        return worst;
    };
    if (segdot(Rv, Rv, h.data())) return -1;
    double worst = residuals();
    int it = 0;
    if (worst > eps && it < itermax) {
        launch_copy(nr, Rv, Z, stream_);
        if (precond_block(Z, R)) return -1;
        launch_copy(nr, Z, P, stream_);
    }
    std::vector<double> pap(ns), rz2(ns);
    while (worst > eps && it < itermax) {
        it++;
        for (int c = 0; c < R; c++) spmv(1.0, P + c * n_, 0.0, AP + c * n_);
        if (segdot(Rv, Z, rz.data()) || segdot(AP, P, pap.data())) return -1;
        for (int64_t k = 0; k < ns; k++) a[k] = (rb[k] > eps && pap[k] != 0.0) ? rz[k] / pap[k] : 0.0;
        if (upload(a)) return -1;
        launch_seg_axpy(n_, R, n_, d_seg_, nseg, coef, 1.0, P, X, stream_);
        launch_seg_axpy(n_, R, n_, d_seg_, nseg, coef, -1.0, AP, Rv, stream_);
        // stopping test before M^{-1} r (its result is unused once every
        // pair has converged)
        if (segdot(Rv, Rv, h.data())) return -1;
        worst = residuals();
        if (!(worst > eps && it < itermax)) break;
        launch_copy(nr, Rv, Z, stream_);
        if (precond_block(Z, R)) return -1;
        if (segdot(Rv, Z, rz2.data())) return -1;
        for (int64_t k = 0; k < ns; k++) a[k] = (rz[k] != 0.0) ? rz2[k] / rz[k] : 0.0;
        if (upload(a)) return -1;
        launch_seg_xpby(n_, R, n_, d_seg_, nseg, coef, Z, P, stream_);
    }
    HIPCHK(hipStreamSynchronize(stream_));
    if (niter) *niter = it;
    if (resid) *resid = worst;
    return 0;
}

// Right-hand sides and solutions in FP64 device memory, user ordering: rows
// are permuted into work blocks (reference pastix_task_refine.c:252-284 does
// it in place on the host), an initial solve x = M^{-1} b is applied when
// asked, then the refinement, and the result is permuted back.
int Engine::refine_user(const double *Bu, int64_t ldb, double *Xu, int64_t ldx, int R,
                        const std::vector<int64_t> &segoff, int method, bool initial_solve, bool user_order,
                        int itermax, int im, double eps, int *niter, double *resid) {
    if (R < 1) return 0;
    if (user_order && !d_perm_) return fail("permutation not uploaded");
    const int64_t nr = n_ * (int64_t)R;
    if (sr_cap_ < 2 * nr) {
        release(d_sr_);
        sr_cap_ = 2 * nr;
        d_sr_ = (double *)alloc(sr_cap_ * sizeof(double));
        if (!d_sr_) {
            sr_cap_ = 0;
            return -1;
        }
    }
    double *PB = d_sr_, *PX = d_sr_ + nr;
    auto in = [&](const double *src, int64_t ld, double *dst) {
        if (user_order) launch_permute<double>(n_, R, d_perm_, true, src, ld, dst, n_, stream_);
        else
            for (int c = 0; c < R; c++) launch_copy(n_, src + c * ld, dst + c * n_, stream_);
    };
    in(Bu, ldb, PB);
    if (initial_solve) {
        launch_copy(nr, PB, PX, stream_);
        if (precond_block(PX, R)) return -1;
    } else {
        in(Xu, ldx, PX);
    }
    int it = 0;
    double res = 0;
    if (method == 1) {
        if (refine_cg_batch(PX, PB, R, segoff, itermax, eps, &it, &res)) return -1;
    } else if (method == 0 || method == 2 || method == 3) {
        if (segoff.size() != 2) return fail("GMRES, simple and BiCGSTAB refinement run on one system");
        // one right-hand side at a time, as the reference
        // (refinement/pastix_task_refine.c:157-161)
        for (int c = 0; c < R; c++) {
            int i2 = 0;
            double r2 = 0;
            double *xc = PX + c * n_;
            const double *bc = PB + c * n_;
            const int rc = (method == 0)   ? refine_gmres(xc, bc, itermax, im, eps, &i2, &r2)
                           : (method == 2) ? refine_sr(xc, bc, itermax, eps, &i2, &r2)
                                           : refine_bicgstab(xc, bc, itermax, eps, &i2, &r2);
            if (rc) return -1;
            it = std::max(it, i2);
            res = std::max(res, r2);
        }
    }
    if (user_order) launch_permute<double>(n_, R, d_perm_, false, PX, n_, Xu, ldx, stream_);
    else
        for (int c = 0; c < R; c++) launch_copy(n_, PX + c * n_, Xu + c * ldx, stream_);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    if (niter) *niter = it;
    if (resid) *resid = res;
    return 0;
}

// Backward errors of nrhs device columns in the user ordering into the device
// array dberr, with no allocation: the partial maxima go to the library-static
// scratch of this device (pastix_amd_backward_error_device_noalloc).
int Engine::berr_noalloc(const double *dx, int64_t ldx, const double *db, int64_t ldb, int nrhs, double normInf,
                         double *dberr) {
    if (!have_matrix_) return fail("matrix not uploaded");
    if (!d_iperm_) return fail("permutation not uploaded");
    if (nrhs < 1 || nrhs > NOALLOC_MAX_RHS) return fail("nrhs outside [1, 64]");
    double *scr = noalloc_scratch(dev_);
    if (!scr) return fail("no-allocation scratch not set up on this device");
    HIPCHK(hipSetDevice(dev_));
    launch_berr_user(n_, d_rowptr_, d_cols_, d_vals_, d_iperm_, dx, ldx, db, ldb, nrhs, normInf, scr, dberr,
                     stream_);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

int Engine::berr(const double *dx, const double *db, double out[3]) {
    if (!have_matrix_) return fail("matrix not uploaded");
    launch_berr_parts(n_, d_rowptr_, d_cols_, d_vals_, dx, db, d_red_, d_red_ + 512 * 32, stream_);
    HIPCHK(hipMemcpyAsync(out, d_red_ + 512 * 32, 3 * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return 0;
}

// Preconditioned conjugate gradient, reference refinement/z_refine_grad.c:46-156,
// on device vectors; M^{-1} is one supernodal solve in the factor precision.
int Engine::refine_cg(double *x, const double *b, int itermax, double eps, int *niter, double *resid) {
    if (vec_cap_ < 4 * n_) {
        vec_cap_ = 4 * n_;
        release(d_vec_);
        d_vec_ = (double *)alloc(vec_cap_ * sizeof(double));
        if (!d_svec_) d_svec_ = (float *)alloc(n_ * sizeof(float));
        if (!d_vec_ || !d_svec_) {
            vec_cap_ = 0;
            return -1;
        }
    }
    double *r = d_vec_, *p = d_vec_ + n_, *z = d_vec_ + 2 * n_, *Ap = d_vec_ + 3 * n_;
    const double normb = std::sqrt(dot(b, b));
    const double normx = std::sqrt(dot(x, x));
    launch_copy(n_, b, r, stream_);
    if (normx > 0.) spmv(-1.0, x, 1.0, r);
    double normr = std::sqrt(dot(r, r));
    double rb = (normb > 0) ? normr / normb : normr;
    int it = 0;
    if (rb > eps && it < itermax) {
        launch_copy(n_, r, z, stream_);
        if (precond(z)) return -1;
        launch_copy(n_, z, p, stream_);
    }
    while (rb > eps && it < itermax) {
        it++;
        spmv(1.0, p, 0.0, Ap);
        double rz = dot(r, z);
        double pAp = dot(Ap, p);
        double alpha = rz / pAp;
        launch_axpy(n_, alpha, p, x, stream_);
        launch_axpy(n_, -alpha, Ap, r, stream_);
        // the stopping test reads r only: taken before M^{-1} r, whose
        // result would be discarded when the loop ends here
        normr = std::sqrt(dot(r, r));
        rb = (normb > 0) ? normr / normb : normr;
        if (!(rb > eps && it < itermax)) break;
        launch_copy(n_, r, z, stream_);
        if (precond(z)) return -1;
        double rz2 = dot(r, z);
        double beta = rz2 / rz;
        launch_scal(n_, beta, p, stream_);
        launch_axpy(n_, 1.0, z, p, stream_);
    }
    HIPCHK(hipStreamSynchronize(stream_));
    if (niter) *niter = it;
    if (resid) *resid = rb;
    return 0;
}

// Simple iterative refinement, reference refinement/z_refine_pivot.c:26-119:
// r = b - A x and berr = ||r|| / ||b|| on every pass; while iter < itermax,
// berr > eps and berr has at least halved since the last correction (the first
// pass compares with 3 berr), x += M^{-1} r.  niter counts the passes, the
// last, stopping one included, as the reference's return value does.
int Engine::refine_sr(double *x, const double *b, int itermax, double eps, int *niter, double *resid) {
    if (vec_cap_ < 2 * n_) {
        vec_cap_ = 2 * n_;
        release(d_vec_);
        d_vec_ = (double *)alloc(vec_cap_ * sizeof(double));
        if (!d_svec_) d_svec_ = (float *)alloc(n_ * sizeof(float));
        if (!d_vec_ || !d_svec_) {
            vec_cap_ = 0;
            return -1;
        }
    }
    double *r = d_vec_, *dx = d_vec_ + n_;
    const double normb = std::sqrt(dot(b, b));
    double berr = 0, last = 0;
    int it = 0;
    for (;;) {
        launch_copy(n_, b, r, stream_);
        spmv(-1.0, x, 1.0, r);
        const double normr = std::sqrt(dot(r, r));
        berr = (normb > 0) ? normr / normb : normr;
        if (it == 0) last = 3 * berr;
        const bool go = it < itermax && berr > eps && berr <= last / 2.;
        it++;
        if (!go) break;
        launch_copy(n_, r, dx, stream_);
        if (precond(dx)) return -1;
        launch_axpy(n_, 1.0, dx, x, stream_);
        last = berr;
    }
    HIPCHK(hipStreamSynchronize(stream_));
    if (niter) *niter = it;
    if (resid) *resid = berr;
    return 0;
}

// Preconditioned BiCGSTAB, reference refinement/z_refine_bicgstab.c:25-185,
// step for step: y = M^{-1} p, v = A y, alpha = (r, r2) / (v, r2),
// s = r - alpha v, z = M^{-1} s, t = A z, and the weight
// w = (M^{-1} t, z) / (M^{-1} t, M^{-1} t) from preconditioned vectors (three
// supernodal solves per iteration); x += alpha y + w z, r = s - w t,
// beta = (r, r2) / (r_old, r2) * alpha / w, p = r + beta (p - w v).
int Engine::refine_bicgstab(double *x, const double *b, int itermax, double eps, int *niter, double *resid) {
    if (vec_cap_ < 9 * n_) {
        vec_cap_ = 9 * n_;
        release(d_vec_);
        d_vec_ = (double *)alloc(vec_cap_ * sizeof(double));
        if (!d_svec_) d_svec_ = (float *)alloc(n_ * sizeof(float));
        if (!d_vec_ || !d_svec_) {
            vec_cap_ = 0;
            return -1;
        }
    }
    double *r = d_vec_, *r2 = r + n_, *p = r2 + n_, *y = p + n_, *v = y + n_, *sv = v + n_, *z = sv + n_,
           *t = z + n_, *mt = t + n_;
    const double normb = std::sqrt(dot(b, b));
    const double normx = std::sqrt(dot(x, x));
    launch_copy(n_, b, r, stream_);
    if (normx > 0.) spmv(-1.0, x, 1.0, r);
    double rb = std::sqrt(dot(r, r));
    rb = (normb > 0) ? rb / normb : rb;
    launch_copy(n_, r, r2, stream_);
    launch_copy(n_, r, p, stream_);
    int it = 0;
    while (rb > eps && it < itermax) {
        it++;
        launch_copy(n_, p, y, stream_);
        if (precond(y)) return -1;
        spmv(1.0, y, 0.0, v);
        const double rr2 = dot(r, r2);
        const double alpha = rr2 / dot(v, r2);
        launch_copy(n_, r, sv, stream_);
        launch_axpy(n_, -alpha, v, sv, stream_);
        launch_copy(n_, sv, z, stream_);
        if (precond(z)) return -1;
        spmv(1.0, z, 0.0, t);
        launch_copy(n_, t, mt, stream_);
        if (precond(mt)) return -1;
        const double w = dot(mt, z) / dot(mt, mt);
        launch_axpy(n_, alpha, y, x, stream_);
        launch_axpy(n_, w, z, x, stream_);
        launch_copy(n_, sv, r, stream_);
        launch_axpy(n_, -w, t, r, stream_);
        const double beta = dot(r, r2) / rr2 * (alpha / w);
        launch_axpy(n_, -w, v, p, stream_);
        launch_scal(n_, beta, p, stream_);
        launch_axpy(n_, 1.0, r, p, stream_);
        const double nr = std::sqrt(dot(r, r));
        rb = (normb > 0) ? nr / normb : nr;
    }
    HIPCHK(hipStreamSynchronize(stream_));
    if (niter) *niter = it;
    if (resid) *resid = rb;
    return 0;
}

// Right-preconditioned restarted GMRES, reference
// refinement/z_refine_gmres.c:46-352 step for step: each cycle starts from the
// true residual and ends when it is small enough (:190-205); each step
// orthogonalises A M^{-1} v_i against v_0..v_i one after the other (dot, then
// axpy: :240-249), normalises only above 1e-50 (:252-260), applies the earlier
// Givens rotations and a new one whose denominator is raised to eps when it
// falls below (:263-289); the cycle ends at im steps, on the residual estimate
// or at itermax steps (:300-309), then x += W H^{-1} g (:338-348).  The
// projections and their coefficients stay on the device; the step's column of
// H crosses to the host once per step.  The reference takes the product A x
// of a cycle only when ||x_0|| > 0, tested once (:145,185); here it is taken
// with the current x on every cycle (the same first cycle, since b - A 0 = b).
int Engine::refine_gmres(double *x, const double *b, int itermax, int im, double eps, int *niter, double *resid) {
    const int64_t need = (int64_t)(2 * im + 1) * n_ + (im + 2);
    if (vec_cap_ < need) {
        vec_cap_ = need;
        release(d_vec_);
        d_vec_ = (double *)alloc(vec_cap_ * sizeof(double));
        if (!d_svec_) d_svec_ = (float *)alloc(n_ * sizeof(float));
        if (!d_vec_ || !d_svec_) {
            vec_cap_ = 0;
            return -1;
        }
    }
    double *V = d_vec_;                            // n x (im+1)
    double *W = d_vec_ + (int64_t)(im + 1) * n_;   // n x im: M^{-1} v_i
    double *hd = W + (int64_t)im * n_;             // im+2: column i of H, then y
    const int im1 = im + 1;
    std::vector<double> H((size_t)im1 * im, 0.0), cs(im), sn(im), g(im1), h(im1 + 1);
    const double normb = std::sqrt(dot(b, b));
    auto relres = [&](double r) { return (normb > 0) ? r / normb : r; };
    int iters = 0;
    double rb = 0;
    for (;;) {
        launch_copy(n_, b, V, stream_);
        spmv(-1.0, x, 1.0, V);
        double res = std::sqrt(dot(V, V));
        rb = relres(res);
        if (rb <= eps) break;
        launch_scal(n_, 1.0 / res, V, stream_);
        g[0] = res;
        int i = -1;
        for (;;) {
            i++;
            double *Hi = H.data() + (size_t)i * im1;
            double *vi = V + (int64_t)i * n_, *wi = W + (int64_t)i * n_, *vn = V + (int64_t)(i + 1) * n_;
            launch_copy(n_, vi, wi, stream_);
            if (precond(wi)) return -1;
            spmv(1.0, wi, 0.0, vn);
            for (int j = 0; j <= i; j++) {
                const double *vj = V + (int64_t)j * n_;
                launch_dot(n_, vn, vj, d_red_, hd + j, stream_);
                launch_axpy_dev(n_, hd + j, -1.0, vj, vn, stream_);
            }
            launch_dot(n_, vn, vn, d_red_, hd + i + 1, stream_);
            HIPCHK(hipMemcpyAsync(h.data(), hd, (i + 2) * sizeof(double), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipStreamSynchronize(stream_));
            for (int j = 0; j <= i; j++) Hi[j] = h[j];
            const double nrm = std::sqrt(h[i + 1]);
            Hi[i + 1] = nrm;
            if (nrm > 1e-50) launch_scal(n_, 1.0 / nrm, vn, stream_);
            for (int j = 0; j < i; j++) {
                const double t = Hi[j];
                Hi[j] = cs[j] * t + sn[j] * Hi[j + 1];
                Hi[j + 1] = cs[j] * Hi[j + 1] - sn[j] * t;
            }
            double t = std::sqrt(Hi[i] * Hi[i] + Hi[i + 1] * Hi[i + 1]);
            if (std::fabs(t) <= eps) t = eps;
            cs[i] = Hi[i] / t;
            sn[i] = Hi[i + 1] / t;
            g[i + 1] = -sn[i] * g[i];
            g[i] = cs[i] * g[i];
            Hi[i] = cs[i] * Hi[i] + sn[i] * Hi[i + 1];
            rb = relres(std::fabs(g[i + 1]));
            iters++;
            if (i + 1 >= im || rb <= eps || iters >= itermax) break;
        }
        // y = H^{-1} g (upper triangular, i+1 columns, ld im+1), x += W y
        for (int k = i; k >= 0; k--) {
            double v = g[k];
            for (int l = k + 1; l <= i; l++) v -= H[(size_t)k + (size_t)l * im1] * g[l];
            g[k] = v / H[(size_t)k + (size_t)k * im1];
        }
        HIPCHK(hipMemcpyAsync(hd, g.data(), (i + 1) * sizeof(double), hipMemcpyHostToDevice, stream_));
        launch_gemv_n(n_, i + 1, W, n_, hd, 1.0, x, stream_);
        if (iters >= itermax) break;
    }
    HIPCHK(hipStreamSynchronize(stream_));
    if (niter) *niter = iters;
    if (resid) *resid = rb;
    return 0;
}

}  // namespace pxa
