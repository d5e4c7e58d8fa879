// Device engine: owns the HBM-resident solver structure, factor and
// workspaces of one pastix instance and runs numfact / solve / SpMV / BLAS-1
// on one HIP stream.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "analysis.h"
#include "plan.h"

namespace pxa {

struct HostBcsc;

class Engine {
  public:
    Engine() = default;
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // factotype: 0 POTRF/LLT, 2 GETRF.  prec: 2 (float) or 3 (double) for the factor.
    int init(int device, const Symbol &s, const Layout &L, const Plan &P, int factotype, int prec);
    bool ready() const { return inited_; }
    int prec() const { return prec_; }
    int factotype() const { return facto_; }

    // Coefficient tables (1D layout, factor precision) to the host.
    int download_coef(void *hL, void *hU);
    // Entries [off, off + count) of the tables (U only for LU), factor precision.
    int download_coef_range(int64_t off, int64_t count, void *hL, void *hU);
    // Permuted columns [schur_fcol(), n) hold the Schur complement (none when
    // it equals n); schur_interface: the solve applies the updates of the
    // factored part to the Schur rows and reads them in the backward sweep
    // (PastixSolvModeInterface), else leaves them as they are (Local).
    int64_t schur_fcol() const { return schurfcol_; }
    bool schur_interface = false;
    // Diagonal of the factor's L panels (permuted order) to the host, FP64.
    int download_diag(double *h);
    // Pinned host staging block of at least count doubles, kept and grown
    // between calls (freed with the engine); nullptr when hipHostMalloc fails.
    double *host_block(int64_t count);
    // Numerical factorisation over all levels; seconds = device time.
    int factorize(double thr, int64_t *npiv, int *negdiag, double *seconds);
    // Solve in place on a device block (factor precision, permuted order).
    int solve_device(void *dx, int64_t ldx, int nrhs, double *seconds);
    // Host right-hand sides in the user's ordering (FP64 or factor precision).
    int solve_host(const int64_t *perm, int flttype, void *b, int64_t ldb, int nrhs, double *seconds);
    // One part of the solve on host right-hand sides already in the permuted
    // ordering: 1 forward sweep, 2 backward sweep, 3 backward sweep of LDL^t's
    // L^t (x *= D first), 4 division by the factor's diagonal.
    int solve_host_part(int part, int flttype, void *b, int64_t ldb, int nrhs, double *seconds);

    // Permuted matrix in CSR form for the residual SpMV (always FP64).
    int upload_matrix(const HostBcsc &bc);
    // Forget the uploaded CSR pattern and fill map (the input's pattern
    // changed); the next upload_matrix sends them again.
    void drop_matrix() {
        have_matrix_ = false;
        have_fillmap_ = false;
        have_transpose_ = false;
    }
    // Coefficient tables filled on the device from the uploaded CSR values:
    // cmap gives per CSR entry its position in the L table (>= 0), in the U
    // table (-2 - position) or -1 (outside the symbol).
    bool has_fillmap() const { return have_fillmap_; }
    void drop_fillmap() { have_fillmap_ = false; }
    int upload_fillmap(const std::vector<int64_t> &cmap);
    int fill_coef();
    // Device refinement drivers (vectors in the permuted ordering, FP64).
    int refine_cg(double *dx, const double *db, int itermax, double eps, int *niter, double *resid);
    int refine_gmres(double *dx, const double *db, int itermax, int im, double eps, int *niter, double *resid);
    int refine_sr(double *dx, const double *db, int itermax, double eps, int *niter, double *resid);
    int refine_bicgstab(double *dx, const double *db, int itermax, double eps, int *niter, double *resid);
    // Batched CG over independent systems (consecutive row segments segoff of
    // the permuted order) and R right-hand-side columns (n x R blocks, ld = n):
    // every (system, column) pair runs its own CG recurrence and stops at its
    // own tolerance.  niter: largest iteration count, resid: largest
    // ||r||/||b|| at exit.
    int refine_cg_batch(double *dX, const double *dB, int R, const std::vector<int64_t> &segoff, int itermax,
                        double eps, int *niter, double *resid);
    // FP64 device right-hand sides B and solutions X (ld >= n), in the user
    // ordering when user_order: X = M^{-1} B first if initial_solve (else X
    // holds the initial guesses), then refinement method 1 = batched CG over
    // the row segments segoff, 0 = GMRES, 2 = simple refinement, 3 = BiCGSTAB
    // (these three per column, one segment), -1 = none.
    int refine_user(const double *B, int64_t ldb, double *X, int64_t ldx, int R, const std::vector<int64_t> &segoff,
                    int method, bool initial_solve, bool user_order, int itermax, int im, double eps, int *niter,
                    double *resid);
    // Normwise backward error pieces: [max|b-Ax|, max|x|, max|b|].
    int berr(const double *dx, const double *db, double out[3]);
    int berr_noalloc(const double *dx, int64_t ldx, const double *db, int64_t ldb, int nrhs, double normInf,
                     double *dberr);

    hipStream_t stream() const { return stream_; }
    int64_t n() const { return n_; }
    void *alloc(size_t bytes);
    // frees a buffer from alloc() (no-op on nullptr) and sets it to nullptr
    template <typename P> void release(P *&p) {
        release_ptr((void *)p);
        p = nullptr;
    }
    void release_ptr(void *p);
    // Work blocks of the refinement (grown on demand, kept between calls):
    // the device bytes a call of refine_user on R columns of one system
    // needs, and freeing them (the next call allocates them again).
    int64_t work_bytes(int R) const;
    void release_work();
    int64_t device_bytes() const { return bytes_; }
    const std::string &last_error() const { return err_; }
    int device() const { return dev_; }
    // applies M^{-1} (one solve in factor precision) to an FP64 device vector
    int precond(double *dv);
    // Replace the matrix by A + delta*I for the next factorisation and for the
    // refinement products: restores the filled coefficients from the device
    // snapshot and adds delta to their diagonal and to the diagonal of the
    // FP64 matrix.  Needs snapshot_coef() and upload_matrix() first; every
    // diagonal entry must be stored.
    int shift_diagonal(double delta);
    double diag_sum() const { return diag_sum_; }
    double shift() const { return shift_; }
    // y = alpha*A*x + beta*y on FP64 device vectors
    int spmv(double alpha, const double *x, double beta, double *y);
    // y = alpha*A^T*x + beta*y; needs upload_transpose() for the current pattern
    int spmv_t(double alpha, const double *x, double beta, double *y);
    bool has_transpose() const { return have_transpose_; }
    // transposed pattern of the uploaded CSR (reads the host copy once per pattern)
    int upload_transpose(const HostBcsc &bc);
    double dot(const double *x, const double *y);
    // BLAS-1 / BLAS-2 on FP64 device vectors of any length, on the engine's
    // stream; results of the reductions to the host
    int vec_dot(int64_t n, const double *x, const double *y, double *out);
    int vec_nrm2(int64_t n, const double *x, double *out);
    int vec_axpy(int64_t n, double alpha, const double *x, double *y);
    int vec_scal(int64_t n, double alpha, double *x);
    int vec_copy(int64_t n, const double *x, double *y);
    int vec_gemv(int64_t m, int64_t n, double alpha, const double *A, int64_t lda, const double *x, double beta,
                 double *y);

    // Keep a device copy of the freshly filled coefficients so that the
    // factorisation can be re-run from HBM-resident input (restore_coef).
    int snapshot_coef();
    int restore_coef();
    // Solve with device right-hand sides in the user ordering.
    int solve_user_device(void *db, int64_t ldb, int nrhs, double *seconds);
    int upload_perm(const std::vector<int64_t> &permtab);

    // Device time of the last factorize (seconds) per kernel class when
    // profiling is on, with the number of timed spans of each class.  The
    // update classes time one launch per span; the panel class times the
    // panel steps of one level (diagonal blocks, panel TRSM, trailing panel
    // update) as one span.
    enum KClass { KC_PANEL = 0, KC_DIRECT, KC_WGEMM, KC_GATHER, KC_ACC, KC_LEAF, KC_N };
    bool profile = false;
    // Replay numfact and device solves from captured HIP graphs (one graph
    // launch per factorisation / solve instead of one launch per kernel).
    bool use_graph = false;
    // Overlap each level's updates into later levels with the next level's
    // panel steps (second stream); off: one stream, level by level.
    bool overlap = true;
    // Factor the leaf region (LL^t) in one persistent launch driven by
    // device-side dependency counters; off: level by level like the rest.
    bool use_leaf = true;
    int leaf_grid = 0;  // persistent workgroups of that launch (0: LEAF_GRID)
    static constexpr int LEAF_GRID = 512;

    bool leaf_active() const { return use_leaf && facto_ != 2 && leaf_.nregion > 0; }
    double kt[KC_N] = {0, 0, 0, 0, 0, 0};
    int kn[KC_N] = {0, 0, 0, 0, 0, 0};
    // per dependency level of the last profiled factorisation: panel steps
    // and Schur updates (seconds; the leaf launch counts as level 0's panel)
    std::vector<double> lv_panel_s, lv_upd_s;
    double t_diag = 0, t_upd = 0;
    int n_diag = 0, n_upd = 0;

  private:
    template <typename T> int factorize_t(double thr);
    template <typename T> int solve_t(T *dx, int64_t ldx, int nrhs);
    int fail(const std::string &m);

    bool inited_ = false;
    int dev_ = 0, facto_ = 0, prec_ = 3;
    hipStream_t stream_ = nullptr;
    // numfact issues the updates that the next level's panels do not wait
    // for on a second stream (joined back before the level after)
    hipStream_t side_ = nullptr;
    // the solve nodes' inverse merges, beside the factorisation
    hipStream_t winv_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr, ev_panel_ = nullptr, ev_side_ = nullptr, ev_winv_ = nullptr;
    struct Owned {
        void *p;
        size_t bytes;
    };
    std::vector<Owned> owned_;
    int64_t bytes_ = 0;

    std::string err_;

    int64_t n_ = 0, cblknbr_ = 0, coefnbr_ = 0, winvnbr_ = 0;
    int maxwidth_ = 0;
    // solve levels: widest node (diagonal kernels), widest cblk (products)
    std::vector<int> slv_maxw_, slv_maxcw_;
    int64_t schurfcol_ = 0;
    std::vector<int32_t> lv_fact_cnt_;
    int32_t *d_fact_cblks_ = nullptr;
    int n_fact_ = 0;
    DevSymbol S_{};
    // host copies of the level pointers
    std::vector<int32_t> lv_cblk_ptr_, lv_upd_ptr_, lv_upd_split_, slv_fwd_ptr_, slv_bwd_ptr_;
    std::vector<int32_t> lv_upd_dnext_, lv_upd_mnext_, lv_wt_next_;
    // device plan
    int32_t *d_lv_cblks_ = nullptr, *d_shift_ = nullptr;
    YSlices ys_{};  // update-buffer slices
    int32_t *d_ydst_ = nullptr;
    int64_t nyrows_ = 0;
    void *d_ybuf_ = nullptr, *d_WB_ = nullptr;
    std::vector<int32_t> lv_step_ptr_, step_cb_ptr_, step_trsm_ptr_, step_upd_ptr_;
    std::vector<uint8_t> lv_panel_right_;
    int32_t *d_step_cblks_ = nullptr;
    // inverses of the solve nodes' diagonal blocks: leaves of the nodes of
    // several cblks, merge-round tiles in batches, the rounds' product scratch
    WinvLeaf *d_winv_leaves_ = nullptr;
    int n_winv_leaves_ = 0;
    WinvTile *d_winv_tiles_ = nullptr;
    std::vector<int32_t> winv_round_ptr_, winv_batch_lv_, winv_batch_leaf_, winv_batch_round_;
    void *d_tbuf_ = nullptr;
    PanelItem *d_ptrsm_ = nullptr, *d_pupd_ = nullptr;
    UpdItem *d_upd_ = nullptr;
    UpdEntry *d_ents_ = nullptr;
    // two-phase multi-source update: product groups, tiles, level workspace
    WGroup *d_wgrp_ = nullptr;
    WTile *d_wtiles_ = nullptr;
    std::vector<int32_t> lv_wt_ptr_;
    std::vector<int64_t> lv_wsize_;
    std::vector<uint8_t> lv_single_pass_;
    void *d_wspace_ = nullptr;
    int64_t wcap_ = 0;  // workspace entries
    // leaf region (LL^t): task lists, counters, ready queue, the updates it
    // sends above itself
    LeafDev leaf_{};
    int32_t leaf_levels_ = 0;
    long long leaf_maxwait_ = 0;  // wall-clock ticks of one wait in that launch (two seconds)
    int32_t *d_leaf_ctr_ = nullptr, *d_leaf_queue_ = nullptr, *d_leaf_sync_ = nullptr;
    UpdItem *d_leaf_out_ = nullptr;
    int32_t n_leaf_out_ = 0, leaf_out_split_ = 0;
    FwdItem *d_fwd_ = nullptr;
    DiagItem *d_diag_ = nullptr;
    std::vector<int32_t> slv_diag_ptr_;
    void *d_z_ = nullptr;  // forward-sweep result, n x nrhs_chunk
    BwdItem *d_bwd_ = nullptr;
    int32_t *d_perm_ = nullptr;  // permtab as int32 for device permutations
    int32_t *d_iperm_ = nullptr; // peritab as int32
    void *d_L0_ = nullptr, *d_U0_ = nullptr;  // snapshot of the filled coefficients
    std::vector<hipEvent_t> pev_;
    // captured graphs: numfact (re-captured when the pivot threshold changes)
    // and solves keyed by (right-hand-side pointer, leading dimension, nrhs,
    // user ordering)
    hipGraphExec_t fact_exec_ = nullptr;
    double fact_thr_ = 0;
    bool fact_overlap_ = true;
    int fact_leaf_ = -1;
    struct SolveGraph {
        const void *ptr;
        int64_t ld;
        int nrhs;
        bool user, iface;
        hipGraphExec_t exec;
    };
    std::vector<SolveGraph> solve_graphs_;
    int enqueue_factorize(double thr);
    int enqueue_solve(void *dx, int64_t ldx, int nrhs, bool user);
    hipGraphExec_t solve_graph(void *dx, int64_t ldx, int nrhs, bool user);
    struct Span {
        int cls, e0, e1, lv;
    };
    int cur_lv_ = -1;  // level whose launches are being issued (per-level profiles)
    std::vector<Span> spans_;
    int nev_ = 0;
    int prof_begin(hipStream_t st);
    void prof_end(int cls, int k, hipStream_t st);
    int64_t bwd_part_ = 0;
    // numerics
    void *d_L_ = nullptr, *d_U_ = nullptr, *d_W1_ = nullptr, *d_W2_ = nullptr;
    int *d_stat_ = nullptr;
    void *d_part_ = nullptr;
    int nrhs_chunk_ = 8;
    void *d_rhs_ = nullptr;  // n x nrhs_cap work block in factor precision
    int64_t rhs_cap_ = 0;
    double *d_red_ = nullptr;  // reduction partials + scalars
    // matrix (CSR of the permuted matrix), FP64
    int64_t *d_rowptr_ = nullptr;
    int32_t *d_cols_ = nullptr;
    double *d_vals_ = nullptr;
    int64_t csr_nnz_ = 0;
    bool have_matrix_ = false;
    // transposed pattern: row pointers, columns and CSR positions of A^T
    int64_t *d_trowptr_ = nullptr, *d_tpos_ = nullptr;
    int32_t *d_tcols_ = nullptr;
    bool have_transpose_ = false;
    int64_t *d_cmap_ = nullptr;
    double *d_diagout_ = nullptr;
    double *h_pin_ = nullptr;
    int64_t h_pin_cap_ = 0;
    bool have_fillmap_ = false;
    // diagonal positions: in the coefficient table (per permuted column) and
    // in the CSR values, with the original FP64 diagonal
    int64_t *d_diagpos_ = nullptr, *d_csrdiag_ = nullptr;
    std::vector<int64_t> csrdiag_;  // host copy of the CSR diagonal positions (-1: absent)
    double *d_diag0_ = nullptr;
    double diag_sum_ = 0;
    bool diag_complete_ = false;
    double shift_ = 0;  // delta of the last shift_diagonal
    // refinement vectors
    double *d_vec_ = nullptr;
    int64_t vec_cap_ = 0;
    float *d_svec_ = nullptr;
    // batched refinement work space
    double *d_bvec_ = nullptr, *d_bscal_ = nullptr;
    float *d_bsvec_ = nullptr;
    int64_t bvec_cap_ = 0, bscal_cap_ = 0, bsvec_cap_ = 0, segoff_dcap_ = 0;
    double *d_sr_ = nullptr;  // permuted right-hand sides and solutions of refine_user
    int64_t sr_cap_ = 0;
    int precond_block(double *zb, int R);
    int part_ = 0;  // solve part of the next solve_t (0: whole solve)
    int32_t *d_seg_ = nullptr;
    int64_t *d_segoff_ = nullptr;
    // chunks of the segmented dot products: offsets per segment, then the
    // segment of every chunk; first-pass sums
    int32_t *d_coff_ = nullptr, *d_cseg_ = nullptr;
    int nchunk_ = 0;
    int64_t chunk_cap_ = 0, cpart_cap_ = 0;
    double *d_cpart_ = nullptr;
    std::vector<int64_t> segoff_host_;
};

// Host-side permuted matrix (reference bcsc/bcsc.h:28-47 keeps the same data
// grouped by cblk): rowptr/cols/tvals, the CSR of B = P A P^T with the full
// pattern (duplicates summed in input order).
struct HostBcsc {
    int64_t n = 0;
    int mtxtype = 0;
    std::vector<int64_t> rowptr;
    std::vector<int32_t> cols;
    std::vector<double> tvals;
    double normF = 0, normInf = 0;
    // A later input with the same pattern is refilled by one gather: the
    // input's format, type, size and index arrays as given, and per CSR
    // entry e the input value indices src[sptr[e] .. sptr[e+1]) summed into
    // it, in input order.
    int in_fmt = -1, in_mtx = -1;
    int64_t in_n = -1, in_nnz = -1;
    std::vector<int64_t> in_colptr, in_rowptr;
    std::vector<int64_t> sptr, src;
    // coefficient-table position of every CSR entry (Engine::upload_fillmap)
    // for the factorisation kind cmap_facto; -1: not built
    std::vector<int64_t> cmap;
    int cmap_facto = -1;
};

}  // namespace pxa
