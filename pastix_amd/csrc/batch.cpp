// Batches of independent systems.
//
// The reference solves one matrix per pastix_data_t.  For many small systems
// this engine analyses their block-diagonal union once: every dependency
// level of the union holds the supernodes of all systems at that depth, so
// one launch per kernel and level factors (and solves) the whole batch, and
// the captured-graph mode turns numfact and each solve into a single graph
// launch.  System s owns rows and columns [off[s], off[s+1]) of the union in
// the user numbering, and the same range in the permuted numbering since the
// union ordering is the concatenation of the systems' orderings.
#include "pastix_data.h"

using namespace pxa_api;

namespace {

int publish_union(pastix_data_t *d, int mtx, int64_t N, const std::vector<int64_t> &cp,
                  const std::vector<int32_t> &rows, std::vector<double> &vals);

// The union by concatenation, when every system is a CSC whose rows
// increase strictly within each column (lower-triangle entries only when
// symmetric): that is already the union's sorted, duplicate-free CSC.
bool concat_union(pastix_int_t nsys, const spmatrix_t *const *spms, bool values, std::vector<int64_t> &off,
                  std::vector<int64_t> &cp, std::vector<int32_t> &rows, std::vector<double> &vals) {
    const int mtx = spms[0]->mtxtype;
    int64_t nnz = 0;
    for (pastix_int_t s = 0; s < nsys; s++) {
        const spmatrix_t *A = spms[s];
        if (A->fmttype != SpmCSC) return false;
        const spm_int_t *c = A->colptr, *r = A->rowptr;
        const int64_t base = c[0];
        for (int64_t j = 0; j < A->n; j++)
            for (int64_t k = c[j] - base; k < c[j + 1] - base; k++) {
                const int64_t i = r[k] - base;
                if (i < 0 || i >= A->n || (k > c[j] - base && i <= r[k - 1] - base) || (mtx != SpmGeneral && i < j))
                    return false;
            }
        nnz += c[A->n] - base;
        off[s + 1] = off[s] + A->n;
    }
    cp.assign(1, 0);
    cp.reserve(off[nsys] + 1);
    rows.clear();
    rows.reserve(nnz);
    vals.clear();
    vals.reserve(nnz);
    for (pastix_int_t s = 0; s < nsys; s++) {
        const spmatrix_t *A = spms[s];
        const spm_int_t *c = A->colptr, *r = A->rowptr;
        const int64_t base = c[0], e = c[A->n] - base, r0 = (int64_t)rows.size();
        for (int64_t j = 0; j < A->n; j++) cp.push_back(r0 + c[j + 1] - base);
        for (int64_t k = 0; k < e; k++) rows.push_back((int32_t)(r[k] - base + off[s]));
        if (!values || A->flttype == SpmPattern || !A->values) vals.insert(vals.end(), e, 1.0);
        else if (A->flttype == SpmDouble) vals.insert(vals.end(), (const double *)A->values, (const double *)A->values + e);
        else for (int64_t k = 0; k < e; k++) vals.push_back((double)((const float *)A->values)[k]);
    }
    return true;
}

// Union CSC of the systems (lower triangle when symmetric, FP64 values).
int build_union(pastix_data_t *d, pastix_int_t nsys, const spmatrix_t *const *spms, bool values,
                std::vector<int64_t> &off) {
    off.assign(nsys + 1, 0);
    const int mtx = spms[0]->mtxtype;
    for (pastix_int_t s = 0; s < nsys; s++) {
        const spmatrix_t *A = spms[s];
        if (!A || A->mtxtype != mtx || A->n <= 0 || A->dof != 1) return PASTIX_ERR_BADPARAMETER;
        if (values && A->flttype != SpmDouble && A->flttype != SpmFloat) return PASTIX_ERR_BADPARAMETER;
    }
    std::vector<int64_t> cp;
    std::vector<int32_t> rows;
    std::vector<double> vals;
    if (concat_union(nsys, spms, values, off, cp, rows, vals)) return publish_union(d, mtx, off[nsys], cp, rows, vals);
    Triplets u;
    for (pastix_int_t s = 0; s < nsys; s++) {
        const spmatrix_t *A = spms[s];
        if (!A || A->mtxtype != mtx || A->n <= 0 || A->dof != 1) return PASTIX_ERR_BADPARAMETER;
        if (values && A->flttype != SpmDouble && A->flttype != SpmFloat) return PASTIX_ERR_BADPARAMETER;
        Triplets t;
        const int rc = spm_triplets(A, t, values);
        if (rc) return rc;
        for (size_t k = 0; k < t.i.size(); k++) {
            if (mtx != SpmGeneral && t.i[k] < t.j[k]) continue;
            u.i.push_back(t.i[k] + off[s]);
            u.j.push_back(t.j[k] + off[s]);
            if (values) u.v.push_back(t.v[k]);
        }
        off[s + 1] = off[s] + A->n;
    }
    const int64_t N = off[nsys];
    triplets_to_csc(N, u, nullptr, cp, rows, vals, false);
    if (!values) vals.assign(rows.size(), 1.0);
    return publish_union(d, mtx, N, cp, rows, vals);
}

int publish_union(pastix_data_t *d, int mtx, int64_t N, const std::vector<int64_t> &cp,
                  const std::vector<int32_t> &rows, std::vector<double> &vals) {
    d->u_colptr.assign(cp.begin(), cp.end());
    d->u_rows.assign(rows.begin(), rows.end());
    d->u_vals.swap(vals);
    spmatrix_t &S = d->u_spm;
    memset(&S, 0, sizeof(S));
    S.mtxtype = (spm_mtxtype_t)mtx;
    S.flttype = SpmDouble;
    S.fmttype = SpmCSC;
    S.gN = S.n = S.gNexp = S.nexp = N;
    S.gnnz = S.nnz = S.gnnzexp = S.nnzexp = (spm_int_t)d->u_rows.size();
    S.dof = 1;
    S.layout = SpmColMajor;
    S.colptr = d->u_colptr.data();
    S.rowptr = d->u_rows.data();
    S.values = d->u_vals.data();
    return PASTIX_SUCCESS;
}

// Ordering of one system: the caller's, or nested dissection of its graph.
int system_order(const spmatrix_t *A, const pastix_order_t *o, pxa::Order &out) {
    const int64_t n = A->n;
    out = pxa::Order();
    out.n = n;
    if (o && (o->permtab || o->peritab)) {
        const int64_t base = o->baseval;
        out.permtab.resize(n);
        out.peritab.resize(n);
        if (o->permtab) {
            for (int64_t i = 0; i < n; i++) out.permtab[i] = o->permtab[i] - base;
            for (int64_t i = 0; i < n; i++) out.peritab[out.permtab[i]] = i;
        } else {
            for (int64_t i = 0; i < n; i++) out.peritab[i] = o->peritab[i] - base;
            for (int64_t i = 0; i < n; i++) out.permtab[out.peritab[i]] = i;
        }
        if (o->rangtab && o->treetab && o->cblknbr > 0) {
            out.cblknbr = o->cblknbr;
            out.rangtab.resize(out.cblknbr + 1);
            out.treetab.resize(out.cblknbr);
            for (int64_t c = 0; c <= out.cblknbr; c++) out.rangtab[c] = o->rangtab[c] - base;
            for (int64_t c = 0; c < out.cblknbr; c++) out.treetab[c] = o->treetab[c] < 0 ? -1 : o->treetab[c] - base;
        }
    } else {
        pxa::Graph g;
        const int rc = spm_graph(A, g);
        if (rc) return rc;
        out = pxa::order_nested_dissection(g);
    }
    std::vector<char> seen(n, 0);
    for (int64_t i = 0; i < n; i++) {
        const int64_t p = out.permtab[i];
        if (p < 0 || p >= n || seen[p]) return PASTIX_ERR_BADPARAMETER;
        seen[p] = 1;
    }
    std::string err;
    if (out.cblknbr && !pxa::check_order(out, err)) return PASTIX_ERR_BADPARAMETER;
    // without a supernode partition, supernodes are found on the union graph
    // (they never span two systems: the union's elimination forest has one
    // tree per system at least)
    return PASTIX_SUCCESS;
}

// Column-major n x R block of the union from per-system blocks (FP64).
// B (N x R, ld N): the systems' columns, zeros past each system's nrhs
void pack(const std::vector<int64_t> &off, pastix_int_t R, const pastix_int_t *nrhs, const void *const *b,
          const pastix_int_t *ldb, double *B) {
    const int64_t N = off.back();
    for (size_t s = 0; s + 1 < off.size(); s++) {
        const double *bs = (const double *)b[s];
        const int64_t ns = off[s + 1] - off[s];
        for (pastix_int_t c = 0; c < R; c++) {
            double *col = B + off[s] + (size_t)c * N;
            if (c < nrhs[s]) memcpy(col, bs + c * ldb[s], ns * sizeof(double));
            else memset(col, 0, ns * sizeof(double));
        }
    }
}

void unpack(const std::vector<int64_t> &off, const pastix_int_t *nrhs, const double *B, void *const *b,
            const pastix_int_t *ldb) {
    const int64_t N = off.back();
    for (size_t s = 0; s + 1 < off.size(); s++) {
        double *bs = (double *)b[s];
        const int64_t ns = off[s + 1] - off[s];
        for (pastix_int_t c = 0; c < nrhs[s]; c++) memcpy(bs + c * ldb[s], B + off[s] + (size_t)c * N, ns * sizeof(double));
    }
}

// pinned host staging for the two N x R blocks of a batched call
double *staging(pastix_data_t *d, pastix_int_t R) {
    return d->engine ? d->engine->host_block(2 * d->batch_off.back() * (int64_t)R) : nullptr;
}

bool rhs_ok(const pastix_data_t *d, const pastix_int_t *nrhs, const void *const *b, const pastix_int_t *ldb,
            pastix_int_t &R) {
    R = 0;
    for (size_t s = 0; s + 1 < d->batch_off.size(); s++) {
        if (nrhs[s] < 0 || (nrhs[s] > 0 && (!b[s] || ldb[s] < d->batch_off[s + 1] - d->batch_off[s]))) return false;
        R = std::max(R, nrhs[s]);
    }
    return true;
}

}  // namespace

// Batched CG on n x R blocks in the user numbering (column-major, ld = n):
// permutes into the factor's numbering, refines on the device with one
// recurrence per (segment, column), permutes back.
double *pxa_api::device_block(pastix_data_t *d, int64_t count) {
    pxa::Engine &E = *d->engine;
    if (d->batch_dev_owner != &E || d->batch_dev_cap < count) {
        if (d->batch_dev_owner == &E) E.release(d->batch_dev);
        d->batch_dev = (double *)E.alloc((size_t)count * sizeof(double));
        d->batch_dev_cap = d->batch_dev ? count : 0;
        d->batch_dev_owner = &E;
    }
    return d->batch_dev;
}

void pxa_api::release_device_block(pastix_data_t *d) {
    if (d->engine && d->batch_dev_owner == d->engine.get()) d->engine->release(d->batch_dev);
    d->batch_dev = nullptr;
    d->batch_dev_cap = 0;
}

int pxa_api::refine_method(const pastix_data_t *d) {
    switch (d->iparm[IPARM_REFINEMENT]) {
    case PastixRefineCG: return 1;
    case PastixRefineSR: return 2;
    case PastixRefineBiCGSTAB: return 3;
    default: return 0;  // GMRES, as the reference's refinement table (pastix_task_refine.c:36-66)
    }
}

double pxa_api::refine_eps(const pastix_data_t *d) {
    const double eps = d->dparm[DPARM_EPSILON_REFINEMENT];
    return eps < 0. ? ((d->flttype == SpmFloat) ? 1e-6 : 1e-12) : eps;
}

int pxa_api::refine_block(pastix_data_t *d, int R, const double *B, double *X, const std::vector<int64_t> &segoff,
                          bool initial_solve, int method) {
    if (!(d->steps & STEP_NUMFACT) || !d->engine) return PASTIX_ERR_BADPARAMETER;
    const int64_t N = d->n;
    const size_t bytes = (size_t)N * R * sizeof(double);
    double *dB = device_block(d, 2 * N * R);
    if (!dB) return PASTIX_ERR_OUTOFMEMORY;
    double *dX = dB + (size_t)N * R;
    if (hipMemcpy(dB, B, bytes, hipMemcpyHostToDevice) != hipSuccess) return PASTIX_ERR_INTERNAL;
    if (!initial_solve && hipMemcpy(dX, X, bytes, hipMemcpyHostToDevice) != hipSuccess) return PASTIX_ERR_INTERNAL;
    int it = 0;
    double res = 0;
    if (d->engine->refine_user(dB, N, dX, N, R, segoff, method, initial_solve, true, (int)d->iparm[IPARM_ITERMAX], 1,
                               refine_eps(d), &it, &res))
        return PASTIX_ERR_INTERNAL;
    if (hipMemcpy(X, dX, bytes, hipMemcpyDeviceToHost) != hipSuccess) return PASTIX_ERR_INTERNAL;
    if (method >= 0) {
        d->iparm[IPARM_NBITER] = it;
        d->dparm[DPARM_RELATIVE_ERROR] = res;
    }
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_batch_analyze(pastix_data_t *d, pastix_int_t nsys, const spmatrix_t *const *spms,
                                        const pastix_order_t *const *orders) {
    if (!d || nsys < 1 || !spms) return PASTIX_ERR_BADPARAMETER;
    std::vector<int64_t> off;
    int rc = build_union(d, nsys, spms, true, off);
    if (rc) return rc;
    // concatenated ordering
    const int64_t N = off[nsys];
    std::vector<pastix_int_t> perm(N), rang, tree;
    bool with_tree = true;
    std::vector<pxa::Order> os(nsys);
    for (pastix_int_t s = 0; s < nsys; s++) {
        rc = system_order(spms[s], orders ? orders[s] : nullptr, os[s]);
        if (rc) return rc;
        with_tree = with_tree && os[s].cblknbr > 0;
        // each system's level order, as its own analysis would apply it; the
        // union's would interleave the systems' supernodes
        if (d->ext[PXA_IPARM_LEVEL_ORDER] && os[s].cblknbr > 0)
            pxa::order_apply_level_order(os[s], d->iparm[IPARM_TASKS2D_LEVEL], d->iparm[IPARM_TASKS2D_WIDTH]);
    }
    rang.push_back(0);
    for (pastix_int_t s = 0; s < nsys; s++) {
        for (int64_t i = 0; i < spms[s]->n; i++) perm[off[s] + i] = off[s] + os[s].permtab[i];
        if (with_tree) {
            const pastix_int_t c0 = (pastix_int_t)tree.size();
            for (int64_t c = 0; c < os[s].cblknbr; c++) {
                rang.push_back(off[s] + os[s].rangtab[c + 1]);
                tree.push_back(os[s].treetab[c] < 0 ? -1 : c0 + os[s].treetab[c]);
            }
        }
    }
    pastix_order_t uo;
    memset(&uo, 0, sizeof(uo));
    uo.baseval = 0;
    uo.vertnbr = N;
    uo.permtab = perm.data();
    if (with_tree) {
        uo.cblknbr = (pastix_int_t)tree.size();
        uo.rangtab = rang.data();
        uo.treetab = tree.data();
    }
    const pastix_int_t saved = d->iparm[IPARM_ORDERING], saved_lo = d->ext[PXA_IPARM_LEVEL_ORDER];
    d->iparm[IPARM_ORDERING] = PastixOrderPersonal;
    d->ext[PXA_IPARM_LEVEL_ORDER] = 0;
    rc = pastix_subtask_order(d, &d->u_spm, &uo);
    d->iparm[IPARM_ORDERING] = saved;
    d->ext[PXA_IPARM_LEVEL_ORDER] = saved_lo;
    if (rc) return rc;
    if ((rc = pastix_subtask_symbfact(d)) || (rc = pastix_subtask_reordering(d)) || (rc = pastix_subtask_blend(d)))
        return rc;
    // the union ordering keeps every system in its own range
    for (pastix_int_t s = 0; s < nsys; s++)
        for (int64_t i = off[s]; i < off[s + 1]; i++)
            if (d->order.permtab[i] < off[s] || d->order.permtab[i] >= off[s + 1]) {
                err_print("pastix_amd_batch_analyze", "ordering mixes systems");
                return PASTIX_ERR_INTERNAL;
            }
    d->batch_off = off;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_batch_numfact(pastix_data_t *d, const spmatrix_t *const *spms) {
    if (!d || !spms || d->batch_off.size() < 2) return PASTIX_ERR_BADPARAMETER;
    const pastix_int_t nsys = (pastix_int_t)d->batch_off.size() - 1;
    const size_t nnz = d->u_rows.size();
    std::vector<int64_t> off;
    int rc = build_union(d, nsys, spms, true, off);
    if (rc) return rc;
    if (off != d->batch_off || d->u_rows.size() != nnz) {
        err_print("pastix_amd_batch_numfact", "the systems differ from the analysed batch");
        return PASTIX_ERR_BADPARAMETER;
    }
    return pastix_task_numfact(d, &d->u_spm);
}

extern "C" int pastix_amd_batch_solve(pastix_data_t *d, const pastix_int_t *nrhs, void *const *b,
                                      const pastix_int_t *ldb) {
    if (!d || !nrhs || !b || !ldb || d->batch_off.size() < 2) return PASTIX_ERR_BADPARAMETER;
    pastix_int_t R = 0;
    if (!rhs_ok(d, nrhs, b, ldb, R)) return PASTIX_ERR_BADPARAMETER;
    if (R == 0) return PASTIX_SUCCESS;
    if (!(d->steps & STEP_NUMFACT) || !d->engine) return PASTIX_ERR_BADPARAMETER;
    double *B = staging(d, R);
    if (!B) return PASTIX_ERR_OUTOFMEMORY;
    double *X = B + d->batch_off.back() * (size_t)R;
    pack(d->batch_off, R, nrhs, b, ldb, B);
    // one solve in the factor precision, permutations and conversions on the device
    const int rc = refine_block(d, (int)R, B, X, d->batch_off, true, -1);
    if (rc) return rc;
    unpack(d->batch_off, nrhs, X, b, ldb);
    d->steps |= STEP_SOLVE;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_batch_refine(pastix_data_t *d, const pastix_int_t *nrhs, void *const *b,
                                       const pastix_int_t *ldb, void *const *x, const pastix_int_t *ldx) {
    if (!d || !nrhs || !b || !ldb || !x || !ldx || d->batch_off.size() < 2) return PASTIX_ERR_BADPARAMETER;
    if (!(d->steps & STEP_NUMFACT) || !d->engine) return PASTIX_ERR_BADPARAMETER;
    if (d->iparm[IPARM_REFINEMENT] != PastixRefineCG) {
        err_print("pastix_amd_batch_refine", "batched refinement is conjugate gradient (IPARM_REFINEMENT = PastixRefineCG)");
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
    pastix_int_t R = 0, Rx = 0;
    if (!rhs_ok(d, nrhs, b, ldb, R) || !rhs_ok(d, nrhs, x, ldx, Rx)) return PASTIX_ERR_BADPARAMETER;
    if (R == 0) return PASTIX_SUCCESS;
    const auto t0 = std::chrono::steady_clock::now();
    double *B = staging(d, R);
    if (!B) return PASTIX_ERR_OUTOFMEMORY;
    double *X = B + d->batch_off.back() * (size_t)R;
    pack(d->batch_off, R, nrhs, b, ldb, B);
    pack(d->batch_off, R, nrhs, x, ldx, X);
    const int rc = refine_block(d, (int)R, B, X, d->batch_off, false, 1);
    if (rc) return rc;
    unpack(d->batch_off, nrhs, X, x, ldx);
    d->dparm[DPARM_REFINE_TIME] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_batch_solve_refine(pastix_data_t *d, const pastix_int_t *nrhs, void *const *b,
                                             const pastix_int_t *ldb, void *const *x, const pastix_int_t *ldx) {
    if (!d || !nrhs || !b || !ldb || !x || !ldx || d->batch_off.size() < 2) return PASTIX_ERR_BADPARAMETER;
    if (!(d->steps & STEP_NUMFACT) || !d->engine) return PASTIX_ERR_BADPARAMETER;
    if (d->iparm[IPARM_REFINEMENT] != PastixRefineCG) {
        err_print("pastix_amd_batch_solve_refine", "batched refinement is conjugate gradient (IPARM_REFINEMENT = PastixRefineCG)");
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
    pastix_int_t R = 0, Rx = 0;
    if (!rhs_ok(d, nrhs, b, ldb, R) || !rhs_ok(d, nrhs, x, ldx, Rx)) return PASTIX_ERR_BADPARAMETER;
    if (R == 0) return PASTIX_SUCCESS;
    const auto t0 = std::chrono::steady_clock::now();
    double *B = staging(d, R);
    if (!B) return PASTIX_ERR_OUTOFMEMORY;
    double *X = B + d->batch_off.back() * (size_t)R;
    pack(d->batch_off, R, nrhs, b, ldb, B);
    const int rc = refine_block(d, (int)R, B, X, d->batch_off, true, 1);
    if (rc) return rc;
    unpack(d->batch_off, nrhs, X, x, ldx);
    d->steps |= STEP_SOLVE;
    d->dparm[DPARM_REFINE_TIME] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_batch_offsets(const pastix_data_t *d, int64_t *off) {
    if (!d || d->batch_off.size() < 2) return PASTIX_ERR_BADPARAMETER;
    if (off)
        for (size_t s = 0; s < d->batch_off.size(); s++) off[s] = d->batch_off[s];
    return (int)d->batch_off.size() - 1;
}
