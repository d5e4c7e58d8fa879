// PaStiX C API on top of the host analysis and the device engine.
//
// Task/subtask functions keep the reference contracts: the steps bitmask and
// its prerequisite checks (common/pastixdata.h:25-33,
// sopalin/pastix_task_sopalin.c:90-101,206-217,367-388), iparm/dparm inputs and
// outputs, in-place right-hand sides, and PASTIX_ERR_* return codes.
#include "pastix_data.h"
#include "kernels.h"
#include <algorithm>
#include <new>
#include <stdexcept>
#include <atomic>
#include <functional>
#include <thread>

namespace pxa_api {
double *device_pair(pastix_data_t *d) {
    if (d->dvec && d->dvec_owner == d->engine.get()) return d->dvec;
    d->dvec = (double *)d->engine->alloc(2 * d->n * sizeof(double));
    d->dvec_owner = d->engine.get();
    return d->dvec;
}

void err_print(const char *fn, const char *msg) { fprintf(stderr, "%s: %s\n", fn, msg); }

// Tuning tables: text files of entries
//   entry <key>
//   single_pass <one value per level: -1, 0 or 1>
//   panel_right <one value per level>
//   end
// keyed by the analysed partition and the factorisation it is tuned for.
std::string tuning_key(const pastix_data_t *d) {
    char buf[256];
    snprintf(buf, sizeof(buf), "n=%lld cblknbr=%lld bloknbr=%lld nlevels=%d facto=%lld prec=%lld",
             (long long)d->n, (long long)d->symbol.cblknbr(), (long long)d->symbol.bloknbr(), (int)d->layout.nlevels,
             (long long)d->iparm[IPARM_FACTORIZATION], (long long)d->ext[PXA_IPARM_FACTOR_PRECISION]);
    return buf;
}

// 1: an entry for this analysis was read into k, 0: none, -1: unreadable file
int read_tuning(const pastix_data_t *d, const char *path, pxa::PlanKnobs &k) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    const std::string key = "entry " + tuning_key(d);
    const int nl = d->layout.nlevels;
    char line[1 << 16];
    bool in = false, found = false;
    pxa::PlanKnobs r;
    while (fgets(line, sizeof(line), f)) {
        std::string l(line);
        while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
        if (!in) {
            if (l == key) in = true;
            continue;
        }
        if (l == "end") {
            found = (int)r.single_pass.size() == nl && (int)r.panel_right.size() == nl;
            break;
        }
        std::vector<int8_t> *dst = nullptr;
        size_t pos = 0;
        if (l.rfind("single_pass", 0) == 0) dst = &r.single_pass, pos = 11;
        else if (l.rfind("panel_right", 0) == 0) dst = &r.panel_right, pos = 11;
        if (!dst) continue;
        const char *p = l.c_str() + pos;
        char *e = nullptr;
        for (;;) {
            const long v = strtol(p, &e, 10);
            if (e == p) break;
            dst->push_back((int8_t)std::max(-1L, std::min(1L, v)));
            p = e;
        }
    }
    fclose(f);
    if (found) k = r;
    return found ? 1 : 0;
}
}  // namespace pxa_api
using namespace pxa_api;

static double wall() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// --------------------------------------------------------------------------
// Parameters (reference common/api.c:363-538)
// --------------------------------------------------------------------------
extern "C" void pastixInitParam(pastix_int_t *iparm, double *dparm) {
    memset(iparm, 0, IPARM_SIZE * sizeof(pastix_int_t));
    memset(dparm, 0, DPARM_SIZE * sizeof(double));
    iparm[IPARM_VERBOSE] = PastixVerboseNo;
    iparm[IPARM_IO_STRATEGY] = PastixIONo;
    iparm[IPARM_ORDERING] = PastixOrderScotch;
    iparm[IPARM_ORDERING_DEFAULT] = 1;
    iparm[IPARM_SCOTCH_SWITCH_LEVEL] = 120;
    iparm[IPARM_SCOTCH_CMIN] = 0;
    iparm[IPARM_SCOTCH_CMAX] = 100000;
    iparm[IPARM_SCOTCH_FRAT] = 8;
    iparm[IPARM_METIS_CTYPE] = -1;
    iparm[IPARM_METIS_RTYPE] = -1;
    iparm[IPARM_METIS_NSEPS] = 1;
    iparm[IPARM_METIS_NITER] = 10;
    iparm[IPARM_METIS_UFACTOR] = 200;
    iparm[IPARM_METIS_COMPRESS] = 1;
    iparm[IPARM_METIS_SEED] = 3452;
    iparm[IPARM_AMALGAMATION_LVLCBLK] = 5;
    iparm[IPARM_AMALGAMATION_LVLBLAS] = 5;
    iparm[IPARM_REORDERING_SPLIT] = 0;
    iparm[IPARM_REORDERING_STOP] = PASTIX_INT_MAX;
    iparm[IPARM_SPLITTING_STRATEGY] = PastixSplitNot;
    iparm[IPARM_SPLITTING_LEVELS_PROJECTIONS] = 2;
    iparm[IPARM_SPLITTING_LEVELS_KWAY] = 2;
    iparm[IPARM_SPLITTING_PROJECTIONS_DEPTH] = 3;
    iparm[IPARM_SPLITTING_PROJECTIONS_DISTANCE] = 2;
    iparm[IPARM_SPLITTING_PROJECTIONS_WIDTH] = 1;
    iparm[IPARM_MIN_BLOCKSIZE] = 160;
    iparm[IPARM_MAX_BLOCKSIZE] = 320;
    iparm[IPARM_TASKS2D_LEVEL] = -1;
    iparm[IPARM_TASKS2D_WIDTH] = iparm[IPARM_MIN_BLOCKSIZE];
    iparm[IPARM_FACTORIZATION] = PastixFactLU;
    iparm[IPARM_SCHUR_FACT_MODE] = PastixFactModeLocal;
    iparm[IPARM_SCHUR_SOLV_MODE] = PastixSolvModeLocal;
    iparm[IPARM_APPLYPERM_WS] = 1;
    iparm[IPARM_REFINEMENT] = PastixRefineGMRES;
    iparm[IPARM_ITERMAX] = 250;
    iparm[IPARM_GMRES_IM] = 25;
    iparm[IPARM_SCHEDULER] = PastixSchedStatic;
    iparm[IPARM_THREAD_NBR] = -1;
    iparm[IPARM_GPU_NBR] = 0;
    iparm[IPARM_GPU_MEMORY_PERCENTAGE] = 95;
    iparm[IPARM_GPU_MEMORY_BLOCK_SIZE] = 32 * 1024;
    iparm[IPARM_COMPRESS_MIN_WIDTH] = 120;
    iparm[IPARM_COMPRESS_MIN_HEIGHT] = 20;
    iparm[IPARM_COMPRESS_WHEN] = PastixCompressNever;
    iparm[IPARM_COMPRESS_METHOD] = PastixCompressMethodPQRCP;
    iparm[IPARM_COMPRESS_ORTHO] = PastixCompressOrthoCGS;
    iparm[IPARM_COMPRESS_PRESELECT] = -1;
    iparm[IPARM_MODIFY_PARAMETER] = 1;
    iparm[IPARM_START_TASK] = PastixTaskOrdering;
    iparm[IPARM_END_TASK] = PastixTaskClean;
    iparm[IPARM_FLOAT] = -1;
    iparm[IPARM_MTX_TYPE] = -1;
    iparm[IPARM_DOF_NBR] = 1;
    dparm[DPARM_EPSILON_REFINEMENT] = -1.;
    dparm[DPARM_RELATIVE_ERROR] = -1.;
    dparm[DPARM_EPSILON_MAGN_CTRL] = 0.;
    dparm[DPARM_A_NORM] = -1.;
    dparm[DPARM_COMPRESS_TOLERANCE] = 0.01;
    dparm[DPARM_COMPRESS_MIN_RATIO] = 1.;
}

extern "C" void pastixInitWithAffinity(pastix_data_t **pastix_data, PASTIX_Comm, pastix_int_t *iparm, double *dparm,
                                       const int *) {
    pastix_data_t *d = new pastix_data_s();
    if (iparm[IPARM_MODIFY_PARAMETER] == 0) pastixInitParam(iparm, dparm);
    d->iparm = iparm;
    d->dparm = dparm;
    d->ext[PXA_IPARM_FACTOR_PRECISION] = -1;
    d->ext[PXA_IPARM_DEVICE] = 0;
    d->ext[PXA_IPARM_LEVEL_ORDER] = 1;
    d->ext[PXA_IPARM_OVERLAP] = 1;
    // off: its first GPU runs did not finish (the cause is not found yet)
    d->ext[PXA_IPARM_LEAF_TASKS] = 0;
    d->ext[PXA_IPARM_SOLVE_NODE_MAX] = -1;
    d->steps = STEP_INIT;
    *pastix_data = d;
}

extern "C" void pastixInit(pastix_data_t **pastix_data, PASTIX_Comm comm, pastix_int_t *iparm, double *dparm) {
    pastixInitWithAffinity(pastix_data, comm, iparm, dparm, nullptr);
}

extern "C" void pastixFinalize(pastix_data_t **pastix_data) {
    if (!pastix_data || !*pastix_data) return;
    delete *pastix_data;
    *pastix_data = nullptr;
}

// --------------------------------------------------------------------------
// spm helpers
// --------------------------------------------------------------------------


// Expand the user matrix to 0-based triplets of the full matrix (a symmetric
// matrix stores one triangle, as the SPM library does).
int pxa_api::spm_triplets(const spmatrix_t *spm, Triplets &t, bool values) {
    const int64_t n = spm->n, nnz = spm->nnz;
    if (spm->dof != 1) return PASTIX_ERR_NOTIMPLEMENTED;
    if (spm->flttype != SpmDouble && spm->flttype != SpmFloat && spm->flttype != SpmPattern)
        return PASTIX_ERR_NOTIMPLEMENTED;
    const spm_int_t *cp = spm->colptr, *rp = spm->rowptr;
    t.i.reserve(2 * nnz);
    t.j.reserve(2 * nnz);
    t.src.reserve(2 * nnz);
    if (values) t.v.reserve(2 * nnz);
    auto val = [&](int64_t k) -> double {
        if (!values || spm->flttype == SpmPattern || !spm->values) return 1.0;
        return spm->flttype == SpmDouble ? ((const double *)spm->values)[k] : ((const float *)spm->values)[k];
    };
    auto add = [&](int64_t i, int64_t j, int64_t k) {
        const double v = val(k);
        t.i.push_back(i);
        t.j.push_back(j);
        t.src.push_back(k);
        if (values) t.v.push_back(v);
        if (spm->mtxtype != SpmGeneral && i != j) {
            t.i.push_back(j);
            t.j.push_back(i);
            t.src.push_back(k);
            if (values) t.v.push_back(v);
        }
    };
    if (spm->fmttype == SpmCSC) {
        const int64_t base = cp[0];
        for (int64_t j = 0; j < n; j++)
            for (int64_t k = cp[j] - base; k < cp[j + 1] - base; k++) add(rp[k] - base, j, k);
    } else if (spm->fmttype == SpmCSR) {
        const int64_t base = rp[0];
        for (int64_t i = 0; i < n; i++)
            for (int64_t k = rp[i] - base; k < rp[i + 1] - base; k++) add(i, cp[k] - base, k);
    } else {
        int64_t base = INT64_MAX;
        for (int64_t k = 0; k < nnz; k++) base = std::min<int64_t>(base, std::min<int64_t>(cp[k], rp[k]));
        if (nnz == 0) base = 0;
        for (int64_t k = 0; k < nnz; k++) add(rp[k] - base, cp[k] - base, k);
    }
    for (size_t k = 0; k < t.i.size(); k++)
        if (t.i[k] < 0 || t.i[k] >= n || t.j[k] < 0 || t.j[k] >= n) return PASTIX_ERR_BADPARAMETER;
    return PASTIX_SUCCESS;
}

// Adjacency graph of the input's pattern (symmetrised, no diagonal): CSC
// and CSR arrays are read in place after a bounds check (the graph of A and
// of A^t is the same), IJV through triplets.
int pxa_api::spm_graph(const spmatrix_t *spm, pxa::Graph &g) {
    const int64_t n = spm->n, nnz = spm->nnz;
    if (spm->dof != 1) return PASTIX_ERR_NOTIMPLEMENTED;
    if (spm->flttype != SpmDouble && spm->flttype != SpmFloat && spm->flttype != SpmPattern)
        return PASTIX_ERR_NOTIMPLEMENTED;
    if (spm->fmttype == SpmCSC || spm->fmttype == SpmCSR) {
        const spm_int_t *ptr = spm->fmttype == SpmCSC ? spm->colptr : spm->rowptr;
        const spm_int_t *idx = spm->fmttype == SpmCSC ? spm->rowptr : spm->colptr;
        const int64_t base = ptr[0];
        if (ptr[n] - base != nnz) return PASTIX_ERR_BADPARAMETER;
        for (int64_t j = 0; j < n; j++)
            if (ptr[j + 1] < ptr[j]) return PASTIX_ERR_BADPARAMETER;
        for (int64_t k = 0; k < nnz; k++)
            if (idx[k] < base || idx[k] - base >= n) return PASTIX_ERR_BADPARAMETER;
        static_assert(sizeof(spm_int_t) == sizeof(int64_t), "64-bit spm indices");
        g = pxa::graph_from_csc(n, (const int64_t *)ptr, (const int64_t *)idx, base);
        return PASTIX_SUCCESS;
    }
    Triplets t;
    const int rc = spm_triplets(spm, t, false);
    if (rc) return rc;
    std::vector<int64_t> cp;
    std::vector<int32_t> rw;
    std::vector<double> vv;
    triplets_to_csc(n, t, nullptr, cp, rw, vv, false);
    std::vector<int64_t> rows64(rw.begin(), rw.end());
    g = pxa::graph_from_csc(n, cp.data(), rows64.data(), 0);
    return PASTIX_SUCCESS;
}

// Numbering base of the input (spmFindBase): the first pointer of CSC/CSR,
// the smallest index of IJV.
int64_t pxa_api::spm_baseval(const spmatrix_t *spm) {
    if (spm->fmttype == SpmCSC) return spm->colptr[0];
    if (spm->fmttype == SpmCSR) return spm->rowptr[0];
    int64_t base = INT64_MAX;
    for (int64_t k = 0; k < spm->nnz; k++) base = std::min<int64_t>(base, std::min<int64_t>(spm->colptr[k], spm->rowptr[k]));
    return spm->nnz > 0 ? base : 0;
}

// CSC (sorted, duplicates summed) of the n x n matrix given by triplets,
// optionally permuted: B(perm[i], perm[j]) = A(i, j).
void pxa_api::triplets_to_csc(int64_t n, const Triplets &t, const int64_t *perm, std::vector<int64_t> &colptr,
                              std::vector<int32_t> &rows, std::vector<double> &vals, bool transpose,
                              std::vector<int64_t> *eptr, std::vector<int64_t> *esrc) {
    const size_t m = t.i.size();
    colptr.assign(n + 1, 0);
    std::vector<int64_t> I(m), J(m);
    for (size_t k = 0; k < m; k++) {
        int64_t i = perm ? perm[t.i[k]] : t.i[k];
        int64_t j = perm ? perm[t.j[k]] : t.j[k];
        if (transpose) std::swap(i, j);
        I[k] = i;
        J[k] = j;
        colptr[j + 1]++;
    }
    for (int64_t j = 0; j < n; j++) colptr[j + 1] += colptr[j];
    // bucket by column (one scattering pass; the triplets of a column keep
    // their input order), then sort each column by (row, triplet index):
    // duplicates are summed in input order, so that the sum is a fixed
    // function of the input
    struct Ent {
        int64_t row, k;
        double v;
        bool operator<(const Ent &o) const { return row < o.row || (row == o.row && k < o.k); }
    };
    std::vector<int64_t> pos(colptr.begin(), colptr.end() - 1);
    std::vector<Ent> ent(m);
    for (size_t k = 0; k < m; k++) ent[pos[J[k]]++] = Ent{I[k], (int64_t)k, t.v.empty() ? 1.0 : t.v[k]};
    std::vector<int64_t>().swap(I);
    std::vector<int64_t>().swap(J);
    rows.clear();
    vals.clear();
    rows.reserve(m);
    vals.reserve(m);
    if (eptr) {
        eptr->clear();
        eptr->reserve(m + 1);
    }
    if (esrc) {
        esrc->clear();
        esrc->reserve(m);
    }
    std::vector<int64_t> newptr(n + 1, 0);
    for (int64_t j = 0; j < n; j++) {
        Ent *c0 = ent.data() + colptr[j], *c1 = ent.data() + colptr[j + 1];
        if (!std::is_sorted(c0, c1)) std::sort(c0, c1);
        for (const Ent *e = c0; e < c1; e++) {
            if (e > c0 && e->row == e[-1].row) vals.back() += e->v;
            else {
                if (eptr) eptr->push_back(esrc ? (int64_t)esrc->size() : 0);
                rows.push_back((int32_t)e->row);
                vals.push_back(e->v);
            }
            if (esrc) esrc->push_back(e->k);
        }
        newptr[j + 1] = (int64_t)rows.size();
    }
    if (eptr) eptr->push_back(esrc ? (int64_t)esrc->size() : 0);
    colptr = newptr;
}

// --------------------------------------------------------------------------
// Analysis subtasks
// --------------------------------------------------------------------------
extern "C" int pastix_subtask_order(pastix_data_t *d, const spmatrix_t *spm, pastix_order_t *myorder) {
    if (!d) { err_print("pastix_subtask_order", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!spm) { err_print("pastix_subtask_order", "wrong spm parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_INIT)) { err_print("pastix_subtask_order", "pastixInit() has to be called before"); return PASTIX_ERR_BADPARAMETER; }
    int rc = spm_graph(spm, d->graph);
    if (rc) return rc;
    const int64_t n = spm->n;
    d->n = n;
    d->nnz_in = spm->gnnzexp > 0 ? spm->gnnzexp : spm->nnz;
    d->mtxtype = spm->mtxtype;
    d->flttype = spm->flttype;
    // Schur unknowns (reference order/pastix_subtask_order.c:198-230,
    // graph/graph_isolate.c): the ordering runs on the graph without them,
    // and they are appended last, in increasing index order, as one more
    // supernode that the roots of the elimination forest feed
    // (order/order_add_isolate.c:55-130)
    std::vector<int64_t> keep;  // old index of every remaining vertex, increasing
    std::vector<int64_t> iso;   // the Schur unknowns, increasing
    pxa::Graph sub;
    const pxa::Graph *gord = &d->graph;
    if (!d->schur_list.empty()) {
        const int64_t base = spm_baseval(spm);
        std::vector<char> mark(n, 0);
        for (pastix_int_t v : d->schur_list) {
            const int64_t i = v - base;
            if (i < 0 || i >= n || mark[i]) { err_print("pastix_subtask_order", "invalid or repeated Schur unknown"); return PASTIX_ERR_BADPARAMETER; }
            mark[i] = 1;
        }
        if ((int64_t)d->schur_list.size() >= n) { err_print("pastix_subtask_order", "the Schur complement must leave unknowns to factor"); return PASTIX_ERR_BADPARAMETER; }
        std::vector<int64_t> newid(n, -1);
        for (int64_t i = 0; i < n; i++) {
            if (mark[i]) iso.push_back(i);
            else {
                newid[i] = (int64_t)keep.size();
                keep.push_back(i);
            }
        }
        sub.n = (int64_t)keep.size();
        sub.colptr.assign(sub.n + 1, 0);
        for (int64_t k = 0; k < sub.n; k++) {
            const int64_t j = keep[k];
            for (int64_t e = d->graph.colptr[j]; e < d->graph.colptr[j + 1]; e++)
                if (newid[d->graph.rows[e]] >= 0) sub.rows.push_back(newid[d->graph.rows[e]]);
            sub.colptr[k + 1] = (int64_t)sub.rows.size();
        }
        gord = &sub;
    }
    const int64_t no = gord->n;
    pxa::Order o;
    o.n = no;
    const pastix_int_t ord = d->iparm[IPARM_ORDERING];
    pastix_order_t loaded;
    memset(&loaded, 0, sizeof(loaded));
    if (ord == PastixOrderPersonal && (d->iparm[IPARM_IO_STRATEGY] & PastixIOLoad)) {
        // reference order/pastix_subtask_order.c:308-315: the ordering file
        const int lrc = pastixOrderLoad(d, &loaded);
        if (lrc) return lrc;
        myorder = &loaded;
    }
    if (ord == PastixOrderPersonal) {
        o.permtab.resize(no);
        o.peritab.resize(no);
        const int64_t base = (myorder && myorder->baseval) ? myorder->baseval : 0;
        if (myorder && (myorder->permtab || myorder->peritab) && myorder->vertnbr != no) {
            err_print("pastix_subtask_order", "the given ordering does not match the size of the graph to order (the matrix without its Schur unknowns)");
            pastixOrderExit(&loaded);
            return PASTIX_ERR_BADPARAMETER;
        }
        if (myorder && myorder->permtab) {
            for (int64_t i = 0; i < no; i++) o.permtab[i] = myorder->permtab[i] - base;
            for (int64_t i = 0; i < no; i++) o.peritab[o.permtab[i]] = i;
        } else if (myorder && myorder->peritab) {
            for (int64_t i = 0; i < no; i++) o.peritab[i] = myorder->peritab[i] - base;
            for (int64_t i = 0; i < no; i++) o.permtab[o.peritab[i]] = i;
        } else {
            for (int64_t i = 0; i < no; i++) o.permtab[i] = o.peritab[i] = i;
        }
        if (myorder && myorder->rangtab && myorder->treetab && myorder->cblknbr > 0) {
            o.cblknbr = myorder->cblknbr;
            o.rangtab.resize(o.cblknbr + 1);
            o.treetab.resize(o.cblknbr);
            for (int64_t c = 0; c <= o.cblknbr; c++) o.rangtab[c] = myorder->rangtab[c] - base;
            for (int64_t c = 0; c < o.cblknbr; c++)
                o.treetab[c] = myorder->treetab[c] < 0 ? -1 : myorder->treetab[c] - base;
        } else {
            pxa::order_find_supernodes(*gord, o);
        }
        pastixOrderExit(&loaded);
    } else {
        // Scotch / METIS are not part of this build: graph nested dissection.
        if (d->iparm[IPARM_VERBOSE] > PastixVerboseNo)
            fprintf(stderr, "pastix_subtask_order: ordering library not available, using level-structure nested dissection\n");
        o = pxa::order_nested_dissection(*gord);
    }
    std::string err;
    if (!pxa::check_order(o, err)) { err_print("pastix_subtask_order", err.c_str()); return PASTIX_ERR_BADPARAMETER; }
    // reference order/pastix_subtask_order.c:459-461, on every ordering
    if (d->ext[PXA_IPARM_LEVEL_ORDER])
        pxa::order_apply_level_order(o, d->iparm[IPARM_TASKS2D_LEVEL], d->iparm[IPARM_TASKS2D_WIDTH]);
    if (!iso.empty()) {
        pxa::Order f;
        f.n = n;
        f.permtab.assign(n, 0);
        f.peritab.assign(n, 0);
        for (int64_t k = 0; k < no; k++) f.permtab[keep[k]] = o.permtab[k];
        for (size_t k = 0; k < iso.size(); k++) f.permtab[iso[k]] = no + (int64_t)k;
        for (int64_t i = 0; i < n; i++) f.peritab[f.permtab[i]] = i;
        f.cblknbr = o.cblknbr + 1;
        f.rangtab = o.rangtab;
        f.rangtab.push_back(n);
        f.treetab = o.treetab;
        for (int64_t &t : f.treetab)
            if (t < 0) t = o.cblknbr;
        f.treetab.push_back(-1);
        if (!pxa::check_order(f, err)) { err_print("pastix_subtask_order", err.c_str()); return PASTIX_ERR_INTERNAL; }
        o = std::move(f);
    }
    d->schurfcol = no;
    d->order = std::move(o);
    // reference order/pastix_subtask_order.c:548-552
    if (d->iparm[IPARM_IO_STRATEGY] & PastixIOSave) {
        const int src = save_analysis_file(d, "ordergen", true);
        if (src) return src;
    }
    d->steps |= STEP_ORDERING;
    d->steps &= ~(STEP_SYMBFACT | STEP_ANALYSE | STEP_CSC2BCSC | STEP_BCSC2CTAB | STEP_NUMFACT | STEP_SOLVE | STEP_REFINE);
    d->engine.reset();
    d->bcsc = pxa::HostBcsc();  // the permuted matrix follows the new ordering
    return PASTIX_SUCCESS;
}

extern "C" int pastix_subtask_symbfact(pastix_data_t *d) {
    if (!d) return PASTIX_ERR_BADPARAMETER;
    if (!(d->steps & STEP_ORDERING)) { err_print("pastix_subtask_symbfact", "pastix_subtask_order() has to be called before"); return PASTIX_ERR_BADPARAMETER; }
    if (d->iparm[IPARM_IO_STRATEGY] & PastixIOLoad) {
        // reference symbol/pastix_subtask_symbfact.c:159-166: "symbname"
        FILE *f = fopen("symbname", "r");
        if (!f) { err_print("pastix_subtask_symbfact", "cannot open symbname"); return PASTIX_ERR_FILE; }
        const int lrc = symbol_read(f, d->symbol);
        fclose(f);
        if (lrc || d->symbol.n != d->n) { err_print("pastix_subtask_symbfact", "bad symbol file"); return PASTIX_ERR_FILE; }
    } else {
        d->symbol = pxa::symbol_fax(d->graph, d->order);
    }
    std::string err;
    if (!pxa::check_symbol(d->symbol, err)) { err_print("pastix_subtask_symbfact", err.c_str()); return PASTIX_ERR_INTERNAL; }
    d->iparm[IPARM_NNZEROS] = pxa::symbol_nnz(d->symbol);
    // reference symbol/pastix_subtask_symbfact.c:287-297
    if (d->iparm[IPARM_IO_STRATEGY] & PastixIOSave) {
        const int src = save_analysis_file(d, "symbgen", false);
        if (src) return src;
    }
    d->steps |= STEP_SYMBFACT;
    return PASTIX_SUCCESS;
}

// Reference symbol/pastix_subtask_reordering.c:59-135: the rows of every
// supernode are reordered to compact the blocks facing it, and the symbolic
// factorisation is computed again for the new ordering.
extern "C" int pastix_subtask_reordering(pastix_data_t *d) {
    if (!d) return PASTIX_ERR_BADPARAMETER;
    if (!(d->steps & STEP_SYMBFACT)) { err_print("pastix_subtask_reordering", "pastix_subtask_symbfact() has to be called before"); return PASTIX_ERR_BADPARAMETER; }
    // files saved by an analysis hold the reordered ordering and its symbol
    if (d->iparm[IPARM_IO_STRATEGY] & PastixIOLoad) return PASTIX_SUCCESS;
    pxa::symbol_reordering(d->symbol, d->order, d->iparm[IPARM_REORDERING_SPLIT], d->iparm[IPARM_REORDERING_STOP],
                           d->schurfcol);
    std::string err;
    if (!pxa::check_order(d->order, err)) { err_print("pastix_subtask_reordering", err.c_str()); return PASTIX_ERR_INTERNAL; }
    // reference symbol/pastix_subtask_reordering.c:100-107
    if (d->iparm[IPARM_IO_STRATEGY] & PastixIOSave) {
        const int src = save_analysis_file(d, "ordergen", true);
        if (src) return src;
    }
    d->symbol = pxa::symbol_fax(d->graph, d->order);
    if (!pxa::check_symbol(d->symbol, err)) { err_print("pastix_subtask_reordering", err.c_str()); return PASTIX_ERR_INTERNAL; }
    d->iparm[IPARM_NNZEROS] = pxa::symbol_nnz(d->symbol);
    // the symbolic factorisation runs again and saves its file again
    // (symbol/pastix_subtask_reordering.c:116, pastix_subtask_symbfact.c:287-297)
    if (d->iparm[IPARM_IO_STRATEGY] & PastixIOSave) {
        const int src = save_analysis_file(d, "symbgen", false);
        if (src) return src;
    }
    // the permuted matrix follows the new ordering
    d->engine.reset();
    d->bcsc = pxa::HostBcsc();
    return PASTIX_SUCCESS;
}

// The level schedule of the analysed symbol; its failures (host memory, lists
// beyond the devices' 32-bit indices) come back as error codes.
static int make_plan(pastix_data_t *d, int64_t sc, const char *who) {
    try {
        d->plan = pxa::build_plan(d->symbol, d->layout, sc, d->knobs);
    } catch (const std::bad_alloc &) {
        err_print(who, "out of host memory building the level schedule");
        return PASTIX_ERR_OUTOFMEMORY;
    } catch (const std::exception &e) {
        err_print(who, e.what());
        return PASTIX_ERR_INTERNAL;
    }
    return PASTIX_SUCCESS;
}

extern "C" int pastix_subtask_blend(pastix_data_t *d) {
    if (!d) return PASTIX_ERR_BADPARAMETER;
    if (!(d->steps & STEP_SYMBFACT)) { err_print("pastix_subtask_blend", "pastix_subtask_symbfact() has to be called before"); return PASTIX_ERR_BADPARAMETER; }
    double t0 = wall();
    pastix_int_t bmin = d->iparm[IPARM_MIN_BLOCKSIZE], bmax = d->iparm[IPARM_MAX_BLOCKSIZE];
    if (bmax < 1 || bmin < 1 || bmin > bmax) { err_print("pastix_subtask_blend", "invalid block sizes"); return PASTIX_ERR_BADPARAMETER; }
    d->symbol = pxa::symbol_split(d->symbol, bmin, bmax);
    std::string err;
    if (!pxa::check_symbol(d->symbol, err)) { err_print("pastix_subtask_blend", err.c_str()); return PASTIX_ERR_INTERNAL; }
    d->layout = pxa::build_layout(d->symbol);
    // cblks from the Schur column on (reference CBLK_IN_SCHUR, blend/cand.c:725)
    int64_t sc = d->symbol.cblknbr();
    while (sc > 0 && d->symbol.cblk_fcol[sc - 1] >= d->schurfcol) sc--;
    d->schur_cblk = sc;
    // a tuning table named by PASTIX_AMD_TUNING supplies per-level choices
    // for a matching analysis
    if (const char *tf = getenv("PASTIX_AMD_TUNING")) {
        pxa::PlanKnobs k;
        if (read_tuning(d, tf, k) == 1) d->knobs = k;
    }
    d->knobs.node_max = (int32_t)d->ext[PXA_IPARM_SOLVE_NODE_MAX];
    if (const int prc = make_plan(d, sc, "pastix_subtask_blend")) return prc;
    int facto = (int)d->iparm[IPARM_FACTORIZATION];
    int ft = (facto == PastixFactGETRF) ? 2 : is_ldlt(d) ? 1 : 0;
    d->dparm[DPARM_FACT_THFLOPS] = pxa::symbol_thflops(d->symbol, ft);
    if (facto == PastixFactLU) d->iparm[IPARM_NNZEROS] *= 2;
    d->dparm[DPARM_SOLV_FLOPS] = (double)d->iparm[IPARM_NNZEROS];
    // reference blend/pastix_subtask_blend.c:545: factor entries per entry of
    // the input (gnnzexp: the stored entries of the user's matrix)
    if (d->nnz_in > 0) d->dparm[DPARM_FILL_IN] = (double)d->iparm[IPARM_NNZEROS] / (double)d->nnz_in;
    d->iparm[IPARM_NNZEROS_BLOCK_LOCAL] = d->layout.coefnbr;
    d->dparm[DPARM_ANALYZE_TIME] = wall() - t0;
    d->steps |= STEP_ANALYSE;
    d->engine.reset();
    // table positions follow the new partition
    d->bcsc.cmap.clear();
    d->bcsc.cmap_facto = -1;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_task_analyze(pastix_data_t *d, spmatrix_t *spm) {
    if (!d) { err_print("pastix_task_analyze", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!spm) { err_print("pastix_task_analyze", "wrong spm parameter"); return PASTIX_ERR_BADPARAMETER; }
    double t0 = wall();
    int rc = pastix_subtask_order(d, spm, nullptr);
    if (rc) return rc;
    if ((rc = pastix_subtask_symbfact(d))) return rc;
    if ((rc = pastix_subtask_reordering(d))) return rc;
    if ((rc = pastix_subtask_blend(d))) return rc;
    d->dparm[DPARM_ANALYZE_TIME] = wall() - t0;
    return PASTIX_SUCCESS;
}

// --------------------------------------------------------------------------
// Numerical factorisation subtasks
// --------------------------------------------------------------------------
// Refill of an input with the analysed pattern (the reference rebuilds its
// bcsc from the spm on every numfact, bcsc/bcsc_zinit.c:485-541): the input
// pattern is kept as given, and a later input with the same format, type,
// size and index arrays takes the permuted values by one gather.
static void keep_pattern(const spmatrix_t *spm, pxa::HostBcsc &bc) {
    const int64_t n = spm->n, nnz = spm->nnz;
    const int64_t nc = spm->fmttype == SpmCSC ? n + 1 : nnz, nr = spm->fmttype == SpmCSR ? n + 1 : nnz;
    bc.in_fmt = spm->fmttype;
    bc.in_mtx = spm->mtxtype;
    bc.in_n = n;
    bc.in_nnz = nnz;
    bc.in_colptr.assign(spm->colptr, spm->colptr + nc);
    bc.in_rowptr.assign(spm->rowptr, spm->rowptr + nr);
}

// f(a, b) over [0, n) in up to 16 host threads (one call when n is small);
// callers keep every element independent, so the result does not depend on
// the split
static void parallel_ranges(int64_t n, const std::function<void(int64_t, int64_t)> &f) {
    const int64_t T = std::min<int64_t>(16, std::max(1u, std::thread::hardware_concurrency()));
    if (T == 1 || n < (1 << 18)) {
        f(0, n);
        return;
    }
    const int64_t chunk = (n + T - 1) / T;
    std::vector<std::thread> th;
    for (int64_t a = 0; a < n; a += chunk) th.emplace_back(f, a, std::min(n, a + chunk));
    for (std::thread &t : th) t.join();
}

static bool same_array(const spm_int_t *a, const std::vector<int64_t> &b) {
    std::atomic<bool> same{true};
    parallel_ranges((int64_t)b.size(), [&](int64_t lo, int64_t hi) {
        for (int64_t k = lo; k < hi; k++)
            if (a[k] != b[k]) {
                same = false;
                return;
            }
    });
    return same;
}

static bool same_pattern(const spmatrix_t *spm, const pxa::HostBcsc &bc) {
    if (bc.in_fmt < 0 || spm->fmttype != bc.in_fmt || spm->mtxtype != bc.in_mtx || spm->n != bc.in_n ||
        spm->nnz != bc.in_nnz || spm->dof != 1)
        return false;
    if (spm->flttype != SpmDouble && spm->flttype != SpmFloat && spm->flttype != SpmPattern) return false;
    return same_array(spm->colptr, bc.in_colptr) && same_array(spm->rowptr, bc.in_rowptr);
}

// The values the triplet path gives (spm_triplets + triplets_to_csc): the
// same input values summed in the same order.
static void refill_values(const spmatrix_t *spm, pxa::HostBcsc &bc) {
    const int64_t m = (int64_t)bc.tvals.size();
    if (spm->flttype == SpmPattern || !spm->values) {
        for (int64_t e = 0; e < m; e++) {
            double v = 1.0;
            for (int64_t q = bc.sptr[e] + 1; q < bc.sptr[e + 1]; q++) v += 1.0;
            bc.tvals[e] = v;
        }
        return;
    }
    auto gather = [&](auto *vals) {
        parallel_ranges(m, [&](int64_t lo, int64_t hi) {
            for (int64_t e = lo; e < hi; e++) {
                int64_t q = bc.sptr[e];
                double v = (double)vals[bc.src[q]];
                for (q++; q < bc.sptr[e + 1]; q++) v += (double)vals[bc.src[q]];
                bc.tvals[e] = v;
            }
        });
    };
    if (spm->flttype == SpmDouble) gather((const double *)spm->values);
    else gather((const float *)spm->values);
}

extern "C" int pastix_subtask_spm2bcsc(pastix_data_t *d, spmatrix_t *spm) {
    if (!d) { err_print("pastix_subtask_spm2bcsc", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!spm) { err_print("pastix_subtask_spm2bcsc", "wrong spm parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_ANALYSE)) {
        err_print("pastix_subtask_spm2bcsc", "All steps from pastix_task_init() to pastix_task_blend() have to be called before calling this function");
        return PASTIX_ERR_BADPARAMETER;
    }
    if (spm->n != d->n) { err_print("pastix_subtask_spm2bcsc", "matrix size differs from the analysed one"); return PASTIX_ERR_BADPARAMETER; }
    pxa::HostBcsc &bc = d->bcsc;
    if (same_pattern(spm, bc)) {
        // the analysed pattern again: the permuted values by one gather
        refill_values(spm, bc);
    } else {
        Triplets t;
        int rc = spm_triplets(spm, t, true);
        if (rc) return rc;
        bc.n = d->n;
        bc.mtxtype = spm->mtxtype;
        std::vector<int64_t> tsrc;
        triplets_to_csc(d->n, t, d->order.permtab.data(), bc.rowptr, bc.cols, bc.tvals, true, &bc.sptr, &tsrc);
        bc.src.resize(tsrc.size());
        for (size_t q = 0; q < tsrc.size(); q++) bc.src[q] = t.src[tsrc[q]];
        keep_pattern(spm, bc);
        bc.cmap.clear();
        bc.cmap_facto = -1;
        if (d->engine) d->engine->drop_matrix();
    }
    // norms of the permuted matrix (CSR order)
    double f2 = 0, ninf = 0;
    for (int64_t i = 0; i < d->n; i++) {
        double r = 0;
        for (int64_t k = bc.rowptr[i]; k < bc.rowptr[i + 1]; k++) {
            f2 += bc.tvals[k] * bc.tvals[k];
            r += std::fabs(bc.tvals[k]);
        }
        ninf = std::max(ninf, r);
    }
    bc.normF = std::sqrt(f2);
    bc.normInf = ninf;
    d->dparm[DPARM_A_NORM] = bc.normF;
    if (d->batch_off.size() > 2 && d->batch_off.back() == d->n) {
        // a batch: the static-pivot threshold sqrt(eps) ||A|| (reference
        // sopalin/pastix_task_sopalin.c:401-425) takes the smallest of the
        // systems' own norms, so that no system sees a larger threshold than
        // it would alone (the union's norm grows with the number of systems);
        // the ordering keeps each system in its own row range
        double mn = -1;
        for (size_t q = 0; q + 1 < d->batch_off.size(); q++) {
            double g2 = 0;
            for (int64_t i = d->batch_off[q]; i < d->batch_off[q + 1]; i++)
                for (int64_t k = bc.rowptr[i]; k < bc.rowptr[i + 1]; k++) g2 += bc.tvals[k] * bc.tvals[k];
            if (g2 > 0 && (mn < 0 || g2 < mn)) mn = g2;
        }
        if (mn > 0) d->dparm[DPARM_A_NORM] = std::sqrt(mn);
    }
    d->flttype = spm->flttype == SpmPattern ? SpmDouble : spm->flttype;
    d->csc = spm;
    d->steps |= STEP_CSC2BCSC;
    return PASTIX_SUCCESS;
}

// Position in the coefficient table of every CSR entry of the permuted
// matrix (reference kernels/cpucblk_zinit.c:229-309, cpucblk_zfillin_fr):
// entry (i, j) with i >= j goes to L(i, j) in the cblk of column j; with
// i < j, for LU, B(i, j) goes to U(j, i) in the cblk of column i (the U
// panels hold U^t).  Entries outside the symbol are dropped (-1).
static bool build_fillmap(pastix_data_t *d, int facto) {
    const pxa::Symbol &s = d->symbol;
    const pxa::Layout &L = d->layout;
    pxa::HostBcsc &bc = d->bcsc;
    const int64_t n = d->n;
    std::vector<int32_t> col2cblk(n);
    for (int64_t c = 0; c < s.cblknbr(); c++)
        for (int64_t j = s.cblk_fcol[c]; j < s.cblk_fcol[c + 1]; j++) col2cblk[j] = (int32_t)c;
    auto position = [&](int64_t i, int64_t j) -> int64_t {
        const int64_t c = col2cblk[j];
        int64_t lo = s.cblk_bloknum[c], hi = s.cblk_bloknum[c + 1];
        while (lo < hi) {  // first block whose last row is >= i
            const int64_t mid = (lo + hi) / 2;
            if (s.blok_lrow[mid] < i) lo = mid + 1;
            else hi = mid;
        }
        if (lo >= s.cblk_bloknum[c + 1] || s.blok_frow[lo] > i) return -1;
        return L.cblk_coef[c] + L.blok_roff[lo] + (i - s.blok_frow[lo]) + (j - s.cblk_fcol[c]) * L.cblk_stride[c];
    };
    bc.cmap.assign(bc.tvals.size(), -1);
    for (int64_t i = 0; i < n; i++) {
        for (int64_t e = bc.rowptr[i]; e < bc.rowptr[i + 1]; e++) {
            const int64_t j = bc.cols[e];
            int64_t p = -1;
            if (i >= j) p = position(i, j);
            else if (facto == 2) {
                const int64_t q = position(j, i);
                p = q >= 0 ? -2 - q : -1;
            }
            // every position lies inside the table
            if (p >= L.coefnbr || -2 - p >= L.coefnbr) return false;
            bc.cmap[e] = p;
        }
    }
    bc.cmap_facto = facto;
    return true;
}

// engine switches held in the extension parameters
static void engine_flags(pastix_data_t *d) {
    d->engine->profile = d->ext[PXA_IPARM_PROFILE] != 0;
    d->engine->use_graph = d->ext[PXA_IPARM_GRAPH] != 0;
    d->engine->overlap = d->ext[PXA_IPARM_OVERLAP] != 0;
    d->engine->use_leaf = d->ext[PXA_IPARM_LEAF_TASKS] != 0;
    d->engine->leaf_grid = (int)std::max<pastix_int_t>(0, d->ext[PXA_IPARM_LEAF_GRID]);
}

static int factor_prec(pastix_data_t *d) {
    pastix_int_t p = d->ext[PXA_IPARM_FACTOR_PRECISION];
    if (p == SpmFloat || p == SpmDouble) return (int)p;
    return d->flttype == SpmFloat ? SpmFloat : SpmDouble;
}

extern "C" int pastix_subtask_bcsc2ctab(pastix_data_t *d) {
    if (!d) { err_print("pastix_subtask_bcsc2ctab", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_CSC2BCSC)) {
        err_print("pastix_subtask_bcsc2ctab", "All steps from pastix_task_init() to pastix_stask_spm2bcsc() have to be called before calling this function");
        return PASTIX_ERR_BADPARAMETER;
    }
    // LDL^t runs on the LU engine: on a symmetric matrix, LU without pivoting
    // gives L (unit) with D = diag(U) and U = D L^t, so the L panel ends with
    // the reference's LDL^t layout (kernels/core_zsytrfsp.c:162-168: L below
    // the diagonal, D on it, D L^t above) and the LU solve is the LDL^t solve
    const int facto = (d->iparm[IPARM_FACTORIZATION] == PastixFactGETRF || is_ldlt(d)) ? 2 : 0;
    const int prec = factor_prec(d);
    if (!d->engine || !d->engine->ready() || d->engine->prec() != prec || d->engine->factotype() != facto) {
        d->engine.reset(new pxa::Engine());
        if (d->engine->init((int)d->ext[PXA_IPARM_DEVICE], d->symbol, d->layout, d->plan, facto, prec)) {
            d->engine.reset();
            return PASTIX_ERR_INTERNAL;
        }
    }
    // the permuted values go to the device, which fills the tables
    if (d->engine->upload_matrix(d->bcsc)) return PASTIX_ERR_INTERNAL;
    if (d->bcsc.cmap_facto != facto) {
        if (!build_fillmap(d, facto)) {
            err_print("pastix_subtask_bcsc2ctab", "coefficient position outside the table");
            return PASTIX_ERR_INTERNAL;
        }
        d->engine->drop_fillmap();
    }
    if (!d->engine->has_fillmap() && d->engine->upload_fillmap(d->bcsc.cmap)) return PASTIX_ERR_INTERNAL;
    if (d->engine->fill_coef()) return PASTIX_ERR_INTERNAL;
    if (d->engine->upload_perm(d->order.permtab)) return PASTIX_ERR_INTERNAL;
    if (d->ext[PXA_IPARM_RESIDENT_INPUT] && d->engine->snapshot_coef()) return PASTIX_ERR_INTERNAL;
    engine_flags(d);
    d->steps |= STEP_BCSC2CTAB;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_subtask_sopalin(pastix_data_t *d) {
    if (!d) { err_print("pastix_subtask_sopalin", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_ANALYSE)) { err_print("pastix_subtask_sopalin", "All steps from pastix_task_init() to pastix_task_analyze() have to be called before calling this function"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_BCSC2CTAB)) { err_print("pastix_subtask_sopalin", "All steps from pastix_task_init() to pastix_task_bcsc2ctab() have to be called before calling this function"); return PASTIX_ERR_BADPARAMETER; }
    // static pivoting threshold, reference sopalin/pastix_task_sopalin.c:401-425
    double thr;
    const double magn = d->dparm[DPARM_EPSILON_MAGN_CTRL];
    const int prec = d->engine->prec();
    if (magn < 0.) thr = -magn;
    else if (magn == 0.) {
        double eps = (prec == SpmFloat) ? std::ldexp(1.0, -24) : std::ldexp(1.0, -53);
        thr = std::sqrt(eps) * d->dparm[DPARM_A_NORM];
    } else thr = magn * d->dparm[DPARM_A_NORM];
    int64_t npiv = 0;
    int neg = 0;
    double secs = 0;
    if (d->engine->factorize(thr, &npiv, &neg, &secs)) return PASTIX_ERR_INTERNAL;
    d->dparm[DPARM_FACT_TIME] = secs;
    d->dparm[DPARM_FACT_FLOPS] = (secs > 0) ? d->dparm[DPARM_FACT_THFLOPS] / secs / (1 << 30) : 0.;
    d->iparm[IPARM_STATIC_PIVOTING] = npiv;
    d->fact_times[0] = secs;
    d->fact_times[1] = d->engine->t_diag;
    d->fact_times[2] = 0;
    d->fact_times[3] = d->engine->t_upd;
    d->fact_times[4] = d->engine->n_diag;
    d->fact_times[5] = 0;
    d->fact_times[6] = d->engine->n_upd;
    if (neg && d->iparm[IPARM_FACTORIZATION] != PastixFactGETRF && !is_ldlt(d)) {
        // The reference aborts on a negative pivot (kernels/core_zpotrfsp.c:84-89);
        // this engine reports the failure instead.
        err_print("pastix_subtask_sopalin", "Negative diagonal term");
        return PASTIX_ERR_INTERNAL;
    }
    d->steps |= STEP_NUMFACT;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_task_numfact(pastix_data_t *d, spmatrix_t *spm) {
    if (!d) { err_print("pastix_task_sopalin", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!spm) { err_print("pastix_task_sopalin", "wrong spm parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_ANALYSE)) {
        err_print("pastix_task_sopalin", "All steps from pastix_task_init() to pastix_task_blend() have to be called before calling this function");
        return PASTIX_ERR_BADPARAMETER;
    }
    d->steps &= ~(STEP_CSC2BCSC | STEP_BCSC2CTAB | STEP_NUMFACT | STEP_SOLVE | STEP_REFINE);
    int rc;
    if ((rc = pastix_subtask_spm2bcsc(d, spm))) return rc;
    if ((rc = pastix_subtask_bcsc2ctab(d))) return rc;
    if ((rc = pastix_subtask_sopalin(d))) return rc;
    return PASTIX_SUCCESS;
}

// --------------------------------------------------------------------------
// Solve
// --------------------------------------------------------------------------
extern "C" int pastix_subtask_applyorder(pastix_data_t *d, pastix_coeftype_t flttype, pastix_dir_t dir, pastix_int_t m,
                                         pastix_int_t n, void *b, pastix_int_t ldb) {
    if (!d || !b || m != d->n || ldb < m) return PASTIX_ERR_BADPARAMETER;
    if (!(d->steps & STEP_ORDERING)) return PASTIX_ERR_BADPARAMETER;
    const size_t es = (flttype == SpmFloat) ? 4 : 8;
    std::vector<char> tmp((size_t)m * es);
    for (pastix_int_t r = 0; r < n; r++) {
        char *col = (char *)b + (size_t)r * ldb * es;
        for (pastix_int_t i = 0; i < m; i++) {
            // forward: xp[perm[i]] = b[i]; backward: b[i] = xp[perm[i]]
            if (dir == PastixDirForward) memcpy(&tmp[(size_t)d->order.permtab[i] * es], col + (size_t)i * es, es);
            else memcpy(&tmp[(size_t)i * es], col + (size_t)d->order.permtab[i] * es, es);
        }
        memcpy(col, tmp.data(), (size_t)m * es);
    }
    return PASTIX_SUCCESS;
}

// Schur solve mode of the next solve (IPARM_SCHUR_SOLV_MODE, reference
// include/pastix/api.h:212-236): Local and Interface; the Schur mode would
// solve with the Schur part's coefficients, which numfact leaves unfactored.
bool pxa_api::schur_mode(pastix_data_t *d, const char *fn) {
    const pastix_int_t m = d->iparm[IPARM_SCHUR_SOLV_MODE];
    if (d->engine) d->engine->schur_interface = (m == PastixSolvModeInterface);
    if (d->schurfcol < d->n && m == PastixSolvModeSchur) {
        err_print(fn, "the Schur complement is not factored: use PastixSolvModeLocal or PastixSolvModeInterface");
        return false;
    }
    return true;
}

extern "C" int pastix_subtask_solve(pastix_data_t *d, pastix_int_t nrhs, void *b, pastix_int_t ldb) {
    if (!d || !b || nrhs < 0 || ldb < d->n) return PASTIX_ERR_BADPARAMETER;
    if (!(d->steps & STEP_NUMFACT)) { err_print("pastix_subtask_solve", "All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function"); return PASTIX_ERR_BADPARAMETER; }
    if (!schur_mode(d, "pastix_subtask_solve")) return PASTIX_ERR_BADPARAMETER;
    // b is already in the permuted ordering: identity permutation for the engine
    std::vector<int64_t> id(d->n);
    for (int64_t i = 0; i < d->n; i++) id[i] = i;
    double secs = 0;
    if (d->engine->solve_host(id.data(), d->flttype, b, ldb, (int)nrhs, &secs)) return PASTIX_ERR_INTERNAL;
    d->dparm[DPARM_SOLV_TIME] = secs;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_task_solve(pastix_data_t *d, pastix_int_t nrhs, void *b, pastix_int_t ldb) {
    if (!d) { err_print("pastix_task_solve", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_NUMFACT)) { err_print("pastix_task_solve", "All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function"); return PASTIX_ERR_BADPARAMETER; }
    if (!b || nrhs < 0 || ldb < d->n) return PASTIX_ERR_BADPARAMETER;
    if (!schur_mode(d, "pastix_task_solve")) return PASTIX_ERR_BADPARAMETER;
    if (nrhs == 0) return PASTIX_SUCCESS;
    double secs = 0;
    if (d->engine->solve_host(d->order.permtab.data(), d->flttype, b, ldb, (int)nrhs, &secs)) return PASTIX_ERR_INTERNAL;
    d->dparm[DPARM_SOLV_TIME] = secs;
    d->steps |= STEP_SOLVE;
    return PASTIX_SUCCESS;
}

// One sweep of the solve on b in the permuted ordering (reference
// sopalin/pastix_task_solve.c:219-293).  The sweeps of each factorisation, as
// the reference's solve chains them (:433-516): LL^t forward (Lower, NoTrans,
// NonUnit) and backward (Lower, Trans); LU forward (Lower, NoTrans, Unit) and
// backward (Upper, NoTrans, NonUnit); LDL^t forward (Lower, NoTrans, Unit),
// pastix_subtask_diag, backward (Lower, Trans, Unit).  Only the left side.
extern "C" int pastix_subtask_trsm(pastix_data_t *d, pastix_coeftype_t flttype, pastix_side_t side,
                                   pastix_uplo_t uplo, pastix_trans_t trans, pastix_diag_t diag, pastix_int_t nrhs,
                                   void *b, pastix_int_t ldb) {
    if (!d) { err_print("pastix_subtask_trsm", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!b) { err_print("pastix_subtask_trsm", "wrong b parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_NUMFACT)) { err_print("pastix_subtask_trsm", "All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function"); return PASTIX_ERR_BADPARAMETER; }
    if (nrhs < 0 || ldb < d->n || (flttype != SpmDouble && flttype != SpmFloat)) return PASTIX_ERR_BADPARAMETER;
    if (trans == PastixConjTrans) trans = PastixTrans;  // real arithmetic
    const bool lu = d->iparm[IPARM_FACTORIZATION] == PastixFactGETRF, ldlt = is_ldlt(d);
    int part = 0;
    if (side == PastixLeft) {
        if (uplo == PastixLower && trans == PastixNoTrans && diag == ((lu || ldlt) ? PastixUnit : PastixNonUnit))
            part = 1;
        else if (lu && uplo == PastixUpper && trans == PastixNoTrans && diag == PastixNonUnit)
            part = 2;
        else if (!lu && uplo == PastixLower && trans == PastixTrans && diag == (ldlt ? PastixUnit : PastixNonUnit))
            part = ldlt ? 3 : 2;
    }
    if (!part) {
        err_print("pastix_subtask_trsm", "this side/uplo/trans/diag is not a sweep of the factor");
        return PASTIX_ERR_BADPARAMETER;
    }
    if (!schur_mode(d, "pastix_subtask_trsm")) return PASTIX_ERR_BADPARAMETER;
    if (nrhs == 0) return PASTIX_SUCCESS;
    double secs = 0;
    if (d->engine->solve_host_part(part, flttype, b, ldb, (int)nrhs, &secs)) return PASTIX_ERR_INTERNAL;
    return PASTIX_SUCCESS;
}

// b_i /= the factor's diagonal entry (D of LDL^t), reference
// sopalin/pastix_task_solve.c:330-371 (solve_cblk_zdiag), b permuted.
extern "C" int pastix_subtask_diag(pastix_data_t *d, pastix_coeftype_t flttype, pastix_int_t nrhs, void *b,
                                   pastix_int_t ldb) {
    if (!d) { err_print("pastix_subtask_diag", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!b) { err_print("pastix_subtask_diag", "wrong b parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_NUMFACT)) { err_print("pastix_subtask_diag", "All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function"); return PASTIX_ERR_BADPARAMETER; }
    if (nrhs < 0 || ldb < d->n || (flttype != SpmDouble && flttype != SpmFloat)) return PASTIX_ERR_BADPARAMETER;
    if (!schur_mode(d, "pastix_subtask_diag")) return PASTIX_ERR_BADPARAMETER;
    if (nrhs == 0) return PASTIX_SUCCESS;
    double secs = 0;
    if (d->engine->solve_host_part(4, flttype, b, ldb, (int)nrhs, &secs)) return PASTIX_ERR_INTERNAL;
    return PASTIX_SUCCESS;
}

// --------------------------------------------------------------------------
// Refinement (reference refinement/pastix_task_refine.c:115-288)
// --------------------------------------------------------------------------
static int refine_common(pastix_data_t *d, pastix_int_t nrhs, const void *b, pastix_int_t ldb, void *x,
                         pastix_int_t ldx, bool permute) {
    const int64_t n = d->n;
    if (d->schurfcol < n) {
        err_print("pastix_task_refine", "refinement needs the whole factor: not available with a Schur complement");
        return PASTIX_ERR_BADPARAMETER;
    }
    d->engine->schur_interface = false;
    if (nrhs <= 0) return PASTIX_SUCCESS;
    const bool isf = d->flttype == SpmFloat;
    // FP64 copies in the caller's ordering; the permutation runs on the device
    std::vector<double> hb((size_t)n * nrhs), hx((size_t)n * nrhs);
    for (pastix_int_t r = 0; r < nrhs; r++)
        for (int64_t i = 0; i < n; i++) {
            hb[i + r * n] = isf ? ((const float *)b)[i + r * ldb] : ((const double *)b)[i + r * ldb];
            hx[i + r * n] = isf ? ((const float *)x)[i + r * ldx] : ((const double *)x)[i + r * ldx];
        }
    double *dB = device_block(d, 2 * n * nrhs);
    if (!dB) return PASTIX_ERR_OUTOFMEMORY;
    double *dX = dB + (size_t)n * nrhs;
    const size_t bytes = hb.size() * sizeof(double);
    if (hipMemcpy(dB, hb.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dX, hx.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        return PASTIX_ERR_INTERNAL;
    int it = 0;
    double res = 0;
    const int im = (int)std::max<pastix_int_t>(1, d->iparm[IPARM_GMRES_IM]);
    if (d->engine->refine_user(dB, n, dX, n, (int)nrhs, std::vector<int64_t>{0, n}, refine_method(d), false, permute,
                               (int)d->iparm[IPARM_ITERMAX], im, refine_eps(d), &it, &res))
        return PASTIX_ERR_INTERNAL;
    if (hipMemcpy(hx.data(), dX, bytes, hipMemcpyDeviceToHost) != hipSuccess) return PASTIX_ERR_INTERNAL;
    for (pastix_int_t r = 0; r < nrhs; r++)
        for (int64_t i = 0; i < n; i++) {
            if (isf) ((float *)x)[i + r * ldx] = (float)hx[i + r * n];
            else ((double *)x)[i + r * ldx] = hx[i + r * n];
        }
    d->iparm[IPARM_NBITER] = it;
    d->dparm[DPARM_RELATIVE_ERROR] = res;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_subtask_refine(pastix_data_t *d, pastix_int_t n, pastix_int_t nrhs, const void *b,
                                     pastix_int_t ldb, void *x, pastix_int_t ldx) {
    if (!d || !b || !x || n != d->n) return PASTIX_ERR_BADPARAMETER;
    if (!(d->steps & STEP_NUMFACT)) return PASTIX_ERR_BADPARAMETER;
    double t0 = wall();
    int rc = refine_common(d, nrhs, b, ldb, x, ldx, false);
    d->dparm[DPARM_REFINE_TIME] = wall() - t0;
    return rc;
}

extern "C" int pastix_task_refine(pastix_data_t *d, pastix_int_t n, pastix_int_t nrhs, void *b, pastix_int_t ldb,
                                  void *x, pastix_int_t ldx) {
    if (!d) { err_print("pastix_task_refine", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_NUMFACT)) { err_print("pastix_task_refine", "All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function"); return PASTIX_ERR_BADPARAMETER; }
    if (!b || !x || n != d->n || ldb < n || ldx < n) return PASTIX_ERR_BADPARAMETER;
    double t0 = wall();
    int rc = refine_common(d, nrhs, b, ldb, x, ldx, true);
    d->dparm[DPARM_REFINE_TIME] = wall() - t0;
    if (rc == PASTIX_SUCCESS) d->steps |= STEP_REFINE;
    return rc;
}

// --------------------------------------------------------------------------
// Legacy and auxiliary entry points
// --------------------------------------------------------------------------
extern "C" int pastix(pastix_data_t **, PASTIX_Comm, pastix_int_t, pastix_int_t *, pastix_int_t *, void *,
                      pastix_int_t *, pastix_int_t *, void *, pastix_int_t, pastix_int_t *, double *) {
    err_print("pastix", "the former single-call interface is not provided; use the task interface");
    return PASTIX_ERR_NOTIMPLEMENTED;
}
// Reference sopalin/schur.c:50-58: the list is kept as given (the input's
// numbering base); the next pastix_subtask_order orders these unknowns last.
extern "C" void pastixSetSchurUnknownList(pastix_data_t *d, pastix_int_t n, const pastix_int_t *list) {
    if (!d) return;
    if (n > 0 && list) d->schur_list.assign(list, list + n);
}
// Reference sopalin/schur.c:90-148 with coeftab_zgetschur (sopalin/coeftab_z.c:
// 413-437) and cpucblk_zgetschur_fr (kernels/cpucblk_zschur.c:113-154): the
// Schur cblks' panels, never factored, are copied block by block into the
// schur_n x schur_n array S (leading dimension lds), in the order the
// analysis gave the Schur unknowns (increasing input index); for LU the
// transposed U panels are added (S = L + U^T, the sum the factorisation's
// geadd forms).  For LL^t and LDL^t only the lower triangle is defined (the
// reference leaves partial sums above the diagonal); it is zero here.  The
// values are in the arithmetic of the input.
extern "C" int pastixGetSchur(const pastix_data_t *d, void *S, pastix_int_t lds) {
    if (!d) { err_print("pastix_getSchur", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!S) { err_print("pastix_getSchur", "S parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (lds <= 0) { err_print("pastix_getSchur", "lds parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_NUMFACT) || !d->engine) {
        err_print("pastix_getSchur", "All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function");
        return PASTIX_ERR_BADPARAMETER;
    }
    const pxa::Symbol &sy = d->symbol;
    const pxa::Layout &L = d->layout;
    const int64_t f0 = d->schurfcol, ns = d->n - f0;
    if (ns <= 0) { err_print("pastix_getSchur", "no Schur complement was requested"); return PASTIX_ERR_BADPARAMETER; }
    if (lds < ns) { err_print("pastix_getSchur", "lds parameter"); return PASTIX_ERR_BADPARAMETER; }
    int64_t c0 = sy.cblknbr();
    while (c0 > 0 && sy.cblk_fcol[c0 - 1] >= f0) c0--;
    const int64_t off = L.cblk_coef[c0], cnt = L.coefnbr - off;
    const bool lu = d->engine->factotype() == 2 && !is_ldlt(d);
    std::vector<double> hl(cnt), hu(lu ? cnt : 0);
    if (d->engine->prec() == SpmFloat) {
        std::vector<float> fl(cnt), fu(lu ? cnt : 0);
        if (d->engine->download_coef_range(off, cnt, fl.data(), lu ? fu.data() : nullptr)) return PASTIX_ERR_INTERNAL;
        std::copy(fl.begin(), fl.end(), hl.begin());
        std::copy(fu.begin(), fu.end(), hu.begin());
    } else if (d->engine->download_coef_range(off, cnt, hl.data(), lu ? hu.data() : nullptr)) {
        return PASTIX_ERR_INTERNAL;
    }
    std::vector<double> dense((size_t)ns * ns, 0.0);
    for (int64_t c = c0; c < sy.cblknbr(); c++) {
        const int64_t fc = sy.cblk_fcol[c], K = sy.cblk_fcol[c + 1] - fc, st = L.cblk_stride[c];
        const int64_t base = L.cblk_coef[c] - off;
        for (int64_t b = sy.cblk_bloknum[c]; b < sy.cblk_bloknum[c + 1]; b++) {
            const int64_t fr = sy.blok_frow[b], nr = sy.blok_lrow[b] - fr + 1, ro = L.blok_roff[b];
            for (int64_t j = 0; j < K; j++)
                for (int64_t i = 0; i < nr; i++) {
                    const int64_t gi = fr + i - f0, gj = fc + j - f0;
                    const double v = hl[base + ro + i + j * st];
                    if (lu) {
                        dense[gi + gj * ns] += v;
                        dense[gj + gi * ns] += hu[base + ro + i + j * st];
                    } else if (gi >= gj) {
                        dense[gi + gj * ns] = v;
                    }
                }
        }
    }
    for (int64_t j = 0; j < ns; j++)
        for (int64_t i = 0; i < ns; i++) {
            const double v = dense[i + j * ns];
            if (d->flttype == SpmFloat) ((float *)S)[i + j * lds] = (float)v;
            else ((double *)S)[i + j * lds] = v;
        }
    return PASTIX_SUCCESS;
}
extern "C" void pastixExpand(const pastix_data_t *, spmatrix_t *) {}
// Diagonal of the factor, permuted order, at stride incD, in the instance's
// arithmetic (reference sopalin/diag.c:55-102, coeftab_zgetdiag
// sopalin/coeftab_z.c:462-492): L's diagonal for LL^t, U's for LU, D for
// LDL^t.
extern "C" int pastixGetDiag(const pastix_data_t *d, void *D, pastix_int_t incD) {
    if (!d) { err_print("pastix_getDiag", "wrong pastix_data parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!D) { err_print("pastix_getDiag", "D parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (incD <= 0) { err_print("pastix_getDiag", "incD parameter"); return PASTIX_ERR_BADPARAMETER; }
    if (!(d->steps & STEP_NUMFACT) || !d->engine) {
        err_print("pastix_getDiag", "All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function");
        return PASTIX_ERR_BADPARAMETER;
    }
    std::vector<double> dg(d->n);
    if (d->engine->download_diag(dg.data())) return PASTIX_ERR_INTERNAL;
    if (d->flttype == SpmFloat)
        for (int64_t i = 0; i < d->n; i++) ((float *)D)[i * incD] = (float)dg[i];
    else
        for (int64_t i = 0; i < d->n; i++) ((double *)D)[i * incD] = dg[i];
    return PASTIX_SUCCESS;
}
extern "C" void pastixGetOptions(int, char **, pastix_int_t *, double *, int *check, spm_driver_t *driver,
                                 char **filename) {
    if (check) *check = 1;
    if (driver) *driver = SpmDriverLaplacian;
    if (filename) *filename = nullptr;
}

// --------------------------------------------------------------------------
// Ordering structure helpers (reference order/order.c, order/order_grids.c)
// --------------------------------------------------------------------------
extern "C" int pastixOrderAlloc(pastix_order_t *const o, pastix_int_t vertnbr, pastix_int_t cblknbr) {
    if (!o) return PASTIX_ERR_BADPARAMETER;
    memset(o, 0, sizeof(*o));
    o->vertnbr = vertnbr;
    o->cblknbr = cblknbr;
    if (vertnbr > 0) {
        o->permtab = (pastix_int_t *)malloc(vertnbr * sizeof(pastix_int_t));
        o->peritab = (pastix_int_t *)malloc(vertnbr * sizeof(pastix_int_t));
    }
    if (cblknbr > 0) {
        o->rangtab = (pastix_int_t *)calloc(cblknbr + 1, sizeof(pastix_int_t));
        o->treetab = (pastix_int_t *)calloc(cblknbr, sizeof(pastix_int_t));
    }
    return PASTIX_SUCCESS;
}
extern "C" int pastixOrderInit(pastix_order_t *const o, pastix_int_t baseval, pastix_int_t vertnbr,
                               pastix_int_t cblknbr, pastix_int_t *const permtab, pastix_int_t *const peritab,
                               pastix_int_t *const rangtab, pastix_int_t *const treetab) {
    if (!o) return PASTIX_ERR_BADPARAMETER;
    memset(o, 0, sizeof(*o));
    o->baseval = baseval;
    o->vertnbr = vertnbr;
    o->cblknbr = cblknbr;
    o->permtab = permtab;
    o->peritab = peritab;
    o->rangtab = rangtab;
    o->treetab = treetab;
    return PASTIX_SUCCESS;
}
extern "C" void pastixOrderExit(pastix_order_t *const o) {
    if (!o) return;
    free(o->permtab);
    free(o->peritab);
    free(o->rangtab);
    free(o->treetab);
    free(o->selevtx);
    free(o->sndetab);
    memset(o, 0, sizeof(*o));
}
extern "C" void pastixOrderBase(pastix_order_t *const o, pastix_int_t baseval) {
    if (!o || o->baseval == baseval) return;
    pastix_int_t d = baseval - o->baseval;
    for (pastix_int_t i = 0; i < o->vertnbr; i++) {
        if (o->permtab) o->permtab[i] += d;
        if (o->peritab) o->peritab[i] += d;
    }
    if (o->rangtab)
        for (pastix_int_t i = 0; i <= o->cblknbr; i++) o->rangtab[i] += d;
    if (o->treetab)
        for (pastix_int_t i = 0; i < o->cblknbr; i++)
            if (o->treetab[i] >= 0) o->treetab[i] += d;
    o->baseval = baseval;
}
extern "C" int pastixOrderCheck(const pastix_order_t *const o) {
    if (!o) return PASTIX_ERR_BADPARAMETER;
    pxa::Order t;
    t.n = o->vertnbr;
    t.cblknbr = o->cblknbr;
    t.permtab.assign(o->permtab, o->permtab + o->vertnbr);
    t.peritab.assign(o->peritab, o->peritab + o->vertnbr);
    t.rangtab.assign(o->rangtab, o->rangtab + o->cblknbr + 1);
    t.treetab.assign(o->treetab, o->treetab + o->cblknbr);
    for (auto &v : t.permtab) v -= o->baseval;
    for (auto &v : t.peritab) v -= o->baseval;
    for (auto &v : t.rangtab) v -= o->baseval;
    for (auto &v : t.treetab)
        if (v >= 0) v -= o->baseval;
    std::string err;
    return pxa::check_order(t, err) ? PASTIX_SUCCESS : PASTIX_ERR_BADPARAMETER;
}
extern "C" int pastixOrderGrid(pastix_order_t **myorder, pastix_int_t nx, pastix_int_t ny, pastix_int_t nz) {
    if (!myorder || !*myorder || nx < 1 || ny < 1 || nz < 1) return PASTIX_ERR_BADPARAMETER;
    pxa::Order o = pxa::order_grid(nx, ny, nz);
    pastix_order_t *r = *myorder;
    pastixOrderAlloc(r, o.n, o.cblknbr);
    for (int64_t i = 0; i < o.n; i++) {
        r->permtab[i] = o.permtab[i];
        r->peritab[i] = o.peritab[i];
    }
    for (int64_t c = 0; c <= o.cblknbr; c++) r->rangtab[c] = o.rangtab[c];
    for (int64_t c = 0; c < o.cblknbr; c++) r->treetab[c] = o.treetab[c];
    return PASTIX_SUCCESS;
}
// Reference order/order_apply_level_order.c:81-333, in the order's base.
extern "C" int pastixOrderApplyLevelOrder(pastix_order_t *o, pastix_int_t level_tasks2d,
                                          pastix_int_t width_tasks2d) {
    if (!o || !o->rangtab || !o->treetab || !o->permtab || !o->peritab || o->cblknbr < 0)
        return PASTIX_ERR_BADPARAMETER;
    const pastix_int_t b = o->baseval;
    pxa::Order t;
    t.n = o->vertnbr;
    t.cblknbr = o->cblknbr;
    for (pastix_int_t i = 0; i < o->vertnbr; i++) {
        t.permtab.push_back(o->permtab[i] - b);
        t.peritab.push_back(o->peritab[i] - b);
    }
    for (pastix_int_t c = 0; c <= o->cblknbr; c++) t.rangtab.push_back(o->rangtab[c] - b);
    for (pastix_int_t c = 0; c < o->cblknbr; c++) t.treetab.push_back(o->treetab[c] < b ? -1 : o->treetab[c] - b);
    pxa::order_apply_level_order(t, level_tasks2d, width_tasks2d);
    for (pastix_int_t i = 0; i < o->vertnbr; i++) {
        o->permtab[i] = t.permtab[i] + b;
        o->peritab[i] = t.peritab[i] + b;
    }
    for (pastix_int_t c = 0; c <= o->cblknbr; c++) o->rangtab[c] = t.rangtab[c] + b;
    for (pastix_int_t c = 0; c < o->cblknbr; c++) o->treetab[c] = t.treetab[c] < 0 ? -1 : t.treetab[c] + b;
    return PASTIX_SUCCESS;
}

// --------------------------------------------------------------------------
// Engine extensions (pastix_amd.h)
// --------------------------------------------------------------------------
static int replan(pastix_data_t *d);
extern "C" int pastix_amd_set_iparm(pastix_data_t *d, int key, pastix_int_t value) {
    if (!d || key < 0 || key >= PXA_IPARM_SIZE) return PASTIX_ERR_BADPARAMETER;
    if (key == PXA_IPARM_FACTOR_PRECISION && value != -1 && value != SpmFloat && value != SpmDouble)
        return PASTIX_ERR_BADPARAMETER;
    const bool replan_nodes = key == PXA_IPARM_SOLVE_NODE_MAX && d->ext[key] != value;
    d->ext[key] = value;
    if (replan_nodes) {
        d->knobs.node_max = (int32_t)value;
        if (d->steps & STEP_ANALYSE) return replan(d);
    }
    if (d->engine) engine_flags(d);
    return PASTIX_SUCCESS;
}
extern "C" int pastix_amd_get_iparm(const pastix_data_t *d, int key, pastix_int_t *value) {
    if (!d || !value || key < 0 || key >= PXA_IPARM_SIZE) return PASTIX_ERR_BADPARAMETER;
    *value = d->ext[key];
    return PASTIX_SUCCESS;
}
extern "C" int pastix_amd_symbol_sizes(const pastix_data_t *d, int64_t out[8]) {
    if (!d || !(d->steps & STEP_ANALYSE)) return PASTIX_ERR_BADPARAMETER;
    out[0] = d->n;
    out[1] = d->symbol.cblknbr();
    out[2] = d->symbol.bloknbr();
    out[3] = d->layout.coefnbr;
    out[4] = (int64_t)d->layout.browtab.size();
    out[5] = d->layout.nlevels;
    out[6] = (int64_t)d->plan.upd.size();
    out[7] = (int64_t)d->plan.ents.size();
    return PASTIX_SUCCESS;
}
extern "C" int pastix_amd_symbol_export(const pastix_data_t *d, int64_t *cblk_fcol, int64_t *cblk_bloknum,
                                        int64_t *cblk_stride, int64_t *cblk_coef, int64_t *blok_frow,
                                        int64_t *blok_lrow, int64_t *blok_fcblk, int64_t *blok_roff, int64_t *permtab,
                                        int64_t *peritab) {
    if (!d || !(d->steps & STEP_ANALYSE)) return PASTIX_ERR_BADPARAMETER;
    const pxa::Symbol &s = d->symbol;
    const pxa::Layout &L = d->layout;
    auto cp = [](const std::vector<int64_t> &v, int64_t *o) {
        if (o) std::copy(v.begin(), v.end(), o);
    };
    cp(s.cblk_fcol, cblk_fcol);
    cp(s.cblk_bloknum, cblk_bloknum);
    cp(L.cblk_stride, cblk_stride);
    cp(L.cblk_coef, cblk_coef);
    cp(s.blok_frow, blok_frow);
    cp(s.blok_lrow, blok_lrow);
    cp(s.blok_fcblk, blok_fcblk);
    cp(L.blok_roff, blok_roff);
    cp(d->order.permtab, permtab);
    cp(d->order.peritab, peritab);
    return PASTIX_SUCCESS;
}
extern "C" int pastix_amd_coeftab_export(const pastix_data_t *d, void *L, void *U) {
    if (!d || !(d->steps & STEP_NUMFACT) || !d->engine) return PASTIX_ERR_BADPARAMETER;
    return d->engine->download_coef(L, U) ? PASTIX_ERR_INTERNAL : PASTIX_SUCCESS;
}
extern "C" int pastix_amd_solve_device(pastix_data_t *d, pastix_int_t nrhs, void *d_b, pastix_int_t ldb) {
    if (!d || !d_b || nrhs < 0 || ldb < d->n) return PASTIX_ERR_BADPARAMETER;
    if (!(d->steps & STEP_NUMFACT)) return PASTIX_ERR_BADPARAMETER;
    if (!schur_mode(d, "pastix_amd_solve_device")) return PASTIX_ERR_BADPARAMETER;
    if (nrhs == 0) return PASTIX_SUCCESS;
    double secs = 0;
    if (d->engine->solve_user_device(d_b, ldb, (int)nrhs, &secs)) return PASTIX_ERR_INTERNAL;
    d->dparm[DPARM_SOLV_TIME] = secs;
    return PASTIX_SUCCESS;
}
static int refine_device(pastix_data_t *d, pastix_int_t nrhs, const double *d_b, pastix_int_t ldb, double *d_x,
                         pastix_int_t ldx, bool initial_solve) {
    if (!d || !d_b || !d_x || nrhs < 0 || ldb < d->n || ldx < d->n) return PASTIX_ERR_BADPARAMETER;
    if (!(d->steps & STEP_NUMFACT) || !d->engine) return PASTIX_ERR_BADPARAMETER;
    if (d->schurfcol < d->n) {
        err_print("pastix_amd_refine_device", "refinement needs the whole factor: not available with a Schur complement");
        return PASTIX_ERR_BADPARAMETER;
    }
    d->engine->schur_interface = false;
    if (nrhs == 0) return PASTIX_SUCCESS;
    const double t0 = wall();
    int it = 0;
    double res = 0;
    const int im = (int)std::max<pastix_int_t>(1, d->iparm[IPARM_GMRES_IM]);
    if (d->engine->refine_user(d_b, ldb, d_x, ldx, (int)nrhs, std::vector<int64_t>{0, d->n}, refine_method(d),
                               initial_solve, true, (int)d->iparm[IPARM_ITERMAX], im, refine_eps(d), &it, &res))
        return PASTIX_ERR_INTERNAL;
    d->iparm[IPARM_NBITER] = it;
    d->dparm[DPARM_RELATIVE_ERROR] = res;
    d->dparm[DPARM_REFINE_TIME] = wall() - t0;
    d->steps |= STEP_REFINE;
    return PASTIX_SUCCESS;
}
extern "C" int pastix_amd_refine_device(pastix_data_t *d, pastix_int_t nrhs, const double *d_b, pastix_int_t ldb,
                                        double *d_x, pastix_int_t ldx) {
    return refine_device(d, nrhs, d_b, ldb, d_x, ldx, false);
}
extern "C" int pastix_amd_solve_refine_device(pastix_data_t *d, pastix_int_t nrhs, const double *d_b,
                                              pastix_int_t ldb, double *d_x, pastix_int_t ldx) {
    return refine_device(d, nrhs, d_b, ldb, d_x, ldx, true);
}
// ---- residual SpMV and refinement vector kernels ----------------------------
static bool engine_ready(const pastix_data_t *d, const char *fn) {
    if (!d || !d->engine || !d->engine->ready() || !(d->steps & STEP_BCSC2CTAB)) {
        err_print(fn, "pastix_task_numfact() has to be called before");
        return false;
    }
    return true;
}

extern "C" int pastix_amd_bcsc_spmv(pastix_data_t *d, pastix_trans_t trans, double alpha, const double *x,
                                    double beta, double *y) {
    if (!engine_ready(d, "pastix_amd_bcsc_spmv")) return PASTIX_ERR_BADPARAMETER;
    if (!x || !y) return PASTIX_ERR_BADPARAMETER;
    // A^T = A for the symmetric storage types
    if (trans == PastixNoTrans || d->mtxtype == SpmSymmetric || d->mtxtype == SpmHermitian)
        return d->engine->spmv(alpha, x, beta, y) || hipStreamSynchronize(d->engine->stream()) != hipSuccess
                   ? PASTIX_ERR_INTERNAL
                   : PASTIX_SUCCESS;
    if (trans != PastixTrans && trans != PastixConjTrans) return PASTIX_ERR_BADPARAMETER;
    if (!d->engine->has_transpose() && d->engine->upload_transpose(d->bcsc)) return PASTIX_ERR_INTERNAL;
    if (d->engine->spmv_t(alpha, x, beta, y) || hipStreamSynchronize(d->engine->stream()) != hipSuccess)
        return PASTIX_ERR_INTERNAL;
    return PASTIX_SUCCESS;
}
extern "C" int pastix_amd_bvec_dot(pastix_data_t *d, pastix_int_t n, const double *x, const double *y,
                                   double *result) {
    if (!engine_ready(d, "pastix_amd_bvec_dot") || n < 0 || (n > 0 && (!x || !y)) || !result)
        return PASTIX_ERR_BADPARAMETER;
    return d->engine->vec_dot(n, x, y, result) ? PASTIX_ERR_INTERNAL : PASTIX_SUCCESS;
}
extern "C" int pastix_amd_bvec_nrm2(pastix_data_t *d, pastix_int_t n, const double *x, double *result) {
    if (!engine_ready(d, "pastix_amd_bvec_nrm2") || n < 0 || (n > 0 && !x) || !result) return PASTIX_ERR_BADPARAMETER;
    return d->engine->vec_nrm2(n, x, result) ? PASTIX_ERR_INTERNAL : PASTIX_SUCCESS;
}
extern "C" int pastix_amd_bvec_axpy(pastix_data_t *d, pastix_int_t n, double alpha, const double *x, double *y) {
    if (!engine_ready(d, "pastix_amd_bvec_axpy") || n < 0 || (n > 0 && (!x || !y))) return PASTIX_ERR_BADPARAMETER;
    return d->engine->vec_axpy(n, alpha, x, y) ? PASTIX_ERR_INTERNAL : PASTIX_SUCCESS;
}
extern "C" int pastix_amd_bvec_scal(pastix_data_t *d, pastix_int_t n, double alpha, double *x) {
    if (!engine_ready(d, "pastix_amd_bvec_scal") || n < 0 || (n > 0 && !x)) return PASTIX_ERR_BADPARAMETER;
    return d->engine->vec_scal(n, alpha, x) ? PASTIX_ERR_INTERNAL : PASTIX_SUCCESS;
}
extern "C" int pastix_amd_bvec_copy(pastix_data_t *d, pastix_int_t n, const double *x, double *y) {
    if (!engine_ready(d, "pastix_amd_bvec_copy") || n < 0 || (n > 0 && (!x || !y))) return PASTIX_ERR_BADPARAMETER;
    return d->engine->vec_copy(n, x, y) ? PASTIX_ERR_INTERNAL : PASTIX_SUCCESS;
}
extern "C" int pastix_amd_bvec_gemv(pastix_data_t *d, pastix_int_t m, pastix_int_t n, double alpha, const double *A,
                                    pastix_int_t lda, const double *x, double beta, double *y) {
    if (!engine_ready(d, "pastix_amd_bvec_gemv") || m < 0 || n < 0 || lda < std::max<pastix_int_t>(1, m) ||
        (m > 0 && (!y || (n > 0 && (!A || !x)))))
        return PASTIX_ERR_BADPARAMETER;
    return d->engine->vec_gemv(m, n, alpha, A, lda, x, beta, y) ? PASTIX_ERR_INTERNAL : PASTIX_SUCCESS;
}

extern "C" int pastix_amd_flops(const pastix_data_t *d, double out[4]) {
    if (!d || !(d->steps & STEP_ANALYSE)) return PASTIX_ERR_BADPARAMETER;
    const bool lu = d->iparm[IPARM_FACTORIZATION] == PastixFactGETRF || is_ldlt(d);
    out[0] = d->dparm[DPARM_FACT_THFLOPS];
    out[1] = d->plan.flops_diag;
    out[2] = d->plan.flops_trsm * (lu ? 2.0 : 1.0);
    out[3] = lu ? d->plan.flops_upd_lu : d->plan.flops_upd_potrf;
    return PASTIX_SUCCESS;
}
// Per-level statistics of the update schedule (5 doubles per level): tiles,
// entries, useful GEMM flops, MFMA flops issued (32x32 wave quads, K rounded
// to 4), source width sum.  Returns the number of levels written.
extern "C" int pastix_amd_plan_stats(const pastix_data_t *d, double *out, int maxlevels) {
    if (!d || !(d->steps & STEP_ANALYSE)) return -1;
    const pxa::Plan &P = d->plan;
    const pxa::Symbol &s = d->symbol;
    int nl = std::min<int>(maxlevels, P.nlevels);
    for (int lv = 0; lv < nl; lv++) {
        double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = P.lv_upd_ptr[lv]; t < P.lv_upd_ptr[lv + 1]; t++) {
            const pxa::UpdItem &it = P.upd[t];
            const bool direct = t < P.lv_upd_split[lv];
            st[direct ? 0 : 2] += 1;
            for (int e = it.e0; e < it.e1; e++) {
                const pxa::UpdEntry &en = P.ents[e];
                const double K = (double)(s.cblk_fcol[en.s + 1] - s.cblk_fcol[en.s]);
                const double np = en.p1 - en.p0, nq = en.q1 - en.q0;
                const double fl = 2.0 * np * nq * K;
                if (direct) st[1] += fl;
                else {
                    st[3] += 1;
                    st[4] += fl;
                }
                const double qp = std::ceil(np / 32.0) * 32.0, qq = std::ceil(nq / 32.0) * 32.0;
                st[5] += 2.0 * qp * qq * std::ceil(K / 4.0) * 4.0;
                st[6] += K;
            }
        }
        for (int k = 0; k < 8; k++) out[8 * lv + k] = st[k];
    }
    return nl;
}

extern "C" int pastix_amd_refactorize_shift(pastix_data_t *d, double delta) {
    if (!d || !d->engine || !(d->steps & STEP_BCSC2CTAB)) return PASTIX_ERR_BADPARAMETER;
    pxa::Engine &E = *d->engine;
    if (E.shift_diagonal(delta)) return PASTIX_ERR_BADPARAMETER;
    // ||A + delta I||_F for the static-pivot threshold
    const double nf = d->bcsc.normF;
    d->dparm[DPARM_A_NORM] = std::sqrt(std::max(0.0, nf * nf + 2.0 * delta * E.diag_sum() + (double)d->n * delta * delta));
    engine_flags(d);
    return pastix_subtask_sopalin(d);
}

extern "C" int pastix_amd_refactorize(pastix_data_t *d) {
    if (!d || !d->engine || !(d->steps & STEP_BCSC2CTAB)) return PASTIX_ERR_BADPARAMETER;
    if (d->engine->shift() != 0.0) return pastix_amd_refactorize_shift(d, 0.0);
    if (d->engine->restore_coef()) return PASTIX_ERR_BADPARAMETER;
    engine_flags(d);
    return pastix_subtask_sopalin(d);
}
extern "C" int pastix_amd_backward_error(pastix_data_t *d, const double *b, const double *x, double *berr) {
    if (!d || !b || !x || !berr || !d->engine) return PASTIX_ERR_BADPARAMETER;
    pxa::Engine &E = *d->engine;
    const int64_t n = d->n;
    double *db = device_pair(d);
    if (!db) return PASTIX_ERR_OUTOFMEMORY;
    std::vector<double> hb(n), hx(n);
    for (int64_t i = 0; i < n; i++) {
        hb[d->order.permtab[i]] = b[i];
        hx[d->order.permtab[i]] = x[i];
    }
    if (hipMemcpy(db, hb.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db + n, hx.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return PASTIX_ERR_INTERNAL;
    double o[3];
    if (E.berr(db + n, db, o)) return PASTIX_ERR_INTERNAL;
    double den = d->bcsc.normInf * o[1] + o[2];
    *berr = den > 0 ? o[0] / den : o[0];
    return PASTIX_SUCCESS;
}
// No reference counterpart: pastix_amd_backward_error for callers that cannot
// allocate (HIP graph capture, serving loops).  Device columns in the user
// ordering, results into a device array; the partial maxima live in the
// library-static scratch of the device, so at most 64 columns per call.
extern "C" int pastix_amd_backward_error_device_noalloc(pastix_data_t *d, pastix_int_t nrhs, const double *d_b,
                                                        pastix_int_t ldb, const double *d_x, pastix_int_t ldx,
                                                        double *d_berr) {
    if (!d || !d->engine || !(d->steps & STEP_BCSC2CTAB)) {
        err_print("pastix_amd_backward_error_device_noalloc", "the matrix has not been factorised");
        return PASTIX_ERR_BADPARAMETER;
    }
    if (nrhs < 1 || nrhs > pxa::NOALLOC_MAX_RHS) {
        err_print("pastix_amd_backward_error_device_noalloc",
                  "nrhs must be between 1 and 64 (the size of the library's static scratch)");
        return PASTIX_ERR_BADPARAMETER;
    }
    if (!d_b || !d_x || !d_berr || ldb < d->n || ldx < d->n) return PASTIX_ERR_BADPARAMETER;
    if (d->engine->berr_noalloc(d_x, ldx, d_b, ldb, (int)nrhs, d->bcsc.normInf, d_berr)) return PASTIX_ERR_INTERNAL;
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_kernel_times(const pastix_data_t *d, double out[12]) {
    if (!d || !d->engine) return PASTIX_ERR_BADPARAMETER;
    for (int c = 0; c < pxa::Engine::KC_N; c++) {
        out[2 * c] = d->engine->kt[c];
        out[2 * c + 1] = d->engine->kn[c];
    }
    return PASTIX_SUCCESS;
}

// Rebuild the level schedule with new per-level choices (no reference
// counterpart: the tuner's and the tuning table's hook).  The next numfact
// sets the engine up again.
static int replan(pastix_data_t *d) {
    d->engine.reset();
    d->steps &= ~(STEP_BCSC2CTAB | STEP_NUMFACT | STEP_SOLVE | STEP_REFINE);
    if (const int prc = make_plan(d, d->schur_cblk, "pastix_amd_replan")) {
        d->steps &= ~STEP_ANALYSE;
        return prc;
    }
    return PASTIX_SUCCESS;
}
extern "C" int pastix_amd_set_level_knobs(pastix_data_t *d, int nlevels, const int8_t *single_pass,
                                          const int8_t *panel_right) {
    if (!d || nlevels < 0 || (nlevels > 0 && (!single_pass || !panel_right))) return PASTIX_ERR_BADPARAMETER;
    d->knobs.single_pass.assign(single_pass, single_pass + nlevels);
    d->knobs.panel_right.assign(panel_right, panel_right + nlevels);
    return (d->steps & STEP_ANALYSE) ? replan(d) : PASTIX_SUCCESS;
}
extern "C" int pastix_amd_load_tuning(pastix_data_t *d, const char *path, int *applied) {
    if (!d || !path || !applied) return PASTIX_ERR_BADPARAMETER;
    *applied = 0;
    if (!(d->steps & STEP_ANALYSE)) return PASTIX_ERR_BADPARAMETER;
    pxa::PlanKnobs k;
    const int rc = read_tuning(d, path, k);
    if (rc < 0) return PASTIX_ERR_FILE;
    if (rc == 0) return PASTIX_SUCCESS;
    k.node_max = d->knobs.node_max;  // the table holds per-level choices only
    d->knobs = k;
    *applied = 1;
    return replan(d);
}
extern "C" int pastix_amd_save_tuning(const pastix_data_t *d, const char *path) {
    if (!d || !path || !(d->steps & STEP_ANALYSE)) return PASTIX_ERR_BADPARAMETER;
    const std::string key = "entry " + tuning_key(d);
    std::vector<std::string> keep;
    if (FILE *f = fopen(path, "r")) {
        char line[1 << 16];
        bool skip = false;
        while (fgets(line, sizeof(line), f)) {
            std::string l(line);
            while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
            if (l == key) skip = true;
            if (!skip) keep.push_back(l);
            if (skip && l == "end") skip = false;
        }
        fclose(f);
    }
    FILE *f = fopen(path, "w");
    if (!f) return PASTIX_ERR_FILE;
    for (const std::string &l : keep) fprintf(f, "%s\n", l.c_str());
    const pxa::Plan &P = d->plan;
    fprintf(f, "%s\nsingle_pass", key.c_str());
    for (int lv = 0; lv < P.nlevels; lv++) fprintf(f, " %d", (int)P.lv_single_pass[lv]);
    fprintf(f, "\npanel_right");
    for (int lv = 0; lv < P.nlevels; lv++) fprintf(f, " %d", (int)P.lv_panel_right[lv]);
    fprintf(f, "\nend\n");
    return fclose(f) == 0 ? PASTIX_SUCCESS : PASTIX_ERR_FILE;
}
// Per level: cblks, multi-source update tiles, single-pass (1) or product
// workspace (0), right-looking (1) or left-looking (0) panels.
extern "C" int pastix_amd_level_info(const pastix_data_t *d, int32_t *out, int maxlevels) {
    if (!d || !out || !(d->steps & STEP_ANALYSE)) return -1;
    const pxa::Plan &P = d->plan;
    const int nl = std::min<int>(maxlevels, P.nlevels);
    for (int lv = 0; lv < nl; lv++) {
        out[4 * lv] = P.lv_cblk_ptr[lv + 1] - P.lv_cblk_ptr[lv];
        out[4 * lv + 1] = P.lv_upd_ptr[lv + 1] - P.lv_upd_split[lv];
        out[4 * lv + 2] = P.lv_single_pass[lv];
        out[4 * lv + 3] = P.lv_panel_right[lv];
    }
    return nl;
}
// Solve structure (no reference counterpart: the solve nodes are this
// engine's): out[0] nodes, [1] nodes of several cblks, [2] solve levels,
// [3] merge rounds of the nodes' inverses (summed over the batches), [4] their tiles, [5] entries of
// the inverses of the nodes of several cblks, [6] products of the merge
// rounds (flops, one side), [7] factorisation levels.
extern "C" int pastix_amd_solve_info(const pastix_data_t *d, double out[8]) {
    if (!d || !out || !(d->steps & STEP_ANALYSE)) return PASTIX_ERR_BADPARAMETER;
    const pxa::Plan &P = d->plan;
    const int64_t nn = (int64_t)P.node_cblk.size() - 1;
    int64_t groups = 0;
    for (int64_t k = 0; k < nn; k++) groups += (P.node_cblk[k + 1] - P.node_cblk[k]) > 1;
    out[0] = (double)nn;
    out[1] = (double)groups;
    out[2] = (double)P.nslevels;
    out[3] = (double)(P.winv_round_ptr.empty() ? 0 : P.winv_round_ptr.size() - 1);
    out[4] = (double)P.winv_tiles.size();
    out[5] = (double)(P.winvnbr - P.winv_off.back());
    out[6] = P.flops_winv;
    out[7] = (double)P.nlevels;
    return PASTIX_SUCCESS;
}
// Per level of the last numfact run with PXA_IPARM_PROFILE: panel steps and
// Schur updates (seconds).
extern "C" int pastix_amd_level_times(const pastix_data_t *d, double *out, int maxlevels) {
    if (!d || !out || !d->engine) return -1;
    const pxa::Engine &E = *d->engine;
    const int nl = std::min<int>(maxlevels, (int)E.lv_panel_s.size());
    for (int lv = 0; lv < nl; lv++) {
        out[2 * lv] = E.lv_panel_s[lv];
        out[2 * lv + 1] = E.lv_upd_s[lv];
    }
    return nl;
}

extern "C" int pastix_amd_update_flops(const pastix_data_t *d, double out[8]) {
    if (!d || !(d->steps & STEP_ANALYSE)) return PASTIX_ERR_BADPARAMETER;
    const pxa::Plan &P = d->plan;
    out[0] = P.flops_kind[0][0];
    out[1] = P.flops_kind[0][1];
    out[2] = P.flops_kind[1][0];
    out[3] = P.flops_kind[1][1];
    out[4] = P.flops_wtile;
    out[5] = P.flops_kind[0][2];
    out[6] = P.flops_kind[1][2];
    out[7] = 0;
    if (leaf_tasks_on(d)) {
        // the region's levels run as one persistent launch: their tiles leave
        // the level kernels' counts, its updates above it join them
        out[0] += P.flops_leaf_out[0] - P.flops_kind_leaf[0][0];
        out[1] -= P.flops_kind_leaf[0][1];
        out[4] -= P.flops_wtile_leaf;
        out[5] += P.flops_leaf_out[1] - P.flops_kind_leaf[0][2];
        out[7] = P.flops_leaf_upd;
    }
    return PASTIX_SUCCESS;
}

extern "C" int pastix_amd_fact_times(const pastix_data_t *d, double out[8]) {
    if (!d) return PASTIX_ERR_BADPARAMETER;
    for (int i = 0; i < 8; i++) out[i] = d->fact_times[i];
    return PASTIX_SUCCESS;
}
