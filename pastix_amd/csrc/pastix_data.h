// Internal state of one pastix instance (the opaque pastix_data_t of the
// C API) and helpers shared by the API translation units.
#pragma once
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "analysis.h"
#include "engine.h"
#include "plan.h"

extern "C" {
#include "pastix.h"
#include "pastix_amd.h"
}

#define STEP_INIT (1 << 0)
#define STEP_ORDERING (1 << 1)
#define STEP_SYMBFACT (1 << 2)
#define STEP_ANALYSE (1 << 3)
#define STEP_CSC2BCSC (1 << 4)
#define STEP_BCSC2CTAB (1 << 5)
#define STEP_NUMFACT (1 << 6)
#define STEP_SOLVE (1 << 7)
#define STEP_REFINE (1 << 8)

struct pastix_data_s {
    pastix_int_t *iparm = nullptr;
    double *dparm = nullptr;
    pastix_int_t steps = 0;
    pastix_int_t ext[PXA_IPARM_SIZE] = {0};
    // analysis
    pxa::Graph graph;
    pxa::Order order;
    pxa::Symbol symbol;
    pxa::Layout layout;
    pxa::Plan plan;
    int64_t n = 0;
    int64_t nnz_in = 0;  // stored entries of the analysed input (DPARM_FILL_IN)
    // Schur complement: the unknowns given to pastixSetSchurUnknownList (as
    // given, in the input's numbering base), ordered last by the analysis;
    // permuted columns [schurfcol, n) are the Schur part (schurfcol = n: none)
    std::vector<pastix_int_t> schur_list;
    std::string dir_global;  // output directory of the analysis files (IPARM_IO_STRATEGY)
    int64_t schurfcol = 0;
    int mtxtype = SpmGeneral;
    int flttype = SpmDouble;
    const spmatrix_t *csc = nullptr;
    // numerics
    pxa::HostBcsc bcsc;
    std::unique_ptr<pxa::Engine> engine;
    double fact_times[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double *dvec = nullptr;  // 2n FP64 device vectors for refinement / error checks
    pxa::Engine *dvec_owner = nullptr;
    // batch of independent systems analysed as one block-diagonal matrix:
    // system s owns rows and columns [batch_off[s], batch_off[s+1]) of the
    // union, in the user and in the permuted numbering alike
    std::vector<int64_t> batch_off;
    std::vector<spm_int_t> u_colptr, u_rows;
    std::vector<double> u_vals;
    spmatrix_t u_spm{};
    // per-level schedule choices from a tuning table or the tuner
    // (pastix_amd_set_level_knobs / pastix_amd_load_tuning)
    pxa::PlanKnobs knobs;
    int64_t schur_cblk = 0;  // first Schur cblk of the analysed partition
    double *batch_dev = nullptr;  // 2 x n x R FP64 blocks of the batched refinement
    int64_t batch_dev_cap = 0;
    pxa::Engine *batch_dev_owner = nullptr;
};

// Triplets (0-based, symmetric input expanded to both triangles).
struct Triplets {
    std::vector<int64_t> i, j;
    std::vector<double> v;
    std::vector<int64_t> src;  // index of each triplet's value in the input
};

namespace pxa_api {
double *device_pair(pastix_data_t *d);
void err_print(const char *fn, const char *msg);
int64_t spm_baseval(const spmatrix_t *spm);
// analysis files (ordergen / symbgen formats); base: numbering base written
int order_read(FILE *f, pxa::Order &o, int64_t &base);
int order_write(FILE *f, const pxa::Order &o, int64_t base);
int symbol_read(FILE *f, pxa::Symbol &s);
int symbol_write(FILE *f, const pxa::Symbol &s);
const char *output_dir(pastix_data_t *d);
// writes the instance's ordering (order) or symbol into its output directory
int save_analysis_file(pastix_data_t *d, const char *name, bool order);
// sets the engine's Schur solve mode from IPARM_SCHUR_SOLV_MODE; false (with a
// message) when the mode asks to solve with the unfactored Schur part
bool schur_mode(pastix_data_t *d, const char *fn);
int spm_triplets(const spmatrix_t *spm, Triplets &t, bool values);
// adjacency graph of the input's pattern (symmetrised, without the diagonal)
int spm_graph(const spmatrix_t *spm, pxa::Graph &g);
// esrc (optional): for every output entry the triplets summed into it,
// entry e owning (*esrc)[(*eptr)[e] .. (*eptr)[e+1]) in summation order
void triplets_to_csc(int64_t n, const Triplets &t, const int64_t *perm, std::vector<int64_t> &colptr,
                     std::vector<int32_t> &rows, std::vector<double> &vals, bool transpose,
                     std::vector<int64_t> *eptr = nullptr, std::vector<int64_t> *esrc = nullptr);
// Batched CG refinement of n x R user-ordered FP64 blocks (see batch.cpp).
// Device FP64 block of at least count entries, kept by the instance between
// calls (grown on demand).
double *device_block(pastix_data_t *d, int64_t count);
// frees the block device_block() holds
void release_device_block(pastix_data_t *d);
// Refinement method code of the instance for Engine::refine_user: 1 CG, 0 GMRES.
int refine_method(const pastix_data_t *d);
// LDL^t requested (PastixFactSYTRF / PastixFactHETRF, the same in real arithmetic)
inline bool is_ldlt(const pastix_data_t *d) {
    return d->iparm[IPARM_FACTORIZATION] == PastixFactSYTRF || d->iparm[IPARM_FACTORIZATION] == PastixFactHETRF;
}
double refine_eps(const pastix_data_t *d);
// the leaf region runs as one persistent launch: LL^t, a region analysed,
// PXA_IPARM_LEAF_TASKS on
inline bool leaf_tasks_on(const pastix_data_t *d) {
    return d->ext[PXA_IPARM_LEAF_TASKS] != 0 && d->iparm[IPARM_FACTORIZATION] != PastixFactGETRF && !is_ldlt(d) &&
           d->plan.leaf_levels > 0;
}
// Host FP64 blocks (n x R, ld n, user ordering of the instance): X = M^{-1} B
// when initial_solve (else X holds the guesses), then batched CG over the
// row segments segoff (method 1) or nothing (method -1).
int refine_block(pastix_data_t *d, int R, const double *B, double *X, const std::vector<int64_t> &segoff,
                 bool initial_solve, int method);
}  // namespace pxa_api


