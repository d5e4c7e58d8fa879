// Host analysis: ordering, symbolic factorisation, splitting, solver layout.
//
// The reference keeps this work on the CPU (order/, symbol/, blend/); so does
// this engine.  The algorithms are restated from the reference behaviour cited
// on each function; data structures are this engine's own (std::vector based,
// 1D panel layout for the device).
#include "analysis.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <cstdio>
#include <functional>
#include <numeric>
#include <queue>

namespace pxa {

// ---------------------------------------------------------------------------
// Flop model.  Formulas of reference common/flops.h:47-90,195-252 (note the
// reference defines FADDS_TRSM as FMULS_TRMM, so a TRSM counts 2*FMULS_TRMM),
// combined per cblk as in symbol/symbol_cost.c:53-88 (sum1d) with the
// per-factorisation tables of symbol/symbol_cost_flops.c.
// ---------------------------------------------------------------------------
static double fl_potrf(double n) {
    return n * (((1. / 6.) * n + 0.5) * n + (1. / 3.)) + n * (((1. / 6.) * n) * n - (1. / 6.));
}
static double fl_trsm_right(double m, double n) {  // FLOPS_DTRSM(PastixRight, m, n)
    return 2.0 * (0.5 * m * n * (n + 1.));
}
static double fl_syrk(double k, double n) { return 2.0 * (0.5 * k * n * (n + 1.)); }
static double fl_gemm(double m, double n, double k) { return 2.0 * m * n * k; }
static double fl_getrf(double m, double n) {
    double mul, add;
    if (m < n) {
        mul = 0.5 * m * (m * (n - (1. / 3.) * m - 1.) + n) + (2. / 3.) * m;
        add = 0.5 * m * (m * (n - (1. / 3.) * m) - n) + (1. / 6.) * m;
    } else {
        mul = 0.5 * n * (n * (m - (1. / 3.) * n - 1.) + m) + (2. / 3.) * n;
        add = 0.5 * n * (n * (m - (1. / 3.) * n) - m) + (1. / 6.) * n;
    }
    return mul + add;
}

double symbol_thflops(const Symbol &s, int factotype) {
    double total = 0.;
    for (int64_t c = 0; c < s.cblknbr(); c++) {
        double N = (double)(s.cblk_fcol[c + 1] - s.cblk_fcol[c]);
        double M = 0.;
        for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++)
            M += (double)(s.blok_lrow[b] - s.blok_frow[b] + 1);
        if (factotype == 2) {  // GETRF
            total += fl_getrf(N, N);
            if (M > 0) total += 2. * fl_trsm_right(M, N) + fl_gemm(M, M, N);
        } else if (factotype == 1) {  // SYTRF / LDL^t (symbol_cost_flops.c:85-88,142-144,199-201)
            total += N * ((N / 6. + 0.5) * N + 10. / 3.) + N * ((N / 6.) * N - 1. / 6.);
            if (M > 0) total += fl_trsm_right(M, N) + N * M + fl_syrk(N, M) + M * M;
        } else {               // POTRF / LLT
            total += fl_potrf(N);
            if (M > 0) total += fl_trsm_right(M, N) + fl_syrk(N, M);
        }
    }
    return total;
}

int64_t symbol_nnz(const Symbol &s) {
    int64_t nnz = 0;
    for (int64_t c = 0; c < s.cblknbr(); c++) {
        int64_t w = s.cblk_fcol[c + 1] - s.cblk_fcol[c];
        nnz += (w * (w + 1)) / 2 - w;
        for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++)
            nnz += (s.blok_lrow[b] - s.blok_frow[b] + 1) * w;
    }
    return nnz;
}

// ---------------------------------------------------------------------------
// Graph
// ---------------------------------------------------------------------------
Graph graph_from_csc(int64_t n, const int64_t *colptr, const int64_t *rows, int64_t base) {
    // Symmetrise the pattern and drop the diagonal.
    std::vector<int64_t> deg(n, 0);
    for (int64_t j = 0; j < n; j++)
        for (int64_t k = colptr[j] - base; k < colptr[j + 1] - base; k++) {
            int64_t i = rows[k] - base;
            if (i == j || i < 0 || i >= n) continue;
            deg[i]++;
            deg[j]++;
        }
    Graph g;
    g.n = n;
    g.colptr.assign(n + 1, 0);
    for (int64_t j = 0; j < n; j++) g.colptr[j + 1] = g.colptr[j] + deg[j];
    std::vector<int64_t> pos(g.colptr.begin(), g.colptr.end() - 1);
    g.rows.resize(g.colptr[n]);
    for (int64_t j = 0; j < n; j++)
        for (int64_t k = colptr[j] - base; k < colptr[j + 1] - base; k++) {
            int64_t i = rows[k] - base;
            if (i == j || i < 0 || i >= n) continue;
            g.rows[pos[i]++] = j;
            g.rows[pos[j]++] = i;
        }
    // sort + unique per column
    int64_t w = 0;
    std::vector<int64_t> newptr(n + 1, 0);
    for (int64_t j = 0; j < n; j++) {
        auto b = g.rows.begin() + g.colptr[j], e = g.rows.begin() + g.colptr[j + 1];
        std::sort(b, e);
        int64_t start = w;
        int64_t prev = -1;
        for (auto it = b; it != e; ++it) {
            if (*it != prev) g.rows[w++] = *it;
            prev = *it;
        }
        newptr[j] = start;
    }
    newptr[n] = w;
    g.colptr = newptr;
    g.rows.resize(w);
    return g;
}

// ---------------------------------------------------------------------------
// Ordering helpers shared by the grid and the graph nested dissection: the
// recursion emits supernodes in pre-order (separator first) with their depth
// and numbers vertices downwards from n-1; the result is then turned into an
// increasing rangtab and a father-based treetab exactly as the reference does
// at order/order_grids.c:348-380 (father = first later supernode with a
// smaller depth).
// ---------------------------------------------------------------------------
namespace {
struct NDBuilder {
    int64_t n;
    int64_t maxnum;
    std::vector<int64_t> permtab;
    std::vector<int64_t> rang_hi;  // highest number of each emitted supernode (pre-order)
    std::vector<int64_t> depth;
    explicit NDBuilder(int64_t n_) : n(n_), maxnum(n_ - 1), permtab(n_, -1) {}
    void open(int64_t d) {
        rang_hi.push_back(maxnum);
        depth.push_back(d);
    }
    void number(int64_t v) { permtab[v] = maxnum--; }
    Order finish() {
        Order o;
        o.n = n;
        int64_t cb = (int64_t)rang_hi.size();
        o.cblknbr = cb;
        o.permtab = permtab;
        o.peritab.assign(n, 0);
        for (int64_t i = 0; i < n; i++) o.peritab[o.permtab[i]] = i;
        o.rangtab.assign(cb + 1, 0);
        o.treetab.assign(cb, 0);
        for (int64_t i = 0; i < cb; i++) {
            o.rangtab[i + 1] = rang_hi[cb - i - 1] + 1;
            o.treetab[i] = depth[cb - i - 1];
        }
        for (int64_t i = 0; i < cb; i++) {
            int64_t fath = -1;
            for (int64_t j = i + 1; j < cb; j++)
                if (o.treetab[j] < o.treetab[i]) {
                    fath = j;
                    break;
                }
            o.treetab[i] = fath;
        }
        return o;
    }
};
}  // namespace

// Reference order/order_grids.c:137-303 (order_grid3D_classic) and
// :62-75 (order_grid2D_classic): cut the largest dimension at its middle
// plane (ties favour x, then y), number the plane first with the highest
// numbers, recurse into the lower half then the upper half; stop below 15
// vertices.  Vertex index = i + nx*j + nx*ny*k.
Order order_grid(int64_t nx, int64_t ny, int64_t nz) {
    int64_t n = nx * ny * nz;
    NDBuilder b(n);
    const int64_t ldx = nx, ldxy = nx * ny;
    std::function<void(int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t)> rec =
        [&](int64_t x0, int64_t xn, int64_t y0, int64_t yn, int64_t z0, int64_t zn, int64_t d) {
            int64_t dx = xn - x0, dy = yn - y0, dz = zn - z0;
            if (dx * dy * dz < 15) {
                b.open(d);
                for (int64_t i = x0; i < xn; i++)
                    for (int64_t j = y0; j < yn; j++)
                        for (int64_t k = z0; k < zn; k++) b.number(i + ldx * j + ldxy * k);
                return;
            }
            int dir = 0;
            if (dy > dx) dir = 1;
            if (dz > dx && dz > dy) dir = 2;
            b.open(d);
            if (dir == 0) {
                int64_t xs = x0 + dx / 2;
                for (int64_t j = y0; j < yn; j++)
                    for (int64_t k = z0; k < zn; k++) b.number(xs + ldx * j + ldxy * k);
                rec(x0, xs, y0, yn, z0, zn, d + 1);
                rec(xs + 1, xn, y0, yn, z0, zn, d + 1);
            } else if (dir == 1) {
                int64_t ys = y0 + dy / 2;
                for (int64_t i = x0; i < xn; i++)
                    for (int64_t k = z0; k < zn; k++) b.number(i + ldx * ys + ldxy * k);
                rec(x0, xn, y0, ys, z0, zn, d + 1);
                rec(x0, xn, ys + 1, yn, z0, zn, d + 1);
            } else {
                int64_t zs = z0 + dz / 2;
                for (int64_t i = x0; i < xn; i++)
                    for (int64_t j = y0; j < yn; j++) b.number(i + ldx * j + ldxy * zs);
                rec(x0, xn, y0, yn, z0, zs, d + 1);
                rec(x0, xn, y0, yn, zs + 1, zn, d + 1);
            }
        };
    rec(0, nx, 0, ny, 0, nz, 1);
    return b.finish();
}

// Level-structure nested dissection (George's automatic nested dissection):
// pseudo-peripheral vertex, breadth-first levels, the middle level becomes the
// separator.  Disconnected pieces are handled as separate subtrees.
Order order_nested_dissection(const Graph &g) {
    const int64_t n = g.n;
    NDBuilder b(n);
    std::vector<int64_t> mark(n, -1), level(n, -1);
    int64_t stamp = 0;
    const int64_t leafsize = 32;

    // BFS inside the vertex set flagged with `tag` in mark[]; returns vertices in BFS order.
    auto bfs = [&](int64_t root, int64_t tag, std::vector<int64_t> &out, int64_t &nlev) {
        out.clear();
        out.push_back(root);
        level[root] = 0;
        size_t h = 0;
        nlev = 1;
        while (h < out.size()) {
            int64_t v = out[h++];
            for (int64_t k = g.colptr[v]; k < g.colptr[v + 1]; k++) {
                int64_t w = g.rows[k];
                if (mark[w] != tag || level[w] >= 0) continue;
                level[w] = level[v] + 1;
                nlev = std::max(nlev, level[w] + 1);
                out.push_back(w);
            }
        }
    };

    std::function<void(std::vector<int64_t> &, int64_t)> rec = [&](std::vector<int64_t> &V, int64_t d) {
        if ((int64_t)V.size() < leafsize) {
            b.open(d);
            for (int64_t v : V) b.number(v);
            return;
        }
        int64_t tag = ++stamp;
        for (int64_t v : V) {
            mark[v] = tag;
            level[v] = -1;
        }
        std::vector<int64_t> order;
        int64_t nlev = 0;
        bfs(V[0], tag, order, nlev);
        if (order.size() < V.size()) {
            // Disconnected: split into the reached component and the rest.
            std::vector<int64_t> A(order), B;
            for (int64_t v : V)
                if (level[v] < 0) B.push_back(v);
            for (int64_t v : V) level[v] = -1;
            rec(B, d);
            rec(A, d);
            return;
        }
        // Pseudo-peripheral vertex: restart from the last vertex a few times.
        for (int it = 0; it < 3; it++) {
            int64_t far = order.back();
            for (int64_t v : V) level[v] = -1;
            int64_t nl2 = 0;
            bfs(far, tag, order, nl2);
            if (nl2 <= nlev) { nlev = nl2; break; }
            nlev = nl2;
        }
        if (nlev < 3) {
            b.open(d);
            for (int64_t v : V) b.number(v);
            for (int64_t v : V) level[v] = -1;
            return;
        }
        std::vector<int64_t> cnt(nlev, 0);
        for (int64_t v : V) cnt[level[v]]++;
        int64_t half = (int64_t)V.size() / 2, acc = 0, sep = 1;
        for (int64_t l = 0; l < nlev; l++) {
            acc += cnt[l];
            if (acc >= half) { sep = std::max<int64_t>(1, std::min<int64_t>(l, nlev - 2)); break; }
        }
        std::vector<int64_t> S, A, B;
        for (int64_t v : order) {
            if (level[v] == sep) S.push_back(v);
            else if (level[v] < sep) A.push_back(v);
            else B.push_back(v);
        }
        for (int64_t v : V) level[v] = -1;
        b.open(d);
        for (int64_t v : S) b.number(v);
        rec(A, d + 1);
        rec(B, d + 1);
    };
    std::vector<int64_t> all(n);
    std::iota(all.begin(), all.end(), 0);
    if (n > 0) rec(all, 1);
    return b.finish();
}

// Fundamental supernodes of the permuted matrix: elimination tree (Liu),
// column counts by row-subtree traversal, then merge j-1 into j's supernode
// when j is the only child of j-1's parent chain link and the counts nest
// (reference order/order_find_supernodes.c builds the same partition).
void order_find_supernodes(const Graph &g, Order &o) {
    const int64_t n = g.n;
    std::vector<int64_t> parent(n, -1), anc(n, -1);
    for (int64_t j = 0; j < n; j++) {
        int64_t oj = o.peritab[j];
        for (int64_t k = g.colptr[oj]; k < g.colptr[oj + 1]; k++) {
            int64_t i = o.permtab[g.rows[k]];
            if (i >= j) continue;
            // walk from i to its root with path compression
            int64_t r = i;
            while (anc[r] != -1 && anc[r] != j) {
                int64_t nx = anc[r];
                anc[r] = j;
                r = nx;
            }
            if (anc[r] == -1) {
                anc[r] = j;
                parent[r] = j;
            }
        }
    }
    std::vector<int64_t> colcnt(n, 0), mark(n, -1), nchild(n, 0);
    for (int64_t i = 0; i < n; i++) {
        int64_t oi = o.peritab[i];
        mark[i] = i;
        for (int64_t k = g.colptr[oi]; k < g.colptr[oi + 1]; k++) {
            int64_t j = o.permtab[g.rows[k]];
            if (j >= i) continue;
            for (int64_t t = j; t != -1 && t < i && mark[t] != i; t = parent[t]) {
                colcnt[t]++;
                mark[t] = i;
            }
        }
    }
    for (int64_t j = 0; j < n; j++)
        if (parent[j] >= 0) nchild[parent[j]]++;
    std::vector<int64_t> rang;
    std::vector<int64_t> col2sn(n);
    for (int64_t j = 0; j < n; j++) {
        bool start = (j == 0) || parent[j - 1] != j || nchild[j] != 1 || colcnt[j - 1] != colcnt[j] + 1;
        if (start) rang.push_back(j);
        col2sn[j] = (int64_t)rang.size() - 1;
    }
    int64_t cb = (int64_t)rang.size();
    rang.push_back(n);
    o.cblknbr = cb;
    o.rangtab = rang;
    o.treetab.assign(cb, -1);
    for (int64_t s = 0; s < cb; s++) {
        int64_t last = rang[s + 1] - 1;
        o.treetab[s] = parent[last] >= 0 ? col2sn[parent[last]] : -1;
    }
}

// Reference order/order_apply_level_order.c:81-333.  Breadth-first walk of
// the elimination tree from the roots, sons pushed in reverse so that
// brothers keep their order; supernodes are then numbered backwards from
// the end of the walk.  With level_tasks2d < 0 (the reference default) the
// walk keeps two queues: the "2D" supernodes, at least width_tasks2d columns
// wide together with their brothers and every ancestor and its brothers
// (the PASTIX_BLEND_DEEPEST_DISTRIB marking, :155-201), come first in the
// walk, so they take the last numbers; the others follow.
void order_apply_level_order(Order &o, int64_t level_tasks2d, int64_t width_tasks2d) {
    const int64_t cb = o.cblknbr;
    if (cb == 0) return;
    // elimination tree; index cb stands for the virtual root above the roots
    std::vector<int64_t> fath(cb);
    std::vector<std::vector<int64_t>> sons(cb + 1);
    for (int64_t i = 0; i < cb; i++) {
        fath[i] = o.treetab[i] < 0 ? cb : o.treetab[i];
        sons[fath[i]].push_back(i);
    }
    std::vector<int64_t> sorted(cb);
    std::vector<int64_t> newfath(cb, -1);
    if (level_tasks2d < 0) {
        std::vector<char> is2d(cb + 1, 0);
        int64_t tot2d = 0;
        auto mark_sons = [&](int64_t f) {
            for (int64_t s : sons[f])
                if (!is2d[s]) {
                    is2d[s] = 1;
                    tot2d++;
                }
        };
        for (int64_t node = 0; node < cb; node++) {
            if (is2d[node]) continue;
            const int64_t size = o.rangtab[node + 1] - o.rangtab[node];
            if (size < width_tasks2d) continue;
            int64_t f = fath[node];
            mark_sons(f);
            while (f != cb && !is2d[f]) {
                f = fath[f];
                mark_sons(f);
            }
        }
        int64_t p2 = 0, p1 = tot2d;
        for (int64_t r : sons[cb]) sorted[is2d[r] ? p2++ : p1++] = r;
        for (int64_t i = 0; i < cb; i++) {
            const int64_t node = sorted[i];
            const auto &sn = sons[node];
            int64_t n2 = 0;
            for (int64_t s : sn) n2 += is2d[s];
            const int64_t n1 = (int64_t)sn.size() - n2;
            int64_t c2 = 0, c1 = 0;
            for (int64_t s : sn) {
                if (is2d[s]) sorted[p2 + n2 - (c2++) - 1] = s;
                else sorted[p1 + n1 - (c1++) - 1] = s;
                newfath[s] = cb - i - 1;
            }
            p2 += n2;
            p1 += n1;
        }
    } else {
        int64_t pos = 0;
        for (int64_t r : sons[cb]) sorted[pos++] = r;
        for (int64_t i = 0; i < cb; i++) {
            const int64_t node = sorted[i];
            const auto &sn = sons[node];
            for (size_t k = 0; k < sn.size(); k++) {
                sorted[pos + sn.size() - 1 - k] = sn[k];
                newfath[sn[k]] = cb - i - 1;
            }
            pos += (int64_t)sn.size();
        }
    }
    std::vector<int64_t> rang(cb + 1, 0), tree(cb), peri(o.n);
    for (int64_t i = 0; i < cb; i++) {
        int64_t node = sorted[cb - i - 1];
        int64_t size = o.rangtab[node + 1] - o.rangtab[node];
        rang[i + 1] = rang[i] + size;
        tree[i] = newfath[node];
        std::copy(o.peritab.begin() + o.rangtab[node], o.peritab.begin() + o.rangtab[node + 1],
                  peri.begin() + rang[i]);
    }
    o.rangtab = rang;
    o.treetab = tree;
    o.peritab = peri;
    for (int64_t i = 0; i < o.n; i++) o.permtab[o.peritab[i]] = i;
}

// ---------------------------------------------------------------------------
// Symbol reordering (reference symbol/symbol_reordering.c:40-660): inside each
// supernode, rows are reordered so that rows touched by the same
// contributing supernodes become adjacent and the off-diagonal blocks facing
// the supernode fewer and taller.  Each row of supernode c carries the set of
// supernodes with a block facing c over that row, split in an "upper" set
// (contributors no deeper than a split level of the elimination tree) and a
// "lower" set; rows are inserted one at a time into a tour, at the position
// that adds the least Hamming distance between lower sets, among the
// positions whose neighbours' upper sets agree.
// ---------------------------------------------------------------------------
namespace {
using SetVec = std::vector<std::vector<int64_t>>;

// reference hamming_distance (:99-162): size of the symmetric difference of
// two sorted sets, capped at stop; -1 is a ghost row with the empty set
int64_t hamming(const SetVec &v, int64_t xi, int64_t xj, int64_t stop) {
    if (xi == -1) return (int64_t)v[xj].size();
    if (xj == -1) return (int64_t)v[xi].size();
    const std::vector<int64_t> &a = v[xi], &b = v[xj];
    const int64_t sa = (int64_t)a.size(), sb = (int64_t)b.size();
    if (sa - sb >= stop) return stop;
    if (sb - sa >= stop) return stop;
    int64_t sum = 0, i = 0, j = 0;
    while (i < sa && j < sb) {
        if (a[i] == b[j]) {
            i++;
            j++;
        } else if (a[i] < b[j]) {
            while (i < sa && a[i] < b[j]) {
                sum++;
                i++;
            }
        } else {
            while (j < sb && a[i] > b[j]) {
                sum++;
                j++;
            }
        }
        if (sum >= stop) return stop;
    }
    sum += (sa - i) + (sb - j);
    return sum >= stop ? stop : sum;
}

// reference symbol_reorder_tsp (:212-437): peri points at the supernode's
// columns in the ordering's inverse permutation
void reorder_tsp(int64_t size, int64_t *peri, const SetVec &lw, const SetVec &up, int64_t stop) {
    if (size < 3) return;
    std::vector<int64_t> tinvp(size + 1), tlen(size + 1, 0);
    tinvp[0] = -1;
    tinvp[1] = 0;
    const int64_t distance = hamming(lw, 0, -1, stop);
    tlen[0] = distance;
    tlen[1] = distance;
    for (int64_t i = 1; i < size; i++) {
        int64_t lw_before = hamming(lw, i, tinvp[0], stop);
        int64_t lw_after = hamming(lw, i, tinvp[1], stop);
        int64_t up_after = hamming(up, i, tinvp[1], 1);
        int64_t minl = lw_before + lw_after - tlen[0];
        int64_t mpos = 1, min_cut = -1;
        for (int64_t j = 1; j < i; j++) {
            const int64_t up_before = up_after;
            up_after = hamming(up, i, tinvp[j + 1], 1);
            if (up_before < 1 || up_after < 1) {
                // a skipped position leaves the previous distance unknown
                lw_before = (lw_after == -1) ? hamming(lw, i, tinvp[j], stop) : lw_after;
                lw_after = hamming(lw, i, tinvp[j + 1], stop);
                const int64_t l = lw_before + lw_after - tlen[j];
                // the same tour length: the smaller cut between two rows
                if (l == minl) {
                    if (lw_before < min_cut) {
                        min_cut = lw_before;
                        minl = l;
                        mpos = j + 1;
                    }
                    if (lw_after < min_cut) {
                        min_cut = lw_after;
                        minl = l;
                        mpos = j + 1;
                    }
                }
                if (l < minl) {
                    min_cut = lw_before;
                    minl = l;
                    mpos = j + 1;
                    if (lw_after < min_cut) min_cut = lw_after;
                }
                // two equal rows: next to each other, search ends
                if (lw_after == 0) {
                    min_cut = 0;
                    minl = l;
                    mpos = j + 1;
                    j = i;
                }
            } else {
                lw_after = -1;
            }
        }
        // between the last and the first
        const int64_t first_pos = hamming(lw, i, tinvp[0], stop);
        const int64_t last_pos = hamming(lw, i, tinvp[i], stop);
        lw_before = hamming(lw, i, tinvp[mpos - 1], stop);
        lw_after = hamming(lw, i, tinvp[mpos], stop);
        const int64_t l = first_pos + last_pos - tlen[i];
        if (l < minl) {
            minl = l;
            mpos = i + 1;
        }
        if (mpos > 0) tlen[mpos - 1] = lw_before;
        if (mpos < i + 1) {
            int64_t k = i, ll = lw_after;
            for (int64_t j = mpos; j < i + 2; j++) {
                std::swap(tinvp[j], k);
                std::swap(tlen[j], ll);
            }
        } else {
            tinvp[i + 1] = i;
            tlen[i + 1] = first_pos;
        }
    }
    // the tour starts after the ghost row (searched among the first size
    // positions, as the reference does)
    int64_t elected = 0;
    for (int64_t i = 0; i < size; i++)
        if (tinvp[i] == -1) elected = i;
    std::vector<int64_t> sn(size);
    for (int64_t i = 0; i < size; i++) sn[i] = peri[tinvp[(i + 1 + elected) % (size + 1)]];
    std::copy(sn.begin(), sn.end(), peri);
}
}  // namespace

void symbol_reordering(const Symbol &s, Order &o, int64_t split_level, int64_t stop, int64_t schurfcol) {
    const int64_t cb = s.cblknbr();
    if (cb != o.cblknbr) return;
    // depth of every supernode in the elimination tree, roots at 1 (:46-62)
    std::vector<int64_t> levels(cb, 0);
    int64_t maxdepth = 0;
    for (int64_t c = cb - 1; c >= 0; c--) {
        levels[c] = o.treetab[c] < 0 ? 1 : levels[o.treetab[c]] + 1;
        maxdepth = std::max(maxdepth, levels[c]);
    }
    // blocks facing each cblk, in increasing order of their owner (browtab)
    std::vector<int64_t> fptr(cb + 1, 0);
    for (int64_t c = 0; c < cb; c++)
        for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++) fptr[s.blok_fcblk[b] + 1]++;
    for (int64_t c = 0; c < cb; c++) fptr[c + 1] += fptr[c];
    std::vector<int64_t> fown(fptr[cb]), fblk(fptr[cb]);
    {
        std::vector<int64_t> pos(fptr.begin(), fptr.end() - 1);
        for (int64_t c = 0; c < cb; c++)
            for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++) {
                const int64_t f = s.blok_fcblk[b];
                fown[pos[f]] = c;
                fblk[pos[f]++] = b;
            }
    }
    std::vector<int64_t> depthweight(std::max<int64_t>(maxdepth, 1));
    for (int64_t c = 0; c < cb; c++) {
        const int64_t fc = s.cblk_fcol[c], size = s.cblk_fcol[c + 1] - fc;
        if (fc >= schurfcol) continue;
        std::fill(depthweight.begin(), depthweight.end(), 0);
        for (int64_t k = fptr[c]; k < fptr[c + 1]; k++)
            depthweight[levels[fown[k]] - 1] += s.blok_lrow[fblk[k]] - s.blok_frow[fblk[k]] + 1;
        // the split level, moved so that the upper set holds about a fifth
        // of the contributions (:467-503)
        int64_t lsl = split_level;
        {
            int sign = 0;
            for (;;) {
                int64_t up_total = 0, lw_total = 0, i = 0;
                for (; i < lsl && i < maxdepth; i++) up_total += depthweight[i];
                for (; i < maxdepth; i++) lw_total += depthweight[i];
                if (lw_total < 5 * up_total && lw_total > 10 && up_total > 10 && sign <= 0) {
                    lsl--;
                    sign--;
                    continue;
                }
                if (lw_total > 3 * up_total && lw_total > 10 && up_total > 10 && sign >= 0) {
                    lsl++;
                    sign++;
                    continue;
                }
                break;
            }
        }
        SetVec up(size), lw(size);
        for (int64_t k = fptr[c]; k < fptr[c + 1]; k++) {
            const int64_t b = fblk[k];
            SetVec &v = (levels[fown[k]] <= lsl) ? up : lw;
            for (int64_t r = s.blok_frow[b]; r <= s.blok_lrow[b]; r++) v[r - fc].push_back(fown[k]);
        }
        reorder_tsp(size, o.peritab.data() + o.rangtab[c], lw, up, stop);
    }
    for (int64_t i = 0; i < o.n; i++) o.permtab[o.peritab[i]] = i;
}

bool check_order(const Order &o, std::string &err) {
    std::vector<char> seen(o.n, 0);
    for (int64_t i = 0; i < o.n; i++) {
        int64_t p = o.permtab[i];
        if (p < 0 || p >= o.n || seen[p]) { err = "permtab is not a permutation"; return false; }
        seen[p] = 1;
        if (o.peritab[p] != i) { err = "peritab is not the inverse of permtab"; return false; }
    }
    if ((int64_t)o.rangtab.size() != o.cblknbr + 1 || o.rangtab[0] != 0 || o.rangtab[o.cblknbr] != o.n) {
        err = "rangtab does not cover [0,n)";
        return false;
    }
    for (int64_t c = 0; c < o.cblknbr; c++) {
        if (o.rangtab[c + 1] <= o.rangtab[c]) { err = "empty supernode"; return false; }
        if (o.treetab[c] != -1 && o.treetab[c] <= c) { err = "treetab father must follow its son"; return false; }
    }
    return true;
}

// ---------------------------------------------------------------------------
// Symbolic factorisation with dense diagonal blocks.  For supernode c with
// columns [f,l]: rows(c) = { perm(i) > l : A(i, col) != 0, col in c }
//                        U { rows(s) > l : s child of c },
// where the father of s is the supernode holding min(rows(s)).  Blocks are the
// maximal row runs that stay inside one supernode (the structure built by the
// reference symbol/symbol_fax_direct.c + pastixSymbolBuildRowtab).
// ---------------------------------------------------------------------------
Symbol symbol_fax(const Graph &g, const Order &o) {
    const int64_t n = o.n, cb = o.cblknbr;
    std::vector<int64_t> col2cblk(n);
    for (int64_t c = 0; c < cb; c++)
        for (int64_t j = o.rangtab[c]; j < o.rangtab[c + 1]; j++) col2cblk[j] = c;
    std::vector<std::vector<int64_t>> pending(cb);  // children structures waiting for their father
    Symbol s;
    s.n = n;
    s.cblk_fcol.assign(o.rangtab.begin(), o.rangtab.end());
    s.cblk_bloknum.assign(cb + 1, 0);
    std::vector<int64_t> rows;
    for (int64_t c = 0; c < cb; c++) {
        int64_t f = o.rangtab[c], l = o.rangtab[c + 1] - 1;
        rows.clear();
        for (int64_t j = f; j <= l; j++) {
            int64_t oj = o.peritab[j];
            for (int64_t k = g.colptr[oj]; k < g.colptr[oj + 1]; k++) {
                int64_t i = o.permtab[g.rows[k]];
                if (i > l) rows.push_back(i);
            }
        }
        for (int64_t r : pending[c])
            if (r > l) rows.push_back(r);
        std::vector<int64_t>().swap(pending[c]);
        std::sort(rows.begin(), rows.end());
        rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
        // blocks
        s.cblk_bloknum[c] = s.bloknbr();
        s.blok_frow.push_back(f);
        s.blok_lrow.push_back(l);
        s.blok_fcblk.push_back(c);
        for (size_t k = 0; k < rows.size(); k++) {
            int64_t r = rows[k];
            if (k == 0 || r != rows[k - 1] + 1 || col2cblk[r] != col2cblk[rows[k - 1]]) {
                s.blok_frow.push_back(r);
                s.blok_lrow.push_back(r);
                s.blok_fcblk.push_back(col2cblk[r]);
            } else {
                s.blok_lrow.back() = r;
            }
        }
        if (!rows.empty()) {
            int64_t fath = col2cblk[rows[0]];
            auto &p = pending[fath];
            p.insert(p.end(), rows.begin(), rows.end());
        }
    }
    s.cblk_bloknum[cb] = s.bloknbr();
    return s;
}

// Reference blend/splitsymbol.c:79-100 (number of pieces: ceil(width/max),
// minus one if the pieces would fall under min) and :380-470 (walk the
// supernode, recomputing the piece count on the remaining width).  Cut
// positions use the constant step of computeConstantSplit; the reference's
// default cut search (computeSmallestSplit over blocks per line) is not
// restated.  Every split piece keeps a dense block facing each later piece,
// and every block facing a split supernode is cut at the new boundaries.
// Cut rules of reference blend/splitsymbol.c: computeSmallestSplit (the
// default, :195-241), computeSmallestSplit_max (PASTIX_BLEND_SPLIT=UPPER,
// :265-299) and computeConstantSplit (PASTIX_BLEND_SPLIT=CONSTANT,
// :301-313); nbpl points at the cblk's first column, step is the even cut,
// the result is the last column of the piece relative to it.
static int64_t split_smallest(const int64_t *nbpl, int64_t step, int64_t mn, int64_t mx, int64_t pct) {
    const int64_t limit = (step * pct + 99) / 100;
    if (step >= mx) return mx - 1;
    const int64_t lavg = step - 1;
    const int64_t lmin = std::max(lavg - limit - 1, mn), lmax = std::min(lavg + limit + 1, mx);
    int64_t lcol = lavg, best = nbpl[lavg];
    for (int64_t i = lavg + 1; i < lmax; i++)
        if (nbpl[i] < best) {
            lcol = i;
            best = nbpl[i];
        }
    for (int64_t i = lavg - 1; i > lmin; i--)
        if (nbpl[i] < best) {
            lcol = i;
            best = nbpl[i];
        }
    return lcol;
}
static int64_t split_smallest_upper(const int64_t *nbpl, int64_t step, int64_t mn, int64_t mx, int64_t pct) {
    const int64_t limit = (step * pct + 99) / 100;
    if (step >= mx) return mx - 1;
    const int64_t lavg = step - 1;
    const int64_t lmin = std::max(lavg - limit, mn), lmax = std::min(lavg + limit + 1, mx);
    int64_t lcol = lmin, best = nbpl[lmin];
    for (int64_t i = lmin; i < lmax; i++)
        if (nbpl[i] <= best) {
            lcol = i;
            best = nbpl[i];
        }
    return lcol;
}
static int64_t split_constant(const int64_t *, int64_t step, int64_t, int64_t mx, int64_t) {
    return step >= mx ? mx - 1 : step - 1;
}

// Reference blend/splitsymbol.c:79-100 (computeNbSplit) and splitSmart
// (:316-441): each cblk wider than blcolmax is cut into pieces, the cut
// before each piece chosen by the rule above from nblocksperline, the
// number of symbol blocks spanning the boundary between a row and the next
// (:104-139; counted once, then raised by pieces-1 on the rows of a split
// cblk's off-diagonal blocks, which the split repeats).
Symbol symbol_split(const Symbol &s, int64_t blcolmin, int64_t blcolmax) {
    const int64_t cb = s.cblknbr();
    auto nbsplit = [&](int64_t width) -> int64_t {
        if (width <= blcolmax) return 1;
        int64_t nseq = (width + blcolmax - 1) / blcolmax;
        if (nseq > 1 && width / nseq < blcolmin) nseq--;
        return std::max<int64_t>(nseq, 1);
    };
    int64_t (*rule)(const int64_t *, int64_t, int64_t, int64_t, int64_t) = split_smallest;
    if (const char *e = getenv("PASTIX_BLEND_SPLIT")) {
        if (!strcmp(e, "CONSTANT")) rule = split_constant;
        else if (!strcmp(e, "UPPER")) rule = split_smallest_upper;
    }
    int64_t pct = 10;
    if (const char *e = getenv("PASTIX_BLEND_SPLIT_AUTORIZED_PERCENTAGE")) pct = atoll(e);
    std::vector<int64_t> nbpl;  // per absolute row, from the first split cblk on
    int64_t fsplit = -1;
    // New column partition.
    std::vector<int64_t> newf;               // first column of each new cblk
    std::vector<int64_t> first_new(cb + 1);  // first new cblk of each old cblk
    for (int64_t c = 0; c < cb; c++) {
        first_new[c] = (int64_t)newf.size();
        const int64_t f = s.cblk_fcol[c], l = s.cblk_fcol[c + 1] - 1;
        int64_t width = l - f + 1;
        if (nbsplit(width) <= 1) {
            newf.push_back(f);
            continue;
        }
        if (fsplit < 0) {
            fsplit = f;
            nbpl.assign(s.n + 1, 0);
            for (int64_t b = 0; b < s.bloknbr(); b++) {
                if (s.blok_lrow[b] < fsplit) continue;
                for (int64_t line = std::max(s.blok_frow[b], fsplit); line < s.blok_lrow[b]; line++) nbpl[line]++;
            }
        }
        int64_t fc = f, pieces = 0;
        while (fc <= l) {
            const int64_t nseq = nbsplit(width);
            newf.push_back(fc);
            pieces++;
            if (nseq == 1) break;
            const int64_t step = (width + nseq - 1) / nseq;
            const int64_t lwidth = width - blcolmin;
            const int64_t lc = fc + rule(nbpl.data() + fc, step, blcolmin, lwidth, pct);
            width -= lc - fc + 1;
            fc = lc + 1;
        }
        for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++)
            for (int64_t line = s.blok_frow[b]; line < s.blok_lrow[b]; line++) nbpl[line] += pieces - 1;
    }
    first_new[cb] = (int64_t)newf.size();
    const int64_t ncb = (int64_t)newf.size();
    newf.push_back(s.n);
    std::vector<int64_t> col2new(s.n);
    for (int64_t c = 0; c < ncb; c++)
        for (int64_t j = newf[c]; j < newf[c + 1]; j++) col2new[j] = c;

    Symbol t;
    t.n = s.n;
    t.cblk_fcol = newf;
    t.cblk_bloknum.assign(ncb + 1, 0);
    auto push_range = [&](int64_t fr, int64_t lr) {  // split a row range at new cblk boundaries
        int64_t r = fr;
        while (r <= lr) {
            int64_t c = col2new[r];
            int64_t e = std::min(lr, newf[c + 1] - 1);
            t.blok_frow.push_back(r);
            t.blok_lrow.push_back(e);
            t.blok_fcblk.push_back(c);
            r = e + 1;
        }
    };
    for (int64_t c = 0; c < cb; c++) {
        for (int64_t k = first_new[c]; k < first_new[c + 1]; k++) {
            t.cblk_bloknum[k] = t.bloknbr();
            t.blok_frow.push_back(newf[k]);
            t.blok_lrow.push_back(newf[k + 1] - 1);
            t.blok_fcblk.push_back(k);
            if (k + 1 < first_new[c + 1]) push_range(newf[k + 1], s.cblk_fcol[c + 1] - 1);
            for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++)
                push_range(s.blok_frow[b], s.blok_lrow[b]);
        }
    }
    t.cblk_bloknum[ncb] = t.bloknbr();
    return t;
}

bool check_symbol(const Symbol &s, std::string &err) {
    for (int64_t c = 0; c < s.cblknbr(); c++) {
        int64_t b0 = s.cblk_bloknum[c];
        if (s.blok_frow[b0] != s.cblk_fcol[c] || s.blok_lrow[b0] != s.cblk_fcol[c + 1] - 1) {
            err = "first block is not the diagonal block";
            return false;
        }
        for (int64_t b = b0 + 1; b < s.cblk_bloknum[c + 1]; b++) {
            if (s.blok_frow[b] <= s.blok_lrow[b - 1]) { err = "blocks overlap or are unsorted"; return false; }
            int64_t f = s.blok_fcblk[b];
            if (f <= c || s.blok_frow[b] < s.cblk_fcol[f] || s.blok_lrow[b] >= s.cblk_fcol[f + 1]) {
                err = "block does not face a single later cblk";
                return false;
            }
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// Layout + browtab + dependency levels.  ctrbcnt of the reference
// (blend/solver_backup.c:95-101) is the browtab size of each cblk; level(c)
// is 1 + the largest level of the cblks updating c, so every level is a set
// of cblks whose updates have all been applied by lower levels.
// ---------------------------------------------------------------------------
Layout build_layout(const Symbol &s) {
    const int64_t cb = s.cblknbr(), nb = s.bloknbr();
    Layout L;
    L.cblk_stride.assign(cb, 0);
    L.cblk_coef.assign(cb + 1, 0);
    L.blok_lcblk.assign(nb, 0);
    L.blok_roff.assign(nb, 0);
    for (int64_t c = 0; c < cb; c++) {
        int64_t w = s.cblk_fcol[c + 1] - s.cblk_fcol[c];
        int64_t roff = 0;
        for (int64_t b = s.cblk_bloknum[c]; b < s.cblk_bloknum[c + 1]; b++) {
            L.blok_lcblk[b] = c;
            L.blok_roff[b] = roff;
            roff += s.blok_lrow[b] - s.blok_frow[b] + 1;
        }
        L.cblk_stride[c] = roff;
        L.cblk_coef[c + 1] = L.cblk_coef[c] + roff * w;
        L.maxwidth = std::max(L.maxwidth, w);
        L.maxstride = std::max(L.maxstride, roff);
    }
    L.coefnbr = L.cblk_coef[cb];
    // browtab
    std::vector<int64_t> cnt(cb + 1, 0);
    for (int64_t c = 0; c < cb; c++)
        for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++) cnt[s.blok_fcblk[b] + 1]++;
    L.cblk_brow.assign(cb + 1, 0);
    for (int64_t c = 0; c < cb; c++) L.cblk_brow[c + 1] = L.cblk_brow[c] + cnt[c + 1];
    L.browtab.assign(L.cblk_brow[cb], 0);
    std::vector<int64_t> pos(L.cblk_brow.begin(), L.cblk_brow.end() - 1);
    for (int64_t c = 0; c < cb; c++)
        for (int64_t b = s.cblk_bloknum[c] + 1; b < s.cblk_bloknum[c + 1]; b++)
            L.browtab[pos[s.blok_fcblk[b]]++] = b;
    // levels
    L.cblk_level.assign(cb, 0);
    int32_t maxl = -1;
    for (int64_t c = 0; c < cb; c++) {
        int32_t lv = 0;
        for (int64_t k = L.cblk_brow[c]; k < L.cblk_brow[c + 1]; k++)
            lv = std::max(lv, L.cblk_level[L.blok_lcblk[L.browtab[k]]] + 1);
        L.cblk_level[c] = lv;
        maxl = std::max(maxl, lv);
    }
    L.nlevels = maxl + 1;
    return L;
}

}  // namespace pxa
