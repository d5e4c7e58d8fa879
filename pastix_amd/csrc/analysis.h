// Host analysis of the MI355X engine: ordering, supernodal symbolic
// factorization, supernode splitting and the flattened solver layout that the
// device engine and its level scheduler consume.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pxa {

// Symmetrised adjacency graph of the user matrix, 0-based CSC without the
// diagonal (reference graph/graph_prepare.c builds the same object).
struct Graph {
    int64_t n = 0;
    std::vector<int64_t> colptr;  // n+1
    std::vector<int64_t> rows;    // sorted per column
};

// Ordering: the same fields as pastix_order_t, owned by std::vector.
struct Order {
    int64_t n = 0;
    int64_t cblknbr = 0;
    std::vector<int64_t> permtab;  // old -> new
    std::vector<int64_t> peritab;  // new -> old
    std::vector<int64_t> rangtab;  // cblknbr+1
    std::vector<int64_t> treetab;  // cblknbr, father or -1
};

// Block symbol (reference symbol/symbol.h: symbol_cblk_t / symbol_blok_t).
// Row indices are in the permuted numbering.
struct Symbol {
    int64_t n = 0;
    std::vector<int64_t> cblk_fcol;    // cblknbr+1, last entry = n
    std::vector<int64_t> cblk_bloknum; // cblknbr+1, first blok of each cblk is its diagonal block
    std::vector<int64_t> blok_frow;    // first row (inclusive)
    std::vector<int64_t> blok_lrow;    // last row (inclusive)
    std::vector<int64_t> blok_fcblk;   // facing cblk
    int64_t cblknbr() const { return (int64_t)cblk_fcol.size() - 1; }
    int64_t bloknbr() const { return (int64_t)blok_frow.size(); }
};

// Flattened solver layout: one column-major panel per cblk ("1D layout",
// stride x width, ld = stride), every block a contiguous row range of it.
struct Layout {
    std::vector<int64_t> cblk_stride;  // rows in the panel
    std::vector<int64_t> cblk_coef;    // cblknbr+1 element offsets into the coefficient array
    std::vector<int64_t> blok_lcblk;   // owner cblk
    std::vector<int64_t> blok_roff;    // row offset of the block inside its panel
    std::vector<int64_t> cblk_brow;    // cblknbr+1, ranges into browtab
    std::vector<int64_t> browtab;      // blocks facing each cblk, sorted by owner cblk
    std::vector<int32_t> cblk_level;   // dependency level (0 = no incoming update)
    int32_t nlevels = 0;
    int64_t coefnbr = 0;
    int64_t maxwidth = 0, maxstride = 0;
};

// Theoretical factorisation flops (reference symbol/symbol_cost.c:53-88,399-420).
double symbol_thflops(const Symbol &s, int factotype);
// Strictly-lower nonzeros of the factor (reference symbol/symbol_cost.c:327-361).
int64_t symbol_nnz(const Symbol &s);

Graph graph_from_csc(int64_t n, const int64_t *colptr, const int64_t *rows, int64_t baseval);

// Grid nested dissection (reference order/order_grids.c:137-383).
Order order_grid(int64_t nx, int64_t ny, int64_t nz);
// Level-structure nested dissection for general graphs, used when no external
// ordering library is present (Scotch/METIS are not in this build).
Order order_nested_dissection(const Graph &g);
// Supernode partition and tree for an ordering that has none
// (reference order/order_find_supernodes.c, fundamental supernodes).
void order_find_supernodes(const Graph &g, Order &o);
// Breadth-first renumbering of supernodes by level (reference
// order/order_apply_level_order.c:81-333): level_tasks2d < 0 puts the
// supernodes of at least width_tasks2d columns, their brothers and their
// ancestors' families in a queue of their own that takes the last numbers.
void order_apply_level_order(Order &o, int64_t level_tasks2d, int64_t width_tasks2d);

// Symbol reordering (reference symbol/symbol_reordering.c:617-660): the rows
// of every supernode before schurfcol are reordered (o.peritab and permtab
// change inside each supernode's range) to compact the blocks facing it;
// s is the symbolic factorisation of o, before splitting.
void symbol_reordering(const Symbol &s, Order &o, int64_t split_level, int64_t stop, int64_t schurfcol);

// Supernodal symbolic factorisation with dense diagonal blocks
// (reference symbol/symbol_fax_direct.c).
Symbol symbol_fax(const Graph &g, const Order &o);
// Split supernodes wider than blcolmax (reference blend/splitsymbol.c:79-100,380-470).
Symbol symbol_split(const Symbol &s, int64_t blcolmin, int64_t blcolmax);
// Layout, browtab and dependency levels.
Layout build_layout(const Symbol &s);

bool check_order(const Order &o, std::string &err);
bool check_symbol(const Symbol &s, std::string &err);

} // namespace pxa
