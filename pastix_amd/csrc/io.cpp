// Ordering and symbol files of the analysis (reference order/order_io.c:42-330,
// symbol/symbol_io.c:50-196, common/integer.c intLoad/intSave, and the
// output directory of common/api.c:65-131).  Text files of whitespace-
// separated integers:
//   ordergen: version (1), cblknbr, vertnbr; rangtab (cblknbr+1); permtab
//             (vertnbr); treetab (cblknbr, version 1 only); the base is
//             rangtab[0];
//   symbgen:  version (1), cblknbr, bloknbr, nodenbr, baseval; per cblk
//             fcolnum lcolnum bloknum; per block frownum lrownum fcblknm.
// Writers use the reference's layout (tabs, a line feed after every eighth
// value of a long array) so the files read the same in either code.
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <string>

#include "pastix_data.h"

using namespace pxa_api;

namespace {
bool read_int(FILE *f, int64_t &v) {
    long long x;
    if (fscanf(f, "%lld", &x) != 1) return false;
    v = (int64_t)x;
    return true;
}

// writes a[0..n) with the reference's separators (order_io.c:236-256)
bool write_array(FILE *f, const int64_t *a, int64_t n) {
    for (int64_t k = 0; k < n; k++)
        if (fprintf(f, "%lld%c", (long long)a[k], (k == n - 1) ? '\n' : (((k & 7) == 7) ? '\n' : '\t')) < 0)
            return false;
    return true;
}
}  // namespace

int pxa_api::order_read(FILE *f, pxa::Order &o, int64_t &base) {
    int64_t vers = 0, cb = 0, n = 0;
    if (!read_int(f, vers) || !read_int(f, cb) || !read_int(f, n) || (vers != 0 && vers != 1) || cb > n || cb < 0)
        return PASTIX_ERR_FILE;
    o = pxa::Order();
    o.n = n;
    o.cblknbr = cb;
    o.rangtab.resize(cb + 1);
    o.permtab.resize(n);
    o.peritab.resize(n);
    for (int64_t c = 0; c <= cb; c++)
        if (!read_int(f, o.rangtab[c])) return PASTIX_ERR_FILE;
    for (int64_t i = 0; i < n; i++)
        if (!read_int(f, o.permtab[i])) return PASTIX_ERR_FILE;
    if (vers == 1) {
        o.treetab.resize(cb);
        for (int64_t c = 0; c < cb; c++)
            if (!read_int(f, o.treetab[c])) return PASTIX_ERR_FILE;
    }
    base = o.rangtab.empty() ? 0 : o.rangtab[0];
    for (int64_t &v : o.rangtab) v -= base;
    for (int64_t &v : o.permtab) {
        v -= base;
        if (v < 0 || v >= n) return PASTIX_ERR_FILE;
    }
    for (int64_t &v : o.treetab) v = (v < base) ? -1 : v - base;
    for (int64_t i = 0; i < n; i++) o.peritab[o.permtab[i]] = i;
    return PASTIX_SUCCESS;
}

int pxa_api::order_write(FILE *f, const pxa::Order &o, int64_t base) {
    if ((int64_t)o.rangtab.size() != o.cblknbr + 1 || (int64_t)o.treetab.size() != o.cblknbr) return PASTIX_ERR_BADPARAMETER;
    std::vector<int64_t> r(o.rangtab), p(o.permtab), t(o.treetab);
    for (int64_t &v : r) v += base;
    for (int64_t &v : p) v += base;
    for (int64_t &v : t) v = (v < 0) ? -1 : v + base;
    if (fprintf(f, "1\n%lld\t%lld\n", (long long)o.cblknbr, (long long)o.n) < 0) return PASTIX_ERR_FILE;
    if (!write_array(f, r.data(), (int64_t)r.size()) || !write_array(f, p.data(), (int64_t)p.size()) ||
        !write_array(f, t.data(), (int64_t)t.size()))
        return PASTIX_ERR_FILE;
    return PASTIX_SUCCESS;
}

int pxa_api::symbol_read(FILE *f, pxa::Symbol &s) {
    int64_t vers = 0, cb = 0, bb = 0, n = 0, base = 0;
    if (!read_int(f, vers) || !read_int(f, cb) || !read_int(f, bb) || !read_int(f, n) || !read_int(f, base) ||
        vers < 0 || vers > 1 || bb < cb || n < cb)
        return PASTIX_ERR_FILE;
    s = pxa::Symbol();
    s.n = n;
    s.cblk_fcol.resize(cb + 1);
    s.cblk_bloknum.resize(cb + 1);
    for (int64_t c = 0; c < cb; c++) {
        int64_t fc, lc, bn;
        if (!read_int(f, fc) || !read_int(f, lc) || !read_int(f, bn) || fc > lc) return PASTIX_ERR_FILE;
        s.cblk_fcol[c] = fc - base;
        s.cblk_bloknum[c] = bn - base;
        if (c + 1 == cb && lc - base + 1 != n) return PASTIX_ERR_FILE;
    }
    s.cblk_fcol[cb] = n;
    s.cblk_bloknum[cb] = bb;
    s.blok_frow.resize(bb);
    s.blok_lrow.resize(bb);
    s.blok_fcblk.resize(bb);
    for (int64_t b = 0; b < bb; b++) {
        int64_t fr, lr, fc;
        if (!read_int(f, fr) || !read_int(f, lr) || !read_int(f, fc) || fr > lr) return PASTIX_ERR_FILE;
        s.blok_frow[b] = fr - base;
        s.blok_lrow[b] = lr - base;
        s.blok_fcblk[b] = fc - base;
    }
    return PASTIX_SUCCESS;
}

int pxa_api::symbol_write(FILE *f, const pxa::Symbol &s) {
    const int64_t cb = s.cblknbr(), bb = s.bloknbr();
    if (fprintf(f, "1\n%lld\t%lld\t%lld\t%lld\n", (long long)cb, (long long)bb, (long long)s.n, 0LL) < 0)
        return PASTIX_ERR_FILE;
    for (int64_t c = 0; c < cb; c++)
        if (fprintf(f, "%lld\t%lld\t%lld\n", (long long)s.cblk_fcol[c], (long long)(s.cblk_fcol[c + 1] - 1),
                    (long long)s.cblk_bloknum[c]) < 0)
            return PASTIX_ERR_FILE;
    for (int64_t b = 0; b < bb; b++)
        if (fprintf(f, "%lld\t%lld\t%lld\n", (long long)s.blok_frow[b], (long long)s.blok_lrow[b],
                    (long long)s.blok_fcblk[b]) < 0)
            return PASTIX_ERR_FILE;
    return PASTIX_SUCCESS;
}

// The instance's output directory, created on first use as pastix-XXXXXX in
// the working directory (reference common/api.c:65-131, mkdtemp branch).
const char *pxa_api::output_dir(pastix_data_t *d) {
    if (d->dir_global.empty()) {
        char tmpl[] = "pastix-XXXXXX";
        const mode_t old = umask(S_IWGRP | S_IWOTH);
        char *r = mkdtemp(tmpl);
        umask(old);
        if (!r) {
            perror("pastix_gendirectories/global/mkdtemp");
            return nullptr;
        }
        d->dir_global = r;
        fprintf(stdout, "OUTPUTDIR: %s\n", r);
    }
    return d->dir_global.c_str();
}

// Reference order/order_io.c:132-181: the file named by PASTIX_FILE_ORDER,
// else "ordername".
extern "C" int pastixOrderLoad(const pastix_data_t *, pastix_order_t *ordeptr) {
    if (!ordeptr) return PASTIX_ERR_BADPARAMETER;
    const char *env = getenv("PASTIX_FILE_ORDER");
    const char *name = env ? env : "ordername";
    FILE *f = fopen(name, "r");
    if (!f) {
        err_print("pastixOrderLoad", "cannot open the ordering file");
        return PASTIX_ERR_FILE;
    }
    pxa::Order o;
    int64_t base = 0;
    const int rc = order_read(f, o, base);
    fclose(f);
    if (rc) {
        err_print("pastixOrderLoad", "bad input");
        return rc;
    }
    pastixOrderAlloc(ordeptr, o.n, o.cblknbr);
    ordeptr->baseval = base;
    for (int64_t i = 0; i < o.n; i++) {
        ordeptr->permtab[i] = o.permtab[i] + base;
        ordeptr->peritab[i] = o.peritab[i] + base;
    }
    for (int64_t c = 0; c <= o.cblknbr && ordeptr->rangtab; c++) ordeptr->rangtab[c] = o.rangtab[c] + base;
    if (o.treetab.empty()) {
        free(ordeptr->treetab);
        ordeptr->treetab = nullptr;
    } else {
        for (int64_t c = 0; c < o.cblknbr; c++) ordeptr->treetab[c] = o.treetab[c] < 0 ? -1 : o.treetab[c] + base;
    }
    return PASTIX_SUCCESS;
}

// Reference order/order_io.c:294-330: "ordergen" (or PASTIX_FILE_ORDER) in
// the instance's output directory.
extern "C" int pastixOrderSave(pastix_data_t *d, const pastix_order_t *o) {
    if (!d || !o || !o->permtab || !o->rangtab || !o->treetab) return PASTIX_ERR_BADPARAMETER;
    const char *dir = output_dir(d);
    if (!dir) return PASTIX_ERR_FILE;
    const char *env = getenv("PASTIX_FILE_ORDER");
    const std::string path = std::string(dir) + "/" + (env ? env : "ordergen");
    pxa::Order t;
    const int64_t b = o->baseval;
    t.n = o->vertnbr;
    t.cblknbr = o->cblknbr;
    for (pastix_int_t i = 0; i < o->vertnbr; i++) t.permtab.push_back(o->permtab[i] - b);
    for (pastix_int_t c = 0; c <= o->cblknbr; c++) t.rangtab.push_back(o->rangtab[c] - b);
    for (pastix_int_t c = 0; c < o->cblknbr; c++) t.treetab.push_back(o->treetab[c] < b ? -1 : o->treetab[c] - b);
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return PASTIX_ERR_FILE;
    const int rc = order_write(f, t, b);
    fclose(f);
    return rc;
}

int pxa_api::save_analysis_file(pastix_data_t *d, const char *name, bool order) {
    const char *dir = output_dir(d);
    if (!dir) return PASTIX_ERR_FILE;
    const std::string path = std::string(dir) + "/" + name;
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return PASTIX_ERR_FILE;
    const int rc = order ? order_write(f, d->order, 0) : symbol_write(f, d->symbol);
    fclose(f);
    return rc;
}
