"""CPU oracle of the MI355X engine -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, as the checker and as the timed CPU baseline.  The product
(pastix_amd, libpastix_amd.so) never imports it.

`Oracle` wraps liboracle.so (oracle.c: a plain-C restatement of the
reference sequential supernodal factorisation, solve, SpMV, flop count and
the CG, GMRES, simple and BiCGSTAB refinement drivers) over the block structure exported by the engine's analysis.
`bcsc()` restates the reference permuted block-CSC construction
(bcsc/bcsc_zinit.c:485-541) with scipy: B = P A P^T with its full pattern.
"""
import ctypes as ct
import os

import numpy as np
import scipy.sparse as sps

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


class ora_symbol_t(ct.Structure):
    _fields_ = [("n", ct.c_int64), ("cblknbr", ct.c_int64), ("bloknbr", ct.c_int64), ("coefnbr", ct.c_int64)] + \
               [(k, ct.POINTER(ct.c_int64)) for k in ("cblk_fcol", "cblk_bloknum", "cblk_stride", "cblk_coef",
                                                       "blok_frow", "blok_lrow", "blok_fcblk", "blok_roff")]


def _lib():
    if not os.path.exists(_LIB):
        raise ImportError(f"{_LIB} missing: run make -C oracle")
    lib = ct.CDLL(_LIB)
    P = ct.c_void_p
    lib.ora_fill.argtypes = [P] * 9
    lib.ora_fill.restype = ct.c_int
    lib.ora_potrf.argtypes = [P, P, ct.c_double, P]
    lib.ora_potrf.restype = ct.c_int
    lib.ora_getrf.argtypes = [P, P, P, ct.c_double, P]
    lib.ora_getrf.restype = ct.c_int
    lib.ora_solve.argtypes = [P, ct.c_int, P, P, P, ct.c_int64, ct.c_int]
    lib.ora_solve.restype = ct.c_int
    lib.ora_spmv.argtypes = [ct.c_int64, P, P, P, ct.c_double, P, ct.c_double, P]
    lib.ora_spmv.restype = None
    lib.ora_thflops.argtypes = [P, ct.c_int]
    lib.ora_thflops.restype = ct.c_double
    lib.ora_cg.argtypes = [P, P, ct.c_int64, P, P, P, P, P, ct.c_int, ct.c_double, P]
    lib.ora_cg.restype = ct.c_int
    for f in (lib.ora_sr, lib.ora_bicgstab):
        f.argtypes = [P, ct.c_int, P, P, ct.c_int64, P, P, P, P, P, ct.c_int, ct.c_double, P]
        f.restype = ct.c_int
    lib.ora_gmres.argtypes = [P, ct.c_int, P, P, ct.c_int64, P, P, P, P, P, ct.c_int, ct.c_int, ct.c_double, P]
    lib.ora_gmres.restype = ct.c_int
    lib.ora_factor_static.argtypes = [P, ct.c_int, P, P, ct.c_double, ct.c_int, P]
    lib.ora_factor_static.restype = ct.c_int
    lib.ora_use_blas.argtypes = [ct.c_char_p, ct.c_char_p]
    lib.ora_use_blas.restype = ct.c_int
    return lib


def use_blas():
    """Give the static scheduler a sequential CBLAS (scipy's OpenBLAS, one
    thread per call), as the reference's CPU path calls one; returns a
    description, or None when it is not found (plain loops then)."""
    import glob
    import scipy
    libdir = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
    for path in sorted(glob.glob(os.path.join(libdir, "libscipy_openblas*.so"))):
        if _lib().ora_use_blas(path.encode(), b"scipy_") == 0:
            return "OpenBLAS " + os.path.basename(path) + " (scipy wheel), 1 thread per worker"
    return None


def bcsc(A, permtab):
    """B = P A P^T (full pattern) as (csc of B, csc of B^T) index/value arrays."""
    n = A.shape[0]
    p = np.asarray(permtab, dtype=np.int64)
    Pm = sps.csr_matrix((np.ones(n), (p, np.arange(n))), shape=(n, n))
    B = sps.csc_matrix(Pm @ sps.csc_matrix(A) @ Pm.T)
    B.sum_duplicates()
    B.sort_indices()
    Bt = sps.csc_matrix(B.T)
    Bt.sort_indices()
    as64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)
    return {"B": B, "colptr": as64(B.indptr), "rows": as64(B.indices), "vals": np.ascontiguousarray(B.data),
            "tcolptr": as64(Bt.indptr), "trows": as64(Bt.indices), "tvals": np.ascontiguousarray(Bt.data)}


class Oracle:
    def __init__(self, sym):
        self.lib = _lib()
        self.sym = {k: np.ascontiguousarray(v, dtype=np.int64) for k, v in sym.items() if isinstance(v, np.ndarray)}
        s = ora_symbol_t()
        s.n, s.cblknbr, s.bloknbr, s.coefnbr = sym["n"], sym["cblknbr"], sym["bloknbr"], sym["coefnbr"]
        for k in ("cblk_fcol", "cblk_bloknum", "cblk_stride", "cblk_coef", "blok_frow", "blok_lrow", "blok_fcblk",
                  "blok_roff"):
            setattr(s, k, self.sym[k].ctypes.data_as(ct.POINTER(ct.c_int64)))
        self.s = s
        self.n = sym["n"]
        self.coefnbr = sym["coefnbr"]

    def fill(self, bc, lu=False):
        L = np.zeros(self.coefnbr)
        U = np.zeros(self.coefnbr) if lu else None
        rc = self.lib.ora_fill(ct.byref(self.s), bc["colptr"].ctypes.data, bc["rows"].ctypes.data,
                               bc["vals"].ctypes.data, bc["tcolptr"].ctypes.data, bc["trows"].ctypes.data,
                               bc["tvals"].ctypes.data, L.ctypes.data, U.ctypes.data if lu else None)
        return L, U, rc

    def potrf(self, L, crit):
        npiv = ct.c_int64()
        rc = self.lib.ora_potrf(ct.byref(self.s), L.ctypes.data, crit, ct.byref(npiv))
        return rc, npiv.value

    def getrf(self, L, U, crit):
        npiv = ct.c_int64()
        rc = self.lib.ora_getrf(ct.byref(self.s), L.ctypes.data, U.ctypes.data, crit, ct.byref(npiv))
        return rc, npiv.value

    def factor_static(self, L, U, crit, nthreads):
        """The reference static scheduler on nthreads threads (sopalin/sequential_zpotrf.c:71-120);
        LU when U is given.  Returns (rc, static pivots)."""
        npiv = ct.c_int64()
        rc = self.lib.ora_factor_static(ct.byref(self.s), 1 if U is not None else 0, L.ctypes.data,
                                        U.ctypes.data if U is not None else None, crit, nthreads, ct.byref(npiv))
        return rc, npiv.value

    def solve(self, L, U, x):
        """In-place solve of a Fortran-ordered (n, nrhs) block in the permuted ordering."""
        x = np.asfortranarray(x, dtype=np.float64)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        rc = self.lib.ora_solve(ct.byref(self.s), 1 if U is not None else 0, L.ctypes.data,
                                U.ctypes.data if U is not None else None, x.ctypes.data, self.n, nrhs)
        assert rc == 0
        return x

    def spmv(self, bc, x, alpha=1.0, beta=0.0, y=None):
        """y = alpha * B x + beta * y with B from bcsc(): rows of B are the columns of B^T."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n) if y is None else np.ascontiguousarray(y, dtype=np.float64).copy()
        self.lib.ora_spmv(self.n, bc["tcolptr"].ctypes.data, bc["trows"].ctypes.data, bc["tvals"].ctypes.data,
                          alpha, x.ctypes.data, beta, y.ctypes.data)
        return y

    def thflops(self, lu=False):
        return self.lib.ora_thflops(ct.byref(self.s), 1 if lu else 0)

    def cg(self, L, bc, b, x, itermax=250, eps=1e-12):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64)
        res = ct.c_double()
        it = self.lib.ora_cg(ct.byref(self.s), L.ctypes.data, self.n, bc["tcolptr"].ctypes.data,
                             bc["trows"].ctypes.data, bc["tvals"].ctypes.data, b.ctypes.data, x.ctypes.data,
                             itermax, eps, ct.byref(res))
        return x, it, res.value

    def _refine(self, fn, L, U, bc, b, x, itermax, eps):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64)
        res = ct.c_double()
        it = fn(ct.byref(self.s), 1 if U is not None else 0, L.ctypes.data, U.ctypes.data if U is not None else None,
                self.n, bc["tcolptr"].ctypes.data, bc["trows"].ctypes.data, bc["tvals"].ctypes.data, b.ctypes.data,
                x.ctypes.data, itermax, eps, ct.byref(res))
        return x, it, res.value

    def sr(self, L, U, bc, b, x, itermax=250, eps=1e-12):
        """Simple refinement (reference refinement/z_refine_pivot.c); returns (x, passes, ||r||/||b||)."""
        return self._refine(self.lib.ora_sr, L, U, bc, b, x, itermax, eps)

    def bicgstab(self, L, U, bc, b, x, itermax=250, eps=1e-12):
        """Preconditioned BiCGSTAB (reference refinement/z_refine_bicgstab.c); returns (x, iterations, ||r||/||b||)."""
        return self._refine(self.lib.ora_bicgstab, L, U, bc, b, x, itermax, eps)

    def gmres(self, L, U, bc, b, x, itermax=250, im=25, eps=1e-12):
        """Right-preconditioned restarted GMRES (reference refinement/z_refine_gmres.c);
        returns (x, steps, last residual estimate / ||b||)."""
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        b = np.ascontiguousarray(b, dtype=np.float64)
        res = ct.c_double()
        it = self.lib.ora_gmres(ct.byref(self.s), 1 if U is not None else 0, L.ctypes.data,
                                U.ctypes.data if U is not None else None, self.n, bc["tcolptr"].ctypes.data,
                                bc["trows"].ctypes.data, bc["tvals"].ctypes.data, b.ctypes.data, x.ctypes.data,
                                itermax, im, eps, ct.byref(res))
        return x, it, res.value

    # ---- helpers on the 1D panel layout ----
    def to_dense(self, L, lower_only=True):
        """Scatter a panel table into a dense n x n matrix (permuted numbering)."""
        s = self.sym
        D = np.zeros((self.n, self.n))
        for c in range(len(s["cblk_stride"])):
            fc, lc, st = s["cblk_fcol"][c], s["cblk_fcol"][c + 1], s["cblk_stride"][c]
            P = L[s["cblk_coef"][c]:s["cblk_coef"][c + 1]].reshape((lc - fc, st)).T
            for b in range(s["cblk_bloknum"][c], s["cblk_bloknum"][c + 1]):
                r0, nr = s["blok_roff"][b], s["blok_lrow"][b] - s["blok_frow"][b] + 1
                D[s["blok_frow"][b]:s["blok_frow"][b] + nr, fc:lc] = P[r0:r0 + nr, :]
        return np.tril(D) if lower_only else D
