"""Shared helpers for the tests: run the engine's analysis through the C-ABI
and build the oracle on the same structure."""
import numpy as np

from pastix_amd import pastix, spm
import oracle


class RawSpm:
    """An spmatrix_t over given index arrays: fmt SpmCSC (colptr = column
    pointers, rowptr = rows), SpmCSR (rowptr = row pointers, colptr =
    columns) or SpmIJV (colptr = columns, rowptr = rows; entries may repeat
    and add up)."""

    def __init__(self, fmt, colptr, rowptr, vals, n, symmetric=False):
        import ctypes as ct
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.values = np.ascontiguousarray(vals, dtype=np.float64)
        self.dtype = self.values.dtype
        s = spm.spmatrix_t()
        s.mtxtype = spm.SpmSymmetric if symmetric else spm.SpmGeneral
        s.flttype = spm.SpmDouble
        s.fmttype = fmt
        s.gN = s.n = s.gNexp = s.nexp = n
        s.gnnz = s.nnz = s.gnnzexp = s.nnzexp = len(self.values)
        s.dof = 1
        s.layout = spm.SpmColMajor
        s.colptr = self.colptr.ctypes.data_as(ct.POINTER(spm.spm_int))
        s.rowptr = self.rowptr.ctypes.data_as(ct.POINTER(spm.spm_int))
        s.values = self.values.ctypes.data_as(ct.c_void_p)
        self.s = s
        self.n = n

    @property
    def id_ptr(self):
        import ctypes as ct
        return ct.byref(self.s)


def analyse(A, symmetric, facto, grid=None, min_bs=160, max_bs=320, level_order=False):
    """Init + analysis; returns (handle, Spm, symbol dict, grid order or None)."""
    S = spm.Spm(A, symmetric=symmetric)
    ip, dp = pastix.initParam()
    ip[pastix.iparm.factorization] = facto
    ip[pastix.iparm.min_blocksize] = min_bs
    ip[pastix.iparm.max_blocksize] = max_bs
    h = pastix.init(ip, dp)
    if level_order:
        pastix.set_amd_iparm(h, pastix.amd_iparm.level_order, 1)
    o = None
    if grid is not None:
        o = pastix.orderGrid(*grid)
        h.order = o
        ip[pastix.iparm.ordering] = pastix.order.Personal
    pastix.subtask_order(h, S, o)
    pastix.subtask_symbfact(h)
    pastix.subtask_reordering(h)
    pastix.subtask_blend(h)
    return h, S, pastix.symbol(h)


def oracle_factor(A, sym, lu=False, crit=None):
    O = oracle.Oracle(sym)
    bc = oracle.bcsc(A, sym["permtab"])
    L, U, rc = O.fill(bc, lu=lu)
    assert rc == 0, "fill dropped entries outside the symbol"
    if crit is None:
        crit = np.sqrt(np.ldexp(1.0, -53)) * np.linalg.norm(bc["vals"])
    if lu:
        O.getrf(L, U, crit)
    else:
        O.potrf(L, crit)
    return O, bc, L, U


def read_rua(path):
    """Harwell-Boeing real assembled matrix (fixed-width Fortran fields): RUA
    as stored, RSA (lower triangle stored) expanded to both triangles."""
    import re
    import scipy.sparse as sps
    lines = open(path).read().split("\n")
    totcrd, ptrcrd, indcrd, valcrd = [int(v) for v in lines[1].split()[:4]]
    nrow, ncol, nnz = [int(v) for v in lines[2][14:].split()[:3]]
    fmts = re.findall(r"\(([^)]*)\)", lines[3])

    def width(f):
        m = re.match(r"(?:\d+P)?(\d+)[IEDF](\d+)", f.upper())
        return int(m.group(1)), int(m.group(2))

    def read(start, ncard, fmt, count, conv):
        per, w = width(fmt)
        out = []
        for line in lines[start:start + ncard]:
            for k in range(per):
                s = line[k * w:(k + 1) * w].strip()
                if s:
                    out.append(conv(s.replace("D", "E")))
        return out[:count]

    s = 4
    ptr = read(s, ptrcrd, fmts[0], ncol + 1, int)
    ind = read(s + ptrcrd, indcrd, fmts[1], nnz, int)
    val = read(s + ptrcrd + indcrd, valcrd, fmts[2], nnz, float)
    A = sps.csc_matrix((np.array(val), np.array(ind) - 1, np.array(ptr) - 1), shape=(nrow, ncol))
    if lines[2][:3].upper() == "RSA":
        # symmetric assembled: the lower triangle is stored
        A = (A + A.T - sps.diags(A.diagonal())).tocsc()
    return A


def rel_frob(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def lower_mask_panel(sym, L):
    """Zero the strictly-upper part of every diagonal block (never read for LL^t)."""
    L = L.copy()
    for c in range(sym["cblknbr"]):
        fc, lc, st = sym["cblk_fcol"][c], sym["cblk_fcol"][c + 1], sym["cblk_stride"][c]
        P = L[sym["cblk_coef"][c]:sym["cblk_coef"][c + 1]].reshape((lc - fc, st)).T
        K = lc - fc
        P[:K, :K] = np.tril(P[:K, :K])
        L[sym["cblk_coef"][c]:sym["cblk_coef"][c + 1]] = P.T.ravel()
    return L
