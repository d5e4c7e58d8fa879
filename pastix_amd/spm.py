"""Sparse matrix descriptor (spmatrix_t) and the synthetic generators of the
benchmark workloads.

`Spm` builds the C struct consumed by the PaStiX API (layout of the reference
spm/include/spm.h:50-78) from numpy CSC arrays and keeps them alive.  The
generators produce the matrices named by the workloads: constant-coefficient
7-point stencils, the anisotropic 7-point stencil, the 27-point
convection-diffusion operator and shifted 2D 5-point Laplacians.
"""
import ctypes as ct

import numpy as np
import scipy.sparse as sps

SpmPattern, SpmFloat, SpmDouble = 0, 2, 3
SpmCSC, SpmCSR, SpmIJV = 0, 1, 2
SpmGeneral, SpmSymmetric, SpmHermitian = 111, 112, 113
SpmColMajor = 102

spm_int = ct.c_int64


class spmatrix_t(ct.Structure):
    _fields_ = [
        ("mtxtype", ct.c_int), ("flttype", ct.c_int), ("fmttype", ct.c_int),
        ("gN", spm_int), ("n", spm_int), ("gnnz", spm_int), ("nnz", spm_int),
        ("gNexp", spm_int), ("nexp", spm_int), ("gnnzexp", spm_int), ("nnzexp", spm_int),
        ("dof", spm_int), ("dofs", ct.POINTER(spm_int)), ("layout", ct.c_int),
        ("colptr", ct.POINTER(spm_int)), ("rowptr", ct.POINTER(spm_int)),
        ("loc2glob", ct.POINTER(spm_int)), ("values", ct.c_void_p),
    ]


class Spm:
    """A CSC matrix wrapped as spmatrix_t.

    For symmetric matrices only the lower triangle is stored, as the SPM
    library does; `A` keeps the full scipy matrix for checks.
    """

    def __init__(self, A, symmetric=False, dtype=np.float64):
        A = sps.csc_matrix(A)
        A.sort_indices()
        self.A = A
        self.symmetric = symmetric
        S = sps.tril(A, format="csc") if symmetric else A
        S = sps.csc_matrix(S)
        S.sort_indices()
        self.colptr = np.ascontiguousarray(S.indptr, dtype=np.int64)
        self.rowptr = np.ascontiguousarray(S.indices, dtype=np.int64)
        self.values = np.ascontiguousarray(S.data, dtype=dtype)
        self.dtype = np.dtype(dtype)
        n = A.shape[0]
        s = spmatrix_t()
        s.mtxtype = SpmSymmetric if symmetric else SpmGeneral
        s.flttype = SpmDouble if self.dtype == np.float64 else SpmFloat
        s.fmttype = SpmCSC
        s.gN = s.n = s.gNexp = s.nexp = n
        s.gnnz = s.nnz = s.gnnzexp = s.nnzexp = len(self.values)
        s.dof = 1
        s.layout = SpmColMajor
        s.colptr = self.colptr.ctypes.data_as(ct.POINTER(spm_int))
        s.rowptr = self.rowptr.ctypes.data_as(ct.POINTER(spm_int))
        s.values = self.values.ctypes.data_as(ct.c_void_p)
        self.s = s
        self.n = n

    @property
    def id_ptr(self):
        return ct.byref(self.s)

    def set_values(self, A):
        """Refresh the values of an identical pattern (refactorisation)."""
        A = sps.csc_matrix(A)
        A.sort_indices()
        S = sps.tril(A, format="csc") if self.symmetric else A
        S = sps.csc_matrix(S)
        S.sort_indices()
        assert np.array_equal(S.indptr, self.colptr) and np.array_equal(S.indices, self.rowptr)
        self.values[:] = S.data
        self.A = A

    def norm_inf(self):
        return float(abs(self.A).sum(axis=1).max())


def _grid_index(nx, ny, nz):
    return np.arange(nx * ny * nz).reshape(nz, ny, nx)  # index = i + nx*j + nx*ny*k


def laplacian7(nx, ny, nz, coef=(1.0, 1.0, 1.0), diag=None):
    """7-point stencil on an nx x ny x nz grid, Dirichlet truncation.

    Off-diagonals -c_d along direction d; the diagonal is `diag` (constant) or
    2*(cx+cy+cz) when None.  Vertex index i + nx*j + nx*ny*k.
    """
    n = nx * ny * nz
    idx = _grid_index(nx, ny, nz)
    rows, cols, vals = [], [], []
    for d, (c, sl_a, sl_b) in enumerate([
        (coef[0], (slice(None), slice(None), slice(1, None)), (slice(None), slice(None), slice(None, -1))),
        (coef[1], (slice(None), slice(1, None), slice(None)), (slice(None), slice(None, -1), slice(None))),
        (coef[2], (slice(1, None), slice(None), slice(None)), (slice(None, -1), slice(None), slice(None))),
    ]):
        a = idx[sl_a].ravel()
        b = idx[sl_b].ravel()
        rows += [a, b]
        cols += [b, a]
        vals += [np.full(a.size, -c), np.full(a.size, -c)]
    dval = 2.0 * sum(coef) if diag is None else diag
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(np.full(n, dval))
    A = sps.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A.sum_duplicates()
    return A


def laplacian5_2d(k, shift=0.0):
    """2D 5-point Laplacian (diagonal 4, off-diagonal -1) plus shift*I on k x k."""
    return laplacian7(k, k, 1, coef=(1.0, 1.0, 0.0), diag=4.0 + shift)


def convection_diffusion27(nx, ny, nz, seed=0, peclet=10.0, boost=1.05):
    """Nonsymmetric 27-point operator with a symmetric pattern.

    Diffusion weights 1 (face), 1/sqrt(2) (edge), 1/sqrt(3) (corner); first
    order upwind convection with per-node velocity uniform(-1,1)^3 scaled to
    cell Peclet `peclet`; the diagonal is `boost` times the row's
    off-diagonal absolute sum.
    """
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    vel = rng.uniform(-1.0, 1.0, size=(n, 3))
    idx = _grid_index(nx, ny, nz)
    I, J, V = [], [], []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx == dy == dz == 0:
                    continue
                m = abs(dx) + abs(dy) + abs(dz)
                w = {1: 1.0, 2: 1.0 / np.sqrt(2.0), 3: 1.0 / np.sqrt(3.0)}[m]
                zs = slice(max(0, -dz), nz - max(0, dz))
                ys = slice(max(0, -dy), ny - max(0, dy))
                xs = slice(max(0, -dx), nx - max(0, dx))
                src = idx[zs, ys, xs].ravel()
                dst = src + dx + nx * dy + nx * ny * dz
                val = np.full(src.size, -w)
                if m == 1:
                    # upwind: inflow neighbour gets -Pe*|v_d| (cell Peclet scaling)
                    d = 0 if dx else (1 if dy else 2)
                    s = dx + dy + dz
                    vd = vel[src, d]
                    inflow = (vd * s) < 0
                    val = val - np.where(inflow, peclet * np.abs(vd) * 0.5, 0.0)
                I.append(src)
                J.append(dst)
                V.append(val)
    I = np.concatenate(I)
    J = np.concatenate(J)
    V = np.concatenate(V)
    off = sps.csr_matrix((V, (I, J)), shape=(n, n))
    dsum = np.asarray(abs(off).sum(axis=1)).ravel()
    A = off + sps.diags(boost * dsum)
    return sps.csc_matrix(A)


def rhs_from_x(A, nrhs, seed):
    """x_true uniform(-1,1) from `seed` (numpy PCG64), b = A x_true."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, size=(A.shape[0], nrhs))
    b = A @ x
    return np.asfortranarray(x), np.asfortranarray(b)
