"""Python binding of the PaStiX C API served by libpastix_amd.so.

Function names and argument meaning follow the reference pypastix module
(wrappers/python/examples/pypastix/pastix.py): initParam, init, finalize,
task_analyze, task_numfact, task_solve, task_refine and the subtasks.  Calls
go through the C-ABI with ctypes; errors from the library raise
PastixError with the PASTIX_ERR_* code.

The library must have been built (``python -c "import __graft_entry__ as g;
g.build()"``); importing this module without it raises ImportError, there is
no fallback implementation.
"""
import ctypes as ct
import os

import numpy as np

from . import spm as _spm

pastix_int = ct.c_int64
_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libpastix_amd.so")

if not os.path.exists(LIBPATH):
    raise ImportError(f"{LIBPATH} is missing: build it with __graft_entry__.build()")
lib = ct.CDLL(LIBPATH)

# ---------------------------------------------------------------- enums ----
class iparm:
    verbose, io_strategy, nnzeros, nnzeros_block_local, allocated_terms, produce_stats = range(6)
    mc64, ordering, ordering_default = 6, 7, 8
    amalgamation_lvlblas, amalgamation_lvlcblk = 24, 25
    reordering_split, reordering_stop = 26, 27
    min_blocksize, max_blocksize, tasks2d_level, tasks2d_width, allcand = 34, 35, 36, 37, 38
    incomplete, level_of_fill = 39, 40
    factorization, static_pivoting, free_cscuser = 41, 42, 43
    schur_fact_mode, schur_solv_mode, applyperm_ws = 44, 45, 46
    refinement, nbiter, itermax, gmres_im = 47, 48, 49, 50
    scheduler, thread_nbr, autosplit_comm = 51, 52, 53
    gpu_nbr, gpu_memory_percentage, gpu_memory_block_size = 54, 55, 56
    modify_parameter = 65
    dof_nbr = 70
    size = 71


class dparm:
    fill_in, epsilon_refinement, relative_error, epsilon_magn_ctrl = 0, 1, 2, 3
    analyze_time, pred_fact_time, fact_time, solv_time = 4, 5, 6, 7
    fact_flops, fact_thflops, fact_rlflops = 8, 9, 10
    solv_flops, solv_thflops, solv_rlflops, refine_time, a_norm = 11, 12, 13, 14, 15
    compress_tolerance, compress_min_ratio = 16, 17
    size = 18


class factotype:
    POTRF, SYTRF, GETRF, PXTRF, HETRF = 0, 1, 2, 3, 4
    LLH, LDLT, LU, LLT, LDLH = 0, 1, 2, 3, 4


class scheduler:
    Sequential, Static, Parsec, StarPU, Dynamic = 0, 1, 2, 3, 4


class refine:
    GMRES, CG, SR, BiCGSTAB = 0, 1, 2, 3


class order:
    Scotch, Metis, Personal, PtScotch, ParMetis = 0, 1, 2, 3, 4


class dir:
    Forward, Backward = 391, 392


class trans:
    NoTrans, Trans, ConjTrans = 111, 112, 113


class uplo:
    Upper, Lower = 121, 122


class diag:
    NonUnit, Unit = 131, 132


class side:
    Left, Right = 141, 142


class error:
    SUCCESS, UNKNOWN, ALLOC, NOTIMPLEMENTED, OUTOFMEMORY, THREAD, INTERNAL, BADPARAMETER = range(8)


class amd_iparm:
    factor_precision, device, level_order, resident_input, profile, graph, overlap = 0, 1, 2, 3, 4, 5, 6
    leaf_tasks, leaf_grid, solve_node_max = 7, 8, 9


class PastixError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} returned {code}")
        self.code = code


class pastix_order_t(ct.Structure):
    _fields_ = [
        ("baseval", pastix_int), ("vertnbr", pastix_int), ("cblknbr", pastix_int),
        ("permtab", ct.POINTER(pastix_int)), ("peritab", ct.POINTER(pastix_int)),
        ("rangtab", ct.POINTER(pastix_int)), ("treetab", ct.POINTER(pastix_int)),
        ("selevtx", ct.c_void_p), ("sndenbr", pastix_int), ("sndetab", ct.POINTER(pastix_int)),
    ]


_P = ct.c_void_p
_I = pastix_int
_A = np.ctypeslib.ndpointer
_ip = _A(dtype=np.int64, flags="C_CONTIGUOUS")
_dp = _A(dtype=np.float64, flags="C_CONTIGUOUS")

lib.pastixInitParam.argtypes = [_ip, _dp]
lib.pastixInitParam.restype = None
lib.pastixInit.argtypes = [ct.POINTER(_P), ct.c_uint64, _ip, _dp]
lib.pastixInit.restype = None
lib.pastixFinalize.argtypes = [ct.POINTER(_P)]
lib.pastixFinalize.restype = None
for _f in ("pastix_task_analyze", "pastix_task_numfact", "pastix_subtask_spm2bcsc"):
    getattr(lib, _f).argtypes = [_P, ct.c_void_p]
    getattr(lib, _f).restype = ct.c_int
for _f in ("pastix_subtask_symbfact", "pastix_subtask_reordering", "pastix_subtask_blend",
           "pastix_subtask_bcsc2ctab", "pastix_subtask_sopalin"):
    getattr(lib, _f).argtypes = [_P]
    getattr(lib, _f).restype = ct.c_int
lib.pastix_subtask_order.argtypes = [_P, ct.c_void_p, ct.c_void_p]
lib.pastix_subtask_order.restype = ct.c_int
lib.pastix_task_solve.argtypes = [_P, _I, ct.c_void_p, _I]
lib.pastix_task_solve.restype = ct.c_int
lib.pastix_subtask_solve.argtypes = [_P, _I, ct.c_void_p, _I]
lib.pastix_subtask_solve.restype = ct.c_int
lib.pastix_task_refine.argtypes = [_P, _I, _I, ct.c_void_p, _I, ct.c_void_p, _I]
lib.pastix_task_refine.restype = ct.c_int
lib.pastix_subtask_refine.argtypes = [_P, _I, _I, ct.c_void_p, _I, ct.c_void_p, _I]
lib.pastix_subtask_refine.restype = ct.c_int
lib.pastix_subtask_applyorder.argtypes = [_P, ct.c_int, ct.c_int, _I, _I, ct.c_void_p, _I]
lib.pastix_subtask_applyorder.restype = ct.c_int
lib.pastix_subtask_trsm.argtypes = [_P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _I, ct.c_void_p, _I]
lib.pastix_subtask_trsm.restype = ct.c_int
lib.pastix_subtask_diag.argtypes = [_P, ct.c_int, _I, ct.c_void_p, _I]
lib.pastixGetDiag.argtypes = [_P, ct.c_void_p, _I]
lib.pastixGetDiag.restype = ct.c_int
lib.pastix_subtask_diag.restype = ct.c_int
lib.pastixOrderGrid.argtypes = [ct.POINTER(ct.POINTER(pastix_order_t)), _I, _I, _I]
lib.pastixOrderGrid.restype = ct.c_int
lib.pastixOrderExit.argtypes = [ct.POINTER(pastix_order_t)]
lib.pastixOrderExit.restype = None
lib.pastix_amd_set_iparm.argtypes = [_P, ct.c_int, _I]
lib.pastix_amd_set_iparm.restype = ct.c_int
lib.pastix_amd_symbol_sizes.argtypes = [_P, _ip]
lib.pastix_amd_symbol_sizes.restype = ct.c_int
lib.pastix_amd_symbol_export.argtypes = [_P] + [ct.c_void_p] * 10
lib.pastix_amd_symbol_export.restype = ct.c_int
lib.pastix_amd_coeftab_export.argtypes = [_P, ct.c_void_p, ct.c_void_p]
lib.pastix_amd_coeftab_export.restype = ct.c_int
lib.pastix_amd_backward_error.argtypes = [_P, _dp, _dp, ct.POINTER(ct.c_double)]
lib.pastix_amd_backward_error.restype = ct.c_int


def _check(fn, rc):
    if rc != 0:
        raise PastixError(fn, rc)


def _columns(x, nrhs):
    """Right-hand sides to process: all columns of x unless nrhs >= 0 says otherwise."""
    return nrhs if nrhs >= 0 else (x.shape[1] if x.ndim > 1 else 1)


# ---------------------------------------------------------- main API ----
def initParam():
    ip = np.zeros(iparm.size, dtype=np.int64)
    dp = np.zeros(dparm.size, dtype=np.float64)
    lib.pastixInitParam(ip, dp)
    return ip, dp


class _Handle:
    """Keeps the iparm/dparm arrays alive as long as the instance (the library
    stores the caller's pointers, reference common/api.c:733-734)."""

    def __init__(self, ptr, ip, dp):
        self.ptr = ptr
        self.iparm = ip
        self.dparm = dp
        self.order = None


def init(iparm_, dparm_):
    p = _P()
    lib.pastixInit(ct.byref(p), 0, iparm_, dparm_)
    return _Handle(p, iparm_, dparm_)


def finalize(pastix_data, iparm_=None, dparm_=None):
    if pastix_data.ptr:
        lib.pastixFinalize(ct.byref(pastix_data.ptr))
    if pastix_data.order is not None:
        lib.pastixOrderExit(pastix_data.order)
        pastix_data.order = None


def task_analyze(pastix_data, spmA):
    _check("pastix_task_analyze", lib.pastix_task_analyze(pastix_data.ptr, spmA.id_ptr))


def task_numfact(pastix_data, spmA):
    _check("pastix_task_numfact", lib.pastix_task_numfact(pastix_data.ptr, spmA.id_ptr))


def task_solve(pastix_data, spmA, x, nrhs=-1):
    n = x.shape[0]
    nrhs = _columns(x, nrhs)
    assert x.flags.f_contiguous and x.dtype == spmA.dtype
    _check("pastix_task_solve", lib.pastix_task_solve(pastix_data.ptr, nrhs, x.ctypes.data, n))


def task_refine(pastix_data, spmA, b, x, nrhs=-1):
    n = x.shape[0]
    nrhs = _columns(x, nrhs)
    if b.dtype != x.dtype:
        raise TypeError(f"task_refine: right-hand side ({b.dtype}) and solution ({x.dtype}) dtypes differ")
    assert x.flags.f_contiguous and b.flags.f_contiguous
    _check("pastix_task_refine", lib.pastix_task_refine(pastix_data.ptr, n, nrhs, b.ctypes.data, b.shape[0],
                                                        x.ctypes.data, x.shape[0]))


def subtask_order(pastix_data, spmA, myorder=None):
    ptr = ct.byref(myorder) if myorder is not None else None
    _check("pastix_subtask_order", lib.pastix_subtask_order(pastix_data.ptr, spmA.id_ptr, ptr))


def subtask_symbfact(pastix_data):
    _check("pastix_subtask_symbfact", lib.pastix_subtask_symbfact(pastix_data.ptr))


def subtask_reordering(pastix_data):
    _check("pastix_subtask_reordering", lib.pastix_subtask_reordering(pastix_data.ptr))


def subtask_blend(pastix_data):
    _check("pastix_subtask_blend", lib.pastix_subtask_blend(pastix_data.ptr))


def subtask_spm2bcsc(pastix_data, spmA):
    _check("pastix_subtask_spm2bcsc", lib.pastix_subtask_spm2bcsc(pastix_data.ptr, spmA.id_ptr))


def subtask_bcsc2ctab(pastix_data):
    _check("pastix_subtask_bcsc2ctab", lib.pastix_subtask_bcsc2ctab(pastix_data.ptr))


def subtask_sopalin(pastix_data):
    _check("pastix_subtask_sopalin", lib.pastix_subtask_sopalin(pastix_data.ptr))


def subtask_applyorder(pastix_data, direction, b):
    flt = _spm.SpmDouble if b.dtype == np.float64 else _spm.SpmFloat
    m = b.shape[0]
    n = 1 if b.ndim == 1 else b.shape[1]
    _check("pastix_subtask_applyorder",
           lib.pastix_subtask_applyorder(pastix_data.ptr, flt, direction, m, n, b.ctypes.data, m))


def subtask_trsm(pastix_data, side_, uplo_, trans_, diag_, b, nrhs=-1):
    """One sweep of the solve on b in the permuted ordering, in place."""
    nrhs = _columns(b, nrhs)
    flt = _spm.SpmDouble if b.dtype == np.float64 else _spm.SpmFloat
    _check("pastix_subtask_trsm", lib.pastix_subtask_trsm(pastix_data.ptr, flt, side_, uplo_, trans_, diag_, nrhs,
                                                          b.ctypes.data, b.shape[0]))


def subtask_diag(pastix_data, b, nrhs=-1):
    """b /= the factor's diagonal (D of LDL^t), b in the permuted ordering."""
    nrhs = _columns(b, nrhs)
    flt = _spm.SpmDouble if b.dtype == np.float64 else _spm.SpmFloat
    _check("pastix_subtask_diag", lib.pastix_subtask_diag(pastix_data.ptr, flt, nrhs, b.ctypes.data, b.shape[0]))


def getDiag(pastix_data, n, incD=1, dtype=np.float64):
    """Diagonal of the factor in the permuted ordering (pastixGetDiag):
    entries D[0::incD] of an n*incD array of the instance's arithmetic."""
    D = np.zeros(n * incD, dtype=dtype)
    _check("pastixGetDiag", lib.pastixGetDiag(pastix_data.ptr, D.ctypes.data, incD))
    return D


lib.pastixSetSchurUnknownList.argtypes = [_P, _I, ct.c_void_p]
lib.pastixSetSchurUnknownList.restype = None
lib.pastixGetSchur.argtypes = [_P, ct.c_void_p, _I]
lib.pastixGetSchur.restype = ct.c_int


class solv_mode:
    Local, Interface, Schur = 0, 1, 2


def setSchurUnknownList(pastix_data, schur_list):
    """Unknowns (in the matrix's numbering base) to keep out of the
    factorisation; the next analysis orders them last."""
    lst = np.ascontiguousarray(schur_list, dtype=np.int64)
    lib.pastixSetSchurUnknownList(pastix_data.ptr, len(lst), lst.ctypes.data)


def getSchur(pastix_data, S):
    """Fill the Fortran-ordered square array S (leading dimension S.shape[0])
    with the Schur complement after pastix_task_numfact."""
    assert S.flags.f_contiguous
    _check("pastixGetSchur", lib.pastixGetSchur(pastix_data.ptr, S.ctypes.data, S.shape[0]))


def subtask_solve(pastix_data, b, nrhs=-1):
    nrhs = _columns(b, nrhs)
    _check("pastix_subtask_solve", lib.pastix_subtask_solve(pastix_data.ptr, nrhs, b.ctypes.data, b.shape[0]))


# -------------------------------------------------------- orderings ----
def orderGrid(nx, ny, nz):
    """pastixOrderGrid (reference order/order_grids.c:275-383): returns a
    pastix_order_t owned by the caller (free with orderExit)."""
    o = pastix_order_t()
    po = ct.pointer(o)
    _check("pastixOrderGrid", lib.pastixOrderGrid(ct.byref(po), nx, ny, nz))
    return o


def orderExit(o):
    lib.pastixOrderExit(ct.byref(o))


def order_arrays(o):
    n, c = o.vertnbr, o.cblknbr
    return (np.ctypeslib.as_array(o.permtab, (n,)).copy(), np.ctypeslib.as_array(o.peritab, (n,)).copy(),
            np.ctypeslib.as_array(o.rangtab, (c + 1,)).copy(), np.ctypeslib.as_array(o.treetab, (c,)).copy())


# -------------------------------------------------------- extensions ----
def set_amd_iparm(pastix_data, key, value):
    _check("pastix_amd_set_iparm", lib.pastix_amd_set_iparm(pastix_data.ptr, key, value))


def symbol(pastix_data):
    """Analysed structure as a dict of numpy arrays (permuted numbering)."""
    sz = np.zeros(8, dtype=np.int64)
    _check("pastix_amd_symbol_sizes", lib.pastix_amd_symbol_sizes(pastix_data.ptr, sz))
    n, cb, nb = int(sz[0]), int(sz[1]), int(sz[2])
    out = {
        "cblk_fcol": np.zeros(cb + 1, np.int64), "cblk_bloknum": np.zeros(cb + 1, np.int64),
        "cblk_stride": np.zeros(cb, np.int64), "cblk_coef": np.zeros(cb + 1, np.int64),
        "blok_frow": np.zeros(nb, np.int64), "blok_lrow": np.zeros(nb, np.int64),
        "blok_fcblk": np.zeros(nb, np.int64), "blok_roff": np.zeros(nb, np.int64),
        "permtab": np.zeros(n, np.int64), "peritab": np.zeros(n, np.int64),
    }
    keys = ["cblk_fcol", "cblk_bloknum", "cblk_stride", "cblk_coef", "blok_frow", "blok_lrow", "blok_fcblk",
            "blok_roff", "permtab", "peritab"]
    _check("pastix_amd_symbol_export",
           lib.pastix_amd_symbol_export(pastix_data.ptr, *[out[k].ctypes.data for k in keys]))
    out["n"], out["cblknbr"], out["bloknbr"], out["coefnbr"] = n, cb, nb, int(sz[3])
    out["nlevels"], out["upd_items"], out["upd_entries"] = int(sz[5]), int(sz[6]), int(sz[7])
    return out


def coeftab(pastix_data, dtype=np.float64, lu=False):
    """Device factor copied to host: (L, U or None) in the 1D panel layout."""
    sz = np.zeros(8, dtype=np.int64)
    _check("pastix_amd_symbol_sizes", lib.pastix_amd_symbol_sizes(pastix_data.ptr, sz))
    L = np.zeros(int(sz[3]), dtype=dtype)
    U = np.zeros(int(sz[3]), dtype=dtype) if lu else None
    _check("pastix_amd_coeftab_export",
           lib.pastix_amd_coeftab_export(pastix_data.ptr, L.ctypes.data, U.ctypes.data if lu else None))
    return L, U


def backward_error(pastix_data, b, x):
    out = ct.c_double()
    _check("pastix_amd_backward_error",
           lib.pastix_amd_backward_error(pastix_data.ptr, np.ascontiguousarray(b, dtype=np.float64),
                                         np.ascontiguousarray(x, dtype=np.float64), ct.byref(out)))
    return out.value


lib.pastix_amd_backward_error_device_noalloc.argtypes = [_P, _I, ct.c_void_p, _I, ct.c_void_p, _I, ct.c_void_p]
lib.pastix_amd_backward_error_device_noalloc.restype = ct.c_int


def backward_error_device_noalloc(pastix_data, db, dx, nrhs, dberr):
    """Backward errors of the nrhs columns of the DeviceArrays dx against db
    (n x nrhs, user ordering) into the DeviceArray dberr, without any device
    allocation (library-static scratch, nrhs <= 64)."""
    n = db.shape[0]
    _check("pastix_amd_backward_error_device_noalloc",
           lib.pastix_amd_backward_error_device_noalloc(pastix_data.ptr, nrhs, db.ptr, n, dx.ptr, n, dberr.ptr))


lib.pastix_amd_set_level_knobs.argtypes = [_P, ct.c_int, ct.c_void_p, ct.c_void_p]
lib.pastix_amd_set_level_knobs.restype = ct.c_int
lib.pastix_amd_load_tuning.argtypes = [_P, ct.c_char_p, ct.POINTER(ct.c_int)]
lib.pastix_amd_load_tuning.restype = ct.c_int
lib.pastix_amd_save_tuning.argtypes = [_P, ct.c_char_p]
lib.pastix_amd_save_tuning.restype = ct.c_int
lib.pastix_amd_level_info.argtypes = [_P, ct.c_void_p, ct.c_int]
lib.pastix_amd_level_info.restype = ct.c_int
lib.pastix_amd_solve_info.argtypes = [_P, ct.c_void_p]
lib.pastix_amd_solve_info.restype = ct.c_int
lib.pastix_amd_level_times.argtypes = [_P, ct.c_void_p, ct.c_int]
lib.pastix_amd_level_times.restype = ct.c_int


def set_level_knobs(pastix_data, single_pass, panel_right):
    """Per-level schedule choices (-1 keeps the plan's rule); rebuilds the
    level schedule of an analysed instance."""
    sp = np.ascontiguousarray(single_pass, dtype=np.int8)
    pr = np.ascontiguousarray(panel_right, dtype=np.int8)
    if sp.shape != pr.shape:
        raise ValueError("set_level_knobs: one value per level in both arrays")
    _check("pastix_amd_set_level_knobs",
           lib.pastix_amd_set_level_knobs(pastix_data.ptr, len(sp), sp.ctypes.data, pr.ctypes.data))


def load_tuning(pastix_data, path):
    """Apply a tuning table's entry for this analysis; True when one matched."""
    applied = ct.c_int(0)
    _check("pastix_amd_load_tuning", lib.pastix_amd_load_tuning(pastix_data.ptr, str(path).encode(), ct.byref(applied)))
    return bool(applied.value)


def save_tuning(pastix_data, path):
    """Write (or replace) this analysis's entry with its current choices."""
    _check("pastix_amd_save_tuning", lib.pastix_amd_save_tuning(pastix_data.ptr, str(path).encode()))


def level_info(pastix_data):
    """Per level: (cblks, multi-source tiles, single-pass, right-looking)."""
    out = np.zeros((4096, 4), dtype=np.int32)
    nl = lib.pastix_amd_level_info(pastix_data.ptr, out.ctypes.data, 4096)
    if nl < 0:
        raise PastixError("pastix_amd_level_info", nl)
    return out[:nl].copy()


def solve_info(pastix_data):
    """Solve structure: nodes, nodes of several cblks, solve levels, merge
    rounds of the nodes' inverses, their tiles, entries of the merged nodes'
    inverses, merge flops (one side), factorisation levels."""
    out = np.zeros(8)
    _check("pastix_amd_solve_info", lib.pastix_amd_solve_info(pastix_data.ptr, out.ctypes.data))
    keys = ["nodes", "merged_nodes", "solve_levels", "merge_rounds", "merge_tiles", "merged_entries",
            "merge_flops", "levels"]
    return {k: (float(v) if k == "merge_flops" else int(v)) for k, v in zip(keys, out)}


def level_times(pastix_data):
    """Per level of the last profiled numfact: (panel s, update s)."""
    out = np.zeros((4096, 2))
    nl = lib.pastix_amd_level_times(pastix_data.ptr, out.ctypes.data, 4096)
    if nl < 0:
        raise PastixError("pastix_amd_level_times", nl)
    return out[:nl].copy()


lib.pastix_amd_fact_times.argtypes = [_P, _dp]
lib.pastix_amd_fact_times.restype = ct.c_int
lib.pastix_amd_flops.argtypes = [_P, _dp]
lib.pastix_amd_flops.restype = ct.c_int
_vp = ct.POINTER(ct.c_void_p)
_i64p = ct.POINTER(ct.c_int64)
lib.pastix_amd_batch_analyze.argtypes = [_P, ct.c_int64, _vp, _vp]
lib.pastix_amd_batch_analyze.restype = ct.c_int
lib.pastix_amd_batch_numfact.argtypes = [_P, _vp]
lib.pastix_amd_batch_numfact.restype = ct.c_int
lib.pastix_amd_batch_solve.argtypes = [_P, _i64p, _vp, _i64p]
lib.pastix_amd_batch_solve.restype = ct.c_int
lib.pastix_amd_batch_refine.argtypes = [_P, _i64p, _vp, _i64p, _vp, _i64p]
lib.pastix_amd_batch_refine.restype = ct.c_int
lib.pastix_amd_batch_solve_refine.argtypes = [_P, _i64p, _vp, _i64p, _vp, _i64p]
lib.pastix_amd_batch_solve_refine.restype = ct.c_int
lib.pastix_amd_batch_offsets.argtypes = [_P, _i64p]
lib.pastix_amd_batch_offsets.restype = ct.c_int
lib.pastix_amd_kernel_times.argtypes = [_P, _dp]
lib.pastix_amd_kernel_times.restype = ct.c_int
lib.pastix_amd_update_flops.argtypes = [_P, _dp]
lib.pastix_amd_update_flops.restype = ct.c_int
lib.pastix_amd_refactorize.argtypes = [_P]
lib.pastix_amd_refactorize.restype = ct.c_int
lib.pastix_amd_solve_device.argtypes = [_P, _I, ct.c_void_p, _I]
lib.pastix_amd_solve_device.restype = ct.c_int


def fact_times(pastix_data):
    """[total, diag, trsm, update] seconds and launch counts of the last numfact."""
    out = np.zeros(8)
    _check("pastix_amd_fact_times", lib.pastix_amd_fact_times(pastix_data.ptr, out))
    return out


KERNEL_CLASSES = ("panel", "update_direct", "update_wgemm", "update_gather", "update_acc", "leaf_tasks")


def kernel_times(pastix_data):
    """{class: (seconds, timed spans)} of the last numfact (profiling on)."""
    out = np.zeros(12)
    _check("pastix_amd_kernel_times", lib.pastix_amd_kernel_times(pastix_data.ptr, out))
    return {k: (float(out[2 * i]), int(out[2 * i + 1])) for i, k in enumerate(KERNEL_CLASSES)}


def update_flops(pastix_data):
    """Schur-update flops per kernel: L-side single/multi-source tiles, U-side
    single/multi-source tiles, multi-source product tiles as computed, and
    the multi-source tiles of single-pass levels (L, U side)."""
    out = np.zeros(8)
    _check("pastix_amd_update_flops", lib.pastix_amd_update_flops(pastix_data.ptr, out))
    return {"direct": float(out[0]), "multi": float(out[1]), "direct_u": float(out[2]),
            "multi_u": float(out[3]), "wtile": float(out[4]), "single": float(out[5]),
            "single_u": float(out[6]), "leaf": float(out[7])}


def _ptr_array(objs):
    arr = (ct.c_void_p * len(objs))()
    for k, o in enumerate(objs):
        arr[k] = ct.cast(o, ct.c_void_p) if o is not None else None
    return arr


def batch_analyze(pastix_data, spms, orders=None):
    """Analyse independent systems as one block-diagonal batch; orders is a
    list of pastix_order_t (or None entries / None for nested dissection)."""
    sp = _ptr_array([ct.addressof(S.s) for S in spms])
    od = _ptr_array([ct.addressof(o) if o is not None else None for o in orders]) if orders else None
    _check("pastix_amd_batch_analyze", lib.pastix_amd_batch_analyze(pastix_data.ptr, len(spms), sp, od))
    pastix_data.batch = (spms, orders)


def batch_numfact(pastix_data, spms):
    sp = _ptr_array([ct.addressof(S.s) for S in spms])
    _check("pastix_amd_batch_numfact", lib.pastix_amd_batch_numfact(pastix_data.ptr, sp))


def _rhs_lists(xs):
    xs = [np.asfortranarray(np.asarray(x, dtype=np.float64)) for x in xs]
    xs = [x.reshape(x.shape[0], 1, order="F") if x.ndim == 1 else x for x in xs]
    nr = np.array([x.shape[1] for x in xs], dtype=np.int64)
    ld = np.array([x.shape[0] for x in xs], dtype=np.int64)
    return xs, nr, ld


def batch_solve(pastix_data, bs):
    """Solves every system for its right-hand sides (list of n_s x nrhs_s
    arrays); returns the solutions."""
    xs, nr, ld = _rhs_lists([b.copy() for b in bs])
    ptrs = _ptr_array([x.ctypes.data for x in xs])
    _check("pastix_amd_batch_solve", lib.pastix_amd_batch_solve(
        pastix_data.ptr, nr.ctypes.data_as(_i64p), ptrs, ld.ctypes.data_as(_i64p)))
    return xs


def batch_refine(pastix_data, bs, xs):
    """Batched CG refinement of the solutions xs (updated copies returned)."""
    bl, nr, ldb = _rhs_lists(bs)
    xl, _, ldx = _rhs_lists([x.copy() for x in xs])
    pb = _ptr_array([b.ctypes.data for b in bl])
    px = _ptr_array([x.ctypes.data for x in xl])
    _check("pastix_amd_batch_refine", lib.pastix_amd_batch_refine(
        pastix_data.ptr, nr.ctypes.data_as(_i64p), pb, ldb.ctypes.data_as(_i64p), px, ldx.ctypes.data_as(_i64p)))
    return xl


def batch_solve_refine(pastix_data, bs):
    """Solve + batched CG for every system in one call (right-hand sides cross
    to the device once); returns the solutions."""
    bl, nr, ldb = _rhs_lists(bs)
    xl, _, ldx = _rhs_lists([np.zeros_like(b) for b in bs])
    pb = _ptr_array([b.ctypes.data for b in bl])
    px = _ptr_array([x.ctypes.data for x in xl])
    _check("pastix_amd_batch_solve_refine", lib.pastix_amd_batch_solve_refine(
        pastix_data.ptr, nr.ctypes.data_as(_i64p), pb, ldb.ctypes.data_as(_i64p), px, ldx.ctypes.data_as(_i64p)))
    return xl


def batch_offsets(pastix_data):
    n = lib.pastix_amd_batch_offsets(pastix_data.ptr, None)
    if n < 0:
        raise PastixError("pastix_amd_batch_offsets", n)
    off = np.zeros(n + 1, dtype=np.int64)
    lib.pastix_amd_batch_offsets(pastix_data.ptr, off.ctypes.data_as(_i64p))
    return off


def flops(pastix_data):
    """[THFLOPS, diag, trsm, update] algorithmic flops of the analysed problem."""
    out = np.zeros(4)
    _check("pastix_amd_flops", lib.pastix_amd_flops(pastix_data.ptr, out))
    return out


lib.pastix_amd_refactorize_shift.argtypes = [_P, ct.c_double]
lib.pastix_amd_refactorize_shift.restype = ct.c_int


def refactorize_shift(pastix_data, delta):
    """Refactorise A + delta*I from the HBM-resident filled coefficients."""
    _check("pastix_amd_refactorize_shift", lib.pastix_amd_refactorize_shift(pastix_data.ptr, float(delta)))


lib.pastix_amd_device_bytes.argtypes = [_P, _i64p]
lib.pastix_amd_device_bytes.restype = ct.c_int
lib.pastix_amd_server_create.argtypes = [ct.POINTER(ct.c_void_p), ct.POINTER(_P), ct.c_int, ct.c_int64]
lib.pastix_amd_server_create.restype = ct.c_int
lib.pastix_amd_server_submit_solve.argtypes = [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int64, _i64p]
lib.pastix_amd_server_submit_solve.restype = ct.c_int
lib.pastix_amd_server_submit_refactor.argtypes = [ct.c_void_p, ct.c_double, _i64p]
lib.pastix_amd_server_submit_refactor.restype = ct.c_int
lib.pastix_amd_server_wait.argtypes = [ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_double), _i64p]
lib.pastix_amd_server_wait.restype = ct.c_int
lib.pastix_amd_server_stats.argtypes = [ct.c_void_p, _i64p]
lib.pastix_amd_server_stats.restype = ct.c_int
lib.pastix_amd_server_memory.argtypes = [ct.c_void_p, _i64p]
lib.pastix_amd_server_memory.restype = ct.c_int
lib.pastix_amd_server_set_budget.argtypes = [ct.c_void_p, ct.c_int64]
lib.pastix_amd_server_set_budget.restype = ct.c_int
lib.pastix_amd_server_destroy.argtypes = [ct.POINTER(ct.c_void_p)]
lib.pastix_amd_server_destroy.restype = ct.c_int


def device_bytes(pastix_data):
    v = ct.c_int64()
    _check("pastix_amd_device_bytes", lib.pastix_amd_device_bytes(pastix_data.ptr, ct.byref(v)))
    return int(v.value)


class Server:
    """Serving loop over slots (instances factored on the same pattern with
    resident input): continuous batching of solves, refactorisations of
    A + shift*I overlapped with solves on other slots."""

    def __init__(self, slots, max_rhs=64):
        arr = (_P * len(slots))(*[h.ptr for h in slots])
        self._slots = slots
        self.ptr = ct.c_void_p()
        _check("pastix_amd_server_create", lib.pastix_amd_server_create(ct.byref(self.ptr), arr, len(slots), max_rhs))
        self._bufs = {}

    def submit_solve(self, b):
        """b: n x nrhs FP64, overwritten with the solution; returns a ticket."""
        b = np.asfortranarray(b, dtype=np.float64)
        if b.ndim == 1:
            b = b.reshape(-1, 1, order="F")
        t = ct.c_int64()
        _check("pastix_amd_server_submit_solve",
               lib.pastix_amd_server_submit_solve(self.ptr, b.shape[1], b.ctypes.data, b.shape[0], ct.byref(t)))
        self._bufs[t.value] = b
        return t.value

    def submit_refactor(self, shift):
        t = ct.c_int64()
        _check("pastix_amd_server_submit_refactor",
               lib.pastix_amd_server_submit_refactor(self.ptr, float(shift), ct.byref(t)))
        return t.value

    def wait(self, ticket):
        """Returns (result array or None, latency seconds, matrix version)."""
        lat = ct.c_double()
        ver = ct.c_int64()
        rc = lib.pastix_amd_server_wait(self.ptr, ticket, ct.byref(lat), ct.byref(ver))
        if rc != 0:
            raise PastixError("pastix_amd_server_wait", rc)
        return self._bufs.pop(ticket, None), float(lat.value), int(ver.value)

    def set_budget(self, nbytes):
        """Device memory the merged solves' work blocks may hold (all slots)."""
        _check("pastix_amd_server_set_budget", lib.pastix_amd_server_set_budget(self.ptr, int(nbytes)))

    def memory(self):
        out = np.zeros(4, dtype=np.int64)
        _check("pastix_amd_server_memory", lib.pastix_amd_server_memory(self.ptr, out.ctypes.data_as(_i64p)))
        return {"budget": int(out[0]), "held": int(out[1]), "peak": int(out[2]), "waits": int(out[3])}

    def stats(self):
        out = np.zeros(4, dtype=np.int64)
        _check("pastix_amd_server_stats", lib.pastix_amd_server_stats(self.ptr, out.ctypes.data_as(_i64p)))
        return {"batches": int(out[0]), "batched_rhs": int(out[1]), "version": int(out[2]), "pending": int(out[3])}

    def close(self):
        if self.ptr:
            _check("pastix_amd_server_destroy", lib.pastix_amd_server_destroy(ct.byref(self.ptr)))
            self.ptr = ct.c_void_p()


def refactorize(pastix_data):
    _check("pastix_amd_refactorize", lib.pastix_amd_refactorize(pastix_data.ptr))


def solve_device(pastix_data, dbuf, n, nrhs):
    """Solve in place on a DeviceArray holding an n x nrhs column-major block."""
    _check("pastix_amd_solve_device", lib.pastix_amd_solve_device(pastix_data.ptr, nrhs, dbuf.ptr, n))


lib.pastix_amd_refine_device.argtypes = [_P, _I, ct.c_void_p, _I, ct.c_void_p, _I]
lib.pastix_amd_refine_device.restype = ct.c_int
lib.pastix_amd_solve_refine_device.argtypes = [_P, _I, ct.c_void_p, _I, ct.c_void_p, _I]
lib.pastix_amd_solve_refine_device.restype = ct.c_int


def refine_device(pastix_data, db, dx, n, nrhs):
    """Refine the FP64 device block dx (initial guesses) against db, both
    n x nrhs in the user ordering."""
    _check("pastix_amd_refine_device", lib.pastix_amd_refine_device(pastix_data.ptr, nrhs, db.ptr, n, dx.ptr, n))


def solve_refine_device(pastix_data, db, dx, n, nrhs):
    """dx = refined solution of A x = db (FP64 device blocks, user ordering):
    one solve in the factor precision, then the instance's refinement."""
    _check("pastix_amd_solve_refine_device",
           lib.pastix_amd_solve_refine_device(pastix_data.ptr, nrhs, db.ptr, n, dx.ptr, n))


for _f, _args in (("pastix_amd_bcsc_spmv", [_P, ct.c_int, ct.c_double, ct.c_void_p, ct.c_double, ct.c_void_p]),
                  ("pastix_amd_bvec_dot", [_P, _I, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_double)]),
                  ("pastix_amd_bvec_nrm2", [_P, _I, ct.c_void_p, ct.POINTER(ct.c_double)]),
                  ("pastix_amd_bvec_axpy", [_P, _I, ct.c_double, ct.c_void_p, ct.c_void_p]),
                  ("pastix_amd_bvec_scal", [_P, _I, ct.c_double, ct.c_void_p]),
                  ("pastix_amd_bvec_copy", [_P, _I, ct.c_void_p, ct.c_void_p]),
                  ("pastix_amd_bvec_gemv", [_P, _I, _I, ct.c_double, ct.c_void_p, _I, ct.c_void_p, ct.c_double,
                                            ct.c_void_p])):
    getattr(lib, _f).argtypes = _args
    getattr(lib, _f).restype = ct.c_int


def bcsc_spmv(pastix_data, trans_, alpha, dx, beta, dy):
    """dy = alpha op(P A P^T) dx + beta dy on FP64 DeviceArrays (permuted ordering)."""
    _check("pastix_amd_bcsc_spmv", lib.pastix_amd_bcsc_spmv(pastix_data.ptr, trans_, alpha, dx.ptr, beta, dy.ptr))


def bvec_dot(pastix_data, n, dx, dy):
    out = ct.c_double()
    _check("pastix_amd_bvec_dot", lib.pastix_amd_bvec_dot(pastix_data.ptr, n, dx.ptr, dy.ptr, ct.byref(out)))
    return out.value


def bvec_nrm2(pastix_data, n, dx):
    out = ct.c_double()
    _check("pastix_amd_bvec_nrm2", lib.pastix_amd_bvec_nrm2(pastix_data.ptr, n, dx.ptr, ct.byref(out)))
    return out.value


def bvec_axpy(pastix_data, n, alpha, dx, dy):
    _check("pastix_amd_bvec_axpy", lib.pastix_amd_bvec_axpy(pastix_data.ptr, n, alpha, dx.ptr, dy.ptr))


def bvec_scal(pastix_data, n, alpha, dx):
    _check("pastix_amd_bvec_scal", lib.pastix_amd_bvec_scal(pastix_data.ptr, n, alpha, dx.ptr))


def bvec_copy(pastix_data, n, dx, dy):
    _check("pastix_amd_bvec_copy", lib.pastix_amd_bvec_copy(pastix_data.ptr, n, dx.ptr, dy.ptr))


def bvec_gemv(pastix_data, m, n, alpha, dA, lda, dx, beta, dy):
    _check("pastix_amd_bvec_gemv",
           lib.pastix_amd_bvec_gemv(pastix_data.ptr, m, n, alpha, dA.ptr, lda, dx.ptr, beta, dy.ptr))


class DeviceArray:
    """Minimal HIP device buffer (hipMalloc/hipMemcpy through ctypes) used to
    hand device-resident right-hand sides to pastix_amd_solve_device."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = ct.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
            h.hipFree.argtypes = [ct.c_void_p]
            h.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        self.nbytes = host.nbytes
        self.dtype = host.dtype
        self.shape = host.shape
        self.ptr = ct.c_void_p()
        rc = self.hip().hipMalloc(ct.byref(self.ptr), self.nbytes)
        if rc != 0:
            raise MemoryError(f"hipMalloc failed ({rc})")
        self.upload(host)

    def upload(self, host):
        h = np.asfortranarray(host)
        rc = self.hip().hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, 1)
        assert rc == 0

    def copy_from(self, other):
        """Device-to-device copy of another buffer of the same size."""
        if other.nbytes != self.nbytes:
            raise ValueError("copy_from: sizes differ")
        rc = self.hip().hipMemcpy(self.ptr, other.ptr, self.nbytes, 3)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy device-to-device failed ({rc})")
        rc = self.hip().hipDeviceSynchronize()
        if rc != 0:
            raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        rc = self.hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2)
        assert rc == 0
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = ct.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
