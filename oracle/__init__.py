"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes loader of oracle/liboracle.so (and its op-counting twin), the CPU
restatement of the reference's interval arithmetic and covering criteria.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.POINTER
_vp, _sz, _i, _d = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double


def _load(name: str) -> ctypes.CDLL:
    path = os.path.join(_DIR, name)
    if not os.path.exists(path):
        raise ImportError(f"{path} missing; run `make -C oracle`")
    L = ctypes.CDLL(path)
    L.or_c2_eval.argtypes = [_vp, _sz, _sz, _vp, _vp, _i]
    L.or_c2_count_ops.argtypes = [_vp, _sz, _sz, _vp]
    L.or_prim.argtypes = [_i, _i, _d, _d]
    L.or_prim.restype = _d
    L.or_tr_critical_ratio.argtypes = [_d, _d, _vp]
    L.or_tr_critical_ratio.restype = None
    L.or_two_disks_maximize_height.argtypes = [_d, _d, _d]
    L.or_two_disks_maximize_height.restype = _d
    L.or_disk_satisfies_size_bound.argtypes = [_d, _d]
    L.or_bound_required_height.argtypes = [_d]
    L.or_bound_required_height.restype = _d
    L.or_intersection.argtypes = [_vp, _vp, _vp]
    L.or_iv_op.argtypes = [_i, _vp, _vp, _vp]
    L.or_iv_op.restype = None
    L.or_rn.argtypes = [_d, _d, _d, _vp]
    L.or_rn.restype = None
    L.or_rsb_verdicts.argtypes = [_vp, _sz, _vp, _vp, _i]
    L.or_rsb_masks.argtypes = [_vp, _sz, _vp, _i]
    L.or_tr_find_critical.argtypes = [_vp, _sz, _vp, _vp, _sz, _vp, _vp, _i]
    L.or_rsb_grid.argtypes = [_vp, _i, _i, _i, _vp]
    L.or_rsb_grid.restype = _sz
    L.or_rsb_bnb.argtypes = [_vp, _sz, _i, _vp, _sz, _vp, _vp, _vp, _vp, _i]
    L.or_rsb_find_critical.argtypes = [_vp, _vp, _vp, _sz, _vp, _vp, _i]
    L.or_rsb_bisect.argtypes = [_vp, _sz, _vp]
    L.or_rsb_bisect.restype = None
    L.or_subintervals.argtypes = [_d, _d, _i, _vp]
    L.or_subintervals.restype = None
    L.or_tr_bisect.argtypes = [_vp, _sz, _vp]
    L.or_tr_bisect.restype = None
    L.or_tr_merge_adjacent_r1.argtypes = [_vp, _sz, _vp]
    L.or_tr_merge_adjacent_r1.restype = _sz
    L.or_tr_leaf_masks.argtypes = [_vp, _sz, _vp]
    L.or_rsb_c14_probe.argtypes = [_vp, _sz, _vp, _i, _vp, _vp]
    L.or_rsb_c14_probe.restype = None
    L.or_tr_leaf_masks.restype = None
    return L


lib = _load("liboracle.so")
_count_lib = None


def count_lib() -> ctypes.CDLL:
    global _count_lib
    if _count_lib is None:
        _count_lib = _load("liboracle_count.so")
    return _count_lib


def c2_eval(soa: np.ndarray, threads: int = 0):
    soa = np.ascontiguousarray(soa, dtype=np.float64)
    n = soa.shape[1]
    mask = np.empty(n, dtype=np.uint32)
    first = np.empty(n, dtype=np.uint8)
    lib.or_c2_eval(soa.ctypes.data, n, n, mask.ctypes.data, first.ctypes.data, threads)
    return mask, first


def c2_count_ops(soa: np.ndarray) -> np.ndarray:
    """Per-criterion directed-FP64-op totals over the boxes (counting build)."""
    soa = np.ascontiguousarray(soa, dtype=np.float64)
    ops = np.zeros(10, dtype=np.uint64)
    rc = count_lib().or_c2_count_ops(soa.ctypes.data, soa.shape[1], soa.shape[1], ops.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle built without op counting")
    return ops


PRIM_OPS = {"add": 0, "mul": 1, "div": 2, "sqrt": 3}


def prim(op: str, direction: str, a: float, b: float = 0.0) -> float:
    return lib.or_prim(PRIM_OPS[op], 1 if direction == "ru" else 0, a, b)


def iv_op(op: int, a, b=(0.0, 0.0)):
    x = (ctypes.c_double * 2)(*a)
    y = (ctypes.c_double * 2)(*b)
    out = (ctypes.c_double * 3)()
    lib.or_iv_op(op, ctypes.addressof(x), ctypes.addressof(y), ctypes.addressof(out))
    return (out[0], out[1]), bool(out[2])


def rsb_verdicts(cells: np.ndarray, threads: int = 0):
    """Reference-order verdict of (n, 14) rectangle_size_bound cells."""
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
    n = cells.shape[0]
    out = np.empty(n, dtype=np.int32)
    trap = np.empty(n, dtype=np.uint8)
    lib.or_rsb_verdicts(cells.ctypes.data, n, out.ctypes.data, trap.ctypes.data, threads)
    return out, trap


def rsb_masks(cells: np.ndarray, threads: int = 0) -> np.ndarray:
    """Every rectangle_size_bound step evaluated on each cell (bit k = step k,
    bit 20+k = trap in step k)."""
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
    out = np.empty(cells.shape[0], dtype=np.uint64)
    lib.or_rsb_masks(cells.ctypes.data, cells.shape[0], out.ctypes.data, threads)
    return out


def tr_find_critical(cells: np.ndarray, counts, threads: int = 0, cap: int = 1 << 22):
    """Unproven tight_rectangle leaves of a batch of (n, 10) cells, enumerated
    with `counts` = (n1, n2, n3, n4) pieces.  Returns (leaves (m, 10), trap)."""
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 10)
    cnt = np.ascontiguousarray(counts, dtype=np.int32)
    while True:
        out = np.empty((cap, 10))
        n_out = ctypes.c_size_t()
        trap = ctypes.c_int()
        lib.or_tr_find_critical(cells.ctypes.data, cells.shape[0], cnt.ctypes.data, out.ctypes.data, cap,
                                ctypes.addressof(n_out), ctypes.addressof(trap), threads)
        if n_out.value <= cap:
            return out[: n_out.value].copy(), bool(trap.value)
        cap = n_out.value


def subintervals(lb: float, ub: float, n: int) -> np.ndarray:
    """(n, 2) pieces get_subinterval([lb, ub], i, n) (subintervals.hpp:32-43)."""
    out = np.empty((n, 2))
    lib.or_subintervals(lb, ub, n, out.ctypes.data)
    return out


def tr_bisect(cells: np.ndarray) -> np.ndarray:
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 10)
    out = np.empty((2 * cells.shape[0], 10))
    lib.or_tr_bisect(cells.ctypes.data, cells.shape[0], out.ctypes.data)
    return out


def tr_leaf_masks(cells: np.ndarray) -> np.ndarray:
    """Leaf-criterion bits of tight_rectangle leaves ((n, 10) records): see
    or_tr_leaf_masks."""
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    out = np.zeros(cells.shape[0], dtype=np.uint32)
    lib.or_tr_leaf_masks(cells.ctypes.data, cells.shape[0], out.ctypes.data)
    return out


def tr_merge_adjacent_r1(cells: np.ndarray) -> np.ndarray:
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 10)
    out = np.empty_like(cells)
    m = lib.or_tr_merge_adjacent_r1(cells.ctypes.data, cells.shape[0], out.ctypes.data)
    return out[:m].copy()


def rsb_grid(outer, nla: int, nr1: int, nr2: int) -> np.ndarray:
    """Nonempty initial cells of the reference's (lambda, r1, r2) host loop
    over `outer` (14 doubles), in loop order."""
    outer = np.ascontiguousarray(outer, dtype=np.float64).reshape(14)
    out = np.empty((nla * nr1 * nr2, 14))
    m = lib.or_rsb_grid(outer.ctypes.data, nla, nr1, nr2, out.ctypes.data)
    return out[:m].copy()


def rsb_c14_probe(cells: np.ndarray, pieces: np.ndarray):
    """C14's stages on (n, 14) cells with (n, np, 2) r7 pieces: (n, 7) widths
    and ranges, (n, np) piece-strategy bits (or_rsb_c14_probe)."""
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    pieces = np.ascontiguousarray(pieces, dtype=np.float64)
    n, npc = cells.shape[0], pieces.shape[1]
    out = np.zeros((n, 7))
    bits = np.zeros((n, npc), dtype=np.uint32)
    lib.or_rsb_c14_probe(cells.ctypes.data, n, pieces.ctypes.data, npc, out.ctypes.data, bits.ctypes.data)
    return out, bits


def rsb_bnb(cells: np.ndarray, max_depth: int, threads: int = 0, cap: int = 1 << 20):
    """Widest-dimension branch-and-bound; returns (unproven cells at
    max_depth, cells evaluated per depth, verdict histogram (depth, 21),
    number of cells that reached a trap site)."""
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
    while True:
        out = np.empty((cap, 14))
        n_out = ctypes.c_size_t()
        trap = ctypes.c_int()
        per = np.zeros(max_depth + 1, dtype=np.uint64)
        hist = np.zeros((max_depth + 1, 21), dtype=np.uint64)
        lib.or_rsb_bnb(cells.ctypes.data, cells.shape[0], max_depth, out.ctypes.data, cap, ctypes.addressof(n_out),
                       per.ctypes.data, hist.ctypes.data, ctypes.addressof(trap), threads)
        if n_out.value <= cap:
            return out[: n_out.value].copy(), per, hist, int(trap.value)
        cap = n_out.value


def rsb_find_critical(cell, counts, threads: int = 0, cap: int = 1 << 20):
    """Unproven leaves of one rectangle_size_bound kernel launch with
    `counts` pieces for r3..r6.  Returns (leaves (m, 14), trap)."""
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(14)
    cnt = np.ascontiguousarray(counts, dtype=np.int32)
    while True:
        out = np.empty((cap, 14))
        n_out = ctypes.c_size_t()
        trap = ctypes.c_int()
        lib.or_rsb_find_critical(cell.ctypes.data, cnt.ctypes.data, out.ctypes.data, cap, ctypes.addressof(n_out),
                                 ctypes.addressof(trap), threads)
        if n_out.value <= cap:
            return out[: n_out.value].copy(), bool(trap.value)
        cap = n_out.value


def rsb_bisect(cells: np.ndarray) -> np.ndarray:
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
    out = np.empty((2 * cells.shape[0], 14))
    lib.or_rsb_bisect(cells.ctypes.data, cells.shape[0], out.ctypes.data)
    return out


def rsb_count_ops(cells: np.ndarray):
    """(directed ops per cell, verdict per cell) from the counting build."""
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
    ops = np.zeros(cells.shape[0], dtype=np.uint64)
    v = np.zeros(cells.shape[0], dtype=np.int32)
    L = count_lib()
    L.or_rsb_count_ops.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    if L.or_rsb_count_ops(cells.ctypes.data, cells.shape[0], ops.ctypes.data, v.ctypes.data) != 0:
        raise RuntimeError("oracle built without op counting")
    return ops, v


def canonical(cells: np.ndarray) -> np.ndarray:
    """Records sorted lexicographically by their doubles (atomic append order
    on the GPU is not deterministic; the set is)."""
    cells = np.asarray(cells)
    if cells.shape[0] == 0:
        return cells
    return cells[np.lexsort(cells.T[::-1])]


def rn_helpers(a: float, b: float, c: float):
    out = (ctypes.c_double * 3)()
    lib.or_rn(a, b, c, ctypes.addressof(out))
    return tuple(out)


def tr_critical_ratio(lb: float, ub: float):
    out = (ctypes.c_double * 2)()
    lib.or_tr_critical_ratio(lb, ub, ctypes.addressof(out))
    return out[0], out[1]


def intersection(c1, c2):
    a = (ctypes.c_double * 3)(*c1)
    b = (ctypes.c_double * 3)(*c2)
    out = (ctypes.c_double * 8)()
    ok = lib.or_intersection(ctypes.addressof(a), ctypes.addressof(b), ctypes.addressof(out))
    pts = [((out[0], out[1]), (out[2], out[3])), ((out[4], out[5]), (out[6], out[7]))]
    return bool(ok), pts
