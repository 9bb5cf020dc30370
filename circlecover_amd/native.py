"""ctypes bindings of libcirclecover_amd.so (the C ABI in include/circlecover_amd.h)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_PKG, "libcirclecover_amd.so")

CC_OK, CC_E_HIP, CC_E_OVERFLOW, CC_E_TRAP, CC_E_INTERRUPTED, CC_E_ARG, CC_E_NODEVICE = range(7)
CC_BATCH_C2 = 0
CC_PREC_FP64, CC_PREC_FP32_PRE = 0, 1
NOALLOC_MAX_BOXES = 262144  # CC_NOALLOC_MAX_BOXES

# Criterion names of the ten-criterion 4-D box set, in evaluation order
# (bit k of a verdict mask).
C2_CRITERIA = (
    "tr.r1_strategies",
    "tr.r1_r2_strategies_simple",
    "tr.r1_r2_strategies_stacked",
    "tr.r1_r2_r3_strategy_horizontal",
    "tr.r1_r2_r3_strategy_cover_strip",
    "rsb.shortcut_even_split_recursion",
    "rsb.shortcut_uneven_split_recursion/3",
    "rsb.shortcut_uneven_split_recursion/4",
    "rsb.shortcut_r1_r2_opposite_corners_recursion",
    "rsb.shortcut_r1_r2_vertical_strip",
)


class CCError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"circlecover_amd error {code}: {msg}")
        self.code = code


def _load() -> ctypes.CDLL:
    if not os.path.exists(library_path):
        raise ImportError(f"{library_path} is not built; run `python circlecover_amd/build.py`")
    return ctypes.CDLL(library_path)


lib = _load()

_vp, _sz, _i, _u64, _d = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64, ctypes.c_double
_P = ctypes.POINTER
lib.cc_init.argtypes = [_i, _P(_vp)]
lib.cc_destroy.argtypes = [_vp]
lib.cc_last_error.argtypes = [_vp]
lib.cc_last_error.restype = ctypes.c_char_p
lib.cc_synchronize.argtypes = [_vp]
lib.cc_dev_alloc.argtypes = [_vp, _sz, _P(_vp)]
lib.cc_dev_free.argtypes = [_vp, _vp]
lib.cc_h2d.argtypes = [_vp, _vp, _vp, _sz]
lib.cc_d2h.argtypes = [_vp, _vp, _vp, _sz]
lib.cc_timer_start.argtypes = [_vp]
lib.cc_copy_async.argtypes = [_vp, _vp, _vp, _sz]
lib.cc_timer_stop.argtypes = [_vp, _P(ctypes.c_float)]
lib.cc_eval_batch.argtypes = [_vp, _i, _i, _vp, _sz, _sz, _vp, _vp]
lib.cc_eval_batch_async.argtypes = [_vp, _i, _i, _vp, _sz, _sz, _vp, _vp]
lib.cc_eval_batch_noalloc_async.argtypes = [_vp, _i, _i, _vp, _sz, _sz, _vp, _vp]
lib.cc_gen_c2_boxes.argtypes = [_u64, _sz, _vp]
lib.cc_gen_c2_boxes.restype = None
lib.cc_prim_selftest.argtypes = [_vp, _vp, _vp, _sz, _vp]
lib.cc_prim_rn_selftest.argtypes = [_vp, _vp, _vp, _vp, _sz, _vp]
lib.cc_rsb_eval_cells.argtypes = [_vp, _vp, _sz, _vp, _vp]
lib.cc_rsb_eval_cells_async.argtypes = [_vp, _vp, _sz, _vp, _vp]
lib.cc_rsb_eval_masks.argtypes = [_vp, _vp, _sz, _vp]
lib.cc_subintervals.argtypes = [_d, _d, _i, _vp]
lib.cc_subintervals.restype = None
lib.cc_rsb_grid.argtypes = [_vp, _i, _i, _i, _vp, _sz]
lib.cc_rsb_grid.restype = _sz
lib.cc_rsb_find_critical_in.argtypes = [_vp, _vp, _i, _vp, _sz, _P(_sz)]
lib.cc_rsb_enumerate.argtypes = [_vp, _vp, _sz, _vp, _vp, _sz, _P(_sz)]
lib.cc_rsb_bnb.argtypes = [_vp, _vp, _sz, _i, _vp, _sz, _P(_sz), _vp]
lib.cc_rsb_bnb_exact.argtypes = [_vp, _vp, _sz, _i, _vp, _sz, _P(_sz), _vp]
class FastParams(ctypes.Structure):
    _fields_ = [("cpw_a", ctypes.c_int), ("cpw_b", ctypes.c_int), ("blocks_per_cu", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("waves_per_simd", ctypes.c_int)]


lib.cc_rsb_bnb_batched.argtypes = [_vp, _vp, _sz, _i, _vp, _vp, _sz, _P(_sz), _vp]
lib.cc_run_bnb.argtypes = [_vp, _vp, _P(_vp), _P(_sz), _vp]
lib.cc_free.argtypes = [_vp]
lib.cc_free.restype = None
lib.cc_interrupt.argtypes = [_vp]
lib.cc_set_enumeration_form.argtypes = [_vp, _i]
lib.cc_rsb_autotune.argtypes = [_vp, _vp, _sz, _i, _vp, _P(ctypes.c_double)]
RANKER_PARAMS = 5451
RANKER_CLASSES = 11  # criteria 0..9 first proving a box, 10: none proves it


class Ranker(ctypes.Structure):
    _fields_ = [("params", ctypes.c_float * RANKER_PARAMS), ("norm", ctypes.c_float * 16)]

    def params_array(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.params).copy()


class RsbRanker(ctypes.Structure):
    """cc_rsb_ranker: the network and, per class, the step it stands for (-1: the reference's order)."""
    _fields_ = [("net", Ranker), ("class_step", ctypes.c_int32 * 11)]


class FirstStats(ctypes.Structure):
    _fields_ = [("proven", ctypes.c_uint64), ("criterion_evals", ctypes.c_uint64), ("evals_proven", ctypes.c_uint64),
                ("rank_ms", ctypes.c_double), ("eval_ms", ctypes.c_double), ("second_stage", ctypes.c_uint64)]


lib.cc_ranker_train.argtypes = [_vp, _vp, _sz, _sz, _vp, _i, _i, ctypes.c_float, _u64, _vp, _vp]
lib.cc_ranker_gradient.argtypes = [_vp, _vp, _sz, _sz, _vp, _vp, _i, _vp, _vp, _vp]
lib.cc_ranker_rank.argtypes = [_vp, _vp, _sz, _sz, _vp, _vp, _vp]
lib.cc_c2_eval_first.argtypes = [_vp, _vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp]
lib.cc_rsb_bnb_log.argtypes = [_vp, _vp, _sz, _i, _vp, _sz, _P(_sz), _vp, _sz, _P(_sz), _vp]
lib.cc_rsb_ranker_train.argtypes = [_vp, _vp, _sz, _i, _i, ctypes.c_float, _u64, _vp, _vp]
lib.cc_rsb_bnb_learned.argtypes = [_vp, _vp, _sz, _i, _vp, _vp, _vp, _sz, _P(_sz), _vp]
lib.cc_format_interval.argtypes = [_vp, ctypes.c_char_p, _sz, _P(_sz)]
lib.cc_parse_interval.argtypes = [ctypes.c_char_p, _sz, _vp, _P(_sz)]
lib.cc_c2_eval_first_ex.argtypes = [_vp, _vp, _sz, _sz, _vp, _i, _vp, _vp, _vp, _vp]
lib.cc_c2_eval_first_f32.argtypes = [_vp, _vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp]
lib.cc_c2_eval_first_f32_ex.argtypes = [_vp, _vp, _sz, _sz, _i, _vp, _vp, _vp, _vp, _vp, _vp]
lib.cc_c2_f32_report.argtypes = [_vp, _vp, _sz, _sz, _vp, _vp]
lib.cc_tr_find_critical_in.argtypes = [_vp, _vp, _i, _vp, _sz, _P(_sz)]
lib.cc_tr_enumerate.argtypes = [_vp, _vp, _sz, _vp, _vp, _sz, _P(_sz)]
lib.cc_tr_bnb.argtypes = [_vp, _vp, _sz, _vp, _i, _vp, _sz, _P(_sz), _vp]
lib.cc_tr_report_criticals.argtypes = [_vp, _vp, _sz, ctypes.c_char_p, _sz, _P(_sz)]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def subintervals(lb: float, ub: float, n: int) -> np.ndarray:
    """(n, 2) pieces of [lb, ub] as get_subinterval makes them."""
    out = np.empty((n, 2))
    lib.cc_subintervals(lb, ub, n, _ptr(out))
    return out


DBL_MAX = float(np.finfo(np.float64).max)
TR_LAMBDA_MAX = 2.0898841580413818342  # tight_rectangle.cu:462


def tr_initial_cells(la_lo: float, la_hi: float, n: int) -> np.ndarray:
    """(n, 10) tight_rectangle cells: lambda pieces of [la_lo, la_hi], every
    squared radius in [0, DBL_MAX] (tight_rectangle.cu:460-465)."""
    la = subintervals(la_lo, la_hi, n)
    return np.concatenate([la, np.tile([0.0, DBL_MAX], (n, 4))], axis=1)


RSB_DISK_WEIGHT_MAX = 9.0 / 64.0  # rectangle_size_bound/values.hpp:33


def rsb_grid(la_lo: float, la_hi: float, nla: int, nr1: int, nr2: int) -> np.ndarray:
    """(m, 14) initial rectangle_size_bound cells over lambda in [la_lo, la_hi]
    and squared radii in [0, 9/64], as the reference's host loop builds them."""
    outer = np.array([la_lo, la_hi] + [0.0, RSB_DISK_WEIGHT_MAX] * 6)
    out = np.empty((nla * nr1 * nr2, 14))
    m = lib.cc_rsb_grid(_ptr(outer), nla, nr1, nr2, _ptr(out), out.shape[0])
    return out[:m]


class BnbStats(ctypes.Structure):
    _fields_ = [("cells_per_depth", ctypes.c_uint64 * 64), ("verdict_hist", ctypes.c_uint64 * 21),
                ("trap_cells", ctypes.c_uint64), ("cell_evals", ctypes.c_uint64), ("depth_reached", ctypes.c_int),
                ("device_ms", ctypes.c_double), ("wall_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        d = self.depth_reached
        return {"cells_per_depth": list(self.cells_per_depth)[: d + 1], "verdict_hist": list(self.verdict_hist),
                "trap_cells": self.trap_cells, "cell_evals": self.cell_evals, "depth_reached": d,
                "device_ms": self.device_ms, "wall_ms": self.wall_ms}


CC_BNB_PLAIN, CC_BNB_BATCHED, CC_BNB_LEARNED, CC_BNB_PLAIN_EXACT = 0, 1, 2, 3
CC_ENUM_BATCHED, CC_ENUM_PLAIN, CC_ENUM_PLAIN_EXACT = 0, 1, 2


class BnbConfig(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_void_p), ("ncells", ctypes.c_size_t), ("la_lb", ctypes.c_double),
                ("la_ub", ctypes.c_double), ("nla", ctypes.c_int), ("nr1", ctypes.c_int), ("nr2", ctypes.c_int),
                ("max_depth", ctypes.c_int), ("form", ctypes.c_int), ("params", ctypes.c_void_p),
                ("ranker", ctypes.c_void_p)]


class BnbRecord(ctypes.Structure):
    """cc_bnb_record: a logged cell (14 doubles), its first proving step (-1: unproven) and depth."""
    _fields_ = [("cell", ctypes.c_double * 14), ("first_step", ctypes.c_int32), ("depth", ctypes.c_int32)]


BNB_RECORD_DTYPE = np.dtype([("cell", np.float64, (14,)), ("first_step", np.int32), ("depth", np.int32)])


def gen_c2_boxes(seed: int, n: int) -> np.ndarray:
    """Synthetic 4-D boxes as an (8, n) float64 SoA array."""
    out = np.empty((8, n), dtype=np.float64)
    lib.cc_gen_c2_boxes(seed & 0xFFFFFFFFFFFFFFFF, n, _ptr(out))
    return out


class DeviceBuffer:
    def __init__(self, prover: "Prover", nbytes: int):
        self.prover, self.nbytes = prover, nbytes
        p = ctypes.c_void_p()
        prover._check(lib.cc_dev_alloc(prover.ctx, nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, a: np.ndarray) -> None:
        assert a.nbytes <= self.nbytes and a.flags.c_contiguous
        self.prover._check(lib.cc_h2d(self.prover.ctx, self.ptr, _ptr(a), a.nbytes))

    def download(self, a: np.ndarray) -> np.ndarray:
        assert a.nbytes <= self.nbytes and a.flags.c_contiguous
        self.prover._check(lib.cc_d2h(self.prover.ctx, _ptr(a), self.ptr, a.nbytes))
        return a

    def free(self) -> None:
        if self.ptr:
            lib.cc_dev_free(self.prover.ctx, self.ptr)
            self.ptr = None


class Prover:
    """One device context (HIP stream, kernel module, device constants)."""

    def __init__(self, device: int = 0):
        p = ctypes.c_void_p()
        rc = lib.cc_init(device, ctypes.byref(p))
        if rc != CC_OK:
            raise CCError(rc, "cc_init failed (is a gfx950 GPU visible?)")
        self.ctx = p.value

    def close(self) -> None:
        if self.ctx:
            lib.cc_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != CC_OK:
            raise CCError(rc, lib.cc_last_error(self.ctx).decode(errors="replace"))

    # ---- memory / timing
    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def synchronize(self) -> None:
        self._check(lib.cc_synchronize(self.ctx))

    def copy_async(self, dst: "DeviceBuffer", src: "DeviceBuffer", nbytes: int) -> None:
        self._check(lib.cc_copy_async(self.ctx, dst.ptr, src.ptr, nbytes))

    def timer_start(self) -> None:
        self._check(lib.cc_timer_start(self.ctx))

    def timer_stop(self) -> float:
        ms = ctypes.c_float()
        self._check(lib.cc_timer_stop(self.ctx, ctypes.byref(ms)))
        return ms.value

    # ---- batched criteria
    def eval_batch(self, soa: np.ndarray, precision: int = CC_PREC_FP64):
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        assert soa.ndim == 2 and soa.shape[0] == 8
        n = soa.shape[1]
        mask = np.empty(n, dtype=np.uint32)
        first = np.empty(n, dtype=np.uint8)
        self._check(lib.cc_eval_batch(self.ctx, CC_BATCH_C2, precision, _ptr(soa), n, n, _ptr(mask), _ptr(first)))
        return mask, first

    def eval_batch_async(self, d_soa: DeviceBuffer, n: int, ld: int, d_mask: DeviceBuffer, d_first: DeviceBuffer,
                         precision: int = CC_PREC_FP64) -> None:
        self._check(lib.cc_eval_batch_async(self.ctx, CC_BATCH_C2, precision, d_soa.ptr, n, ld, d_mask.ptr,
                                            d_first.ptr))

    def eval_batch_noalloc_async(self, d_soa: DeviceBuffer, n: int, ld: int, d_mask: DeviceBuffer,
                                 d_first: DeviceBuffer, precision: int = CC_PREC_FP64) -> None:
        """cc_eval_batch_noalloc_async: eval_batch_async with the library's
        static per-GPU scratch instead of allocated buffers (n <= NOALLOC_MAX_BOXES)."""
        self._check(lib.cc_eval_batch_noalloc_async(self.ctx, CC_BATCH_C2, precision, d_soa.ptr, n, ld, d_mask.ptr,
                                                    d_first.ptr))

    def rsb_eval_cells(self, cells: np.ndarray):
        """cells: (n, 14) float64 (la, r1..r6 as lb,ub).  Returns (verdict int32, trap uint8)."""
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
        n = cells.shape[0]
        verdict = np.empty(n, dtype=np.int32)
        trap = np.empty(n, dtype=np.uint8)
        self._check(lib.cc_rsb_eval_cells(self.ctx, _ptr(cells), n, _ptr(verdict), _ptr(trap)))
        return verdict, trap

    # ---- tight_rectangle enumeration
    def _grow(self, call, cap: int, width: int, with_trap: bool = False):
        """Runs call(out, cap, n_out) and re-runs it with the reported size on
        CC_E_OVERFLOW (the reference's relaunch protocol).  With with_trap,
        CC_E_TRAP (output complete, a cell reached a trap site) is returned as
        a flag instead of raised."""
        while True:
            out = np.empty((max(cap, 1), width))
            n = ctypes.c_size_t()
            rc = call(out, cap, n)
            if rc == CC_E_OVERFLOW and n.value > cap:
                cap = n.value
                continue
            trap = with_trap and rc == CC_E_TRAP
            if not trap:
                self._check(rc)
            return (out[: n.value], trap) if with_trap else out[: n.value]

    def tr_find_critical_in(self, cell, is_init: bool, cap: int = 1 << 16) -> np.ndarray:
        cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(10)
        return self._grow(lambda o, c, n: lib.cc_tr_find_critical_in(self.ctx, _ptr(cell), int(is_init), _ptr(o), c,
                                                                     ctypes.byref(n)), cap, 10)

    def tr_enumerate(self, cells: np.ndarray, counts, cap: int = 1 << 16) -> np.ndarray:
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 10)
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        return self._grow(lambda o, c, n: lib.cc_tr_enumerate(self.ctx, _ptr(cells), cells.shape[0], _ptr(cnt),
                                                              _ptr(o), c, ctypes.byref(n)), cap, 10)

    def tr_report_criticals(self, cells: np.ndarray) -> str:
        """The tight_rectangle report for a final critical set
        (tight_rectangle.cu:498-536), its Max_distances lines computed on the
        device."""
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 10)
        n = _sz(0)
        cap = 1 << 14
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = lib.cc_tr_report_criticals(self.ctx, _ptr(cells), cells.shape[0], buf, cap, ctypes.byref(n))
            if rc == CC_E_OVERFLOW and n.value >= cap:
                cap = n.value + 1
                continue
            self._check(rc)
            return buf.value.decode("utf-8")

    def tr_bnb(self, cells: np.ndarray, counts, levels: int, cap: int = 1 << 16):
        """Returns (unproven leaves, unproven count after each level)."""
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 10)
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        per = np.zeros(levels + 1, dtype=np.uint64)
        leaves = self._grow(lambda o, c, n: lib.cc_tr_bnb(self.ctx, _ptr(cells), cells.shape[0], _ptr(cnt), levels,
                                                          _ptr(o), c, ctypes.byref(n), _ptr(per)), cap, 10)
        return leaves, per

    def set_enumeration_form(self, form: int) -> None:
        """CC_ENUM_BATCHED (0), CC_ENUM_PLAIN (1) or CC_ENUM_PLAIN_EXACT (2) for
        both proofs' enumerations (cc_set_enumeration_form)."""
        self._check(lib.cc_set_enumeration_form(self.ctx, int(form)))

    # ---- rectangle_size_bound enumeration / branch-and-bound
    def rsb_find_critical_in(self, cell, is_init: bool, cap: int = 1 << 16) -> np.ndarray:
        cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(14)
        return self._grow(lambda o, c, n: lib.cc_rsb_find_critical_in(self.ctx, _ptr(cell), int(is_init), _ptr(o), c,
                                                                      ctypes.byref(n)), cap, 14)

    def rsb_enumerate(self, cells: np.ndarray, counts, cap: int = 1 << 16, with_trap: bool = False):
        """Unproven leaves; with with_trap, (leaves, trap flag) instead of
        raising CC_E_TRAP."""
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
        cnt = np.ascontiguousarray(counts, dtype=np.int32)
        return self._grow(lambda o, c, n: lib.cc_rsb_enumerate(self.ctx, _ptr(cells), cells.shape[0], _ptr(cnt),
                                                               _ptr(o), c, ctypes.byref(n)), cap, 14, with_trap)

    def rsb_bnb(self, cells: np.ndarray, max_depth: int, cap: int = 1 << 16, exact: bool = False):
        """Level-synchronous branch-and-bound (exact: with the batched form's
        exact shortcuts, cc_rsb_bnb_exact).  Returns (unproven cells at
        max_depth, BnbStats)."""
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
        st = BnbStats()
        fn = lib.cc_rsb_bnb_exact if exact else lib.cc_rsb_bnb
        out = self._grow(lambda o, c, n: fn(self.ctx, _ptr(cells), cells.shape[0], max_depth, _ptr(o), c,
                                            ctypes.byref(n), ctypes.byref(st)), cap, 14)
        return out, st

    def run_bnb(self, max_depth: int, form: int = CC_BNB_BATCHED, cells: np.ndarray | None = None,
                grid=None, params=None, ranker=None):
        """cc_run_bnb over `cells` ((m, 14)) or grid = (la_lb, la_ub, nla, nr1, nr2).
        Returns (unproven cells in lexicographic order, BnbStats)."""
        cfg = BnbConfig()
        keep = []
        if cells is not None:
            cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
            keep.append(cells)
            cfg.cells, cfg.ncells = _ptr(cells), cells.shape[0]
        else:
            cfg.la_lb, cfg.la_ub, cfg.nla, cfg.nr1, cfg.nr2 = grid
        cfg.max_depth, cfg.form = max_depth, form
        if ranker is not None:
            keep.append(ranker)
            cfg.ranker = ctypes.cast(ctypes.byref(ranker), ctypes.c_void_p)
        fp = FastParams(*params) if params is not None else None
        cfg.params = ctypes.cast(ctypes.byref(fp), ctypes.c_void_p) if fp is not None else None
        out, n, st = ctypes.c_void_p(), ctypes.c_size_t(), BnbStats()
        self._check(lib.cc_run_bnb(self.ctx, ctypes.byref(cfg), ctypes.byref(out), ctypes.byref(n), ctypes.byref(st)))
        try:
            res = np.empty((n.value, 14))
            if n.value:
                ctypes.memmove(res.ctypes.data, out.value, res.nbytes)
        finally:
            lib.cc_free(out)
        return res, st

    def interrupt(self) -> None:
        """cc_interrupt: may be called from another thread while a call runs."""
        self._check(lib.cc_interrupt(self.ctx))

    def rsb_bnb_batched(self, cells: np.ndarray, max_depth: int, params=None, cap: int = 1 << 16):
        """Batched branch-and-bound; params = (cpw_a, cpw_b, blocks_per_cu,
        use_graph, waves_per_simd) or None for the tuned default.  Returns
        (unproven, BnbStats)."""
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
        st = BnbStats()
        pp = ctypes.byref(FastParams(*params)) if params is not None else None
        out = self._grow(lambda o, c, n: lib.cc_rsb_bnb_batched(self.ctx, _ptr(cells), cells.shape[0], max_depth, pp,
                                                                _ptr(o), c, ctypes.byref(n), ctypes.byref(st)),
                         cap, 14)
        return out, st

    def rsb_autotune(self, cells: np.ndarray, max_depth: int):
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
        best = FastParams()
        ms = ctypes.c_double()
        self._check(lib.cc_rsb_autotune(self.ctx, _ptr(cells), cells.shape[0], max_depth, ctypes.byref(best),
                                        ctypes.byref(ms)))
        return (best.cpw_a, best.cpw_b, best.blocks_per_cu, best.use_graph, best.waves_per_simd), ms.value

    # ---- learned criterion ranker
    def rsb_bnb_log(self, cells: np.ndarray, max_depth: int, log_cap: int | None = None):
        """Level-synchronous branch-and-bound that logs every evaluated cell.
        Returns (unproven, BnbStats, records) with records a BNB_RECORD_DTYPE array."""
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
        st = BnbStats()
        cap = log_cap if log_cap is not None else 8 * cells.shape[0] + 1024
        while True:
            log = np.zeros(cap, dtype=BNB_RECORD_DTYPE)
            n_log = ctypes.c_size_t()
            out = self._grow(lambda o, c, n: lib.cc_rsb_bnb_log(self.ctx, _ptr(cells), cells.shape[0], max_depth,
                                                                _ptr(log), cap, ctypes.byref(n_log), _ptr(o), c,
                                                                ctypes.byref(n), ctypes.byref(st)), 1 << 16, 14)
            if n_log.value <= cap or log_cap is not None:
                return out, st, log[: min(n_log.value, cap)]
            cap = n_log.value

    def rsb_ranker_train(self, records: np.ndarray, epochs: int = 20, batch: int = 4096, lr: float = 1e-3,
                         seed: int = 1):
        """cc_rsb_ranker_train on BNB_RECORD_DTYPE records.  Returns (RsbRanker, epoch losses)."""
        records = np.ascontiguousarray(records, dtype=BNB_RECORD_DTYPE)
        model = RsbRanker()
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        self._check(lib.cc_rsb_ranker_train(self.ctx, _ptr(records), records.shape[0], epochs, batch, lr, seed,
                                            ctypes.byref(model), _ptr(losses)))
        return model, losses[:epochs]

    def rsb_bnb_learned(self, cells: np.ndarray, max_depth: int, ranker: "RsbRanker", params=None,
                        cap: int = 1 << 16):
        """Batched branch-and-bound with the learned step order.  Returns (unproven, BnbStats)."""
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
        st = BnbStats()
        pp = ctypes.byref(FastParams(*params)) if params is not None else None
        out = self._grow(lambda o, c, n: lib.cc_rsb_bnb_learned(self.ctx, _ptr(cells), cells.shape[0], max_depth,
                                                                ctypes.byref(ranker), pp, _ptr(o), c, ctypes.byref(n),
                                                                ctypes.byref(st)), cap, 14)
        return out, st

    def ranker_train(self, soa: np.ndarray, label: np.ndarray, epochs: int = 20, batch: int = 4096,
                     lr: float = 1e-3, seed: int = 1):
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        label = np.ascontiguousarray(label, dtype=np.uint8)
        model = Ranker()
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        self._check(lib.cc_ranker_train(self.ctx, _ptr(soa), soa.shape[1], soa.shape[1], _ptr(label), epochs, batch,
                                        lr, seed, ctypes.byref(model), _ptr(losses)))
        return model, losses[:epochs]

    def ranker_gradient(self, soa: np.ndarray, label: np.ndarray, sample: np.ndarray, model: "Ranker"):
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        label = np.ascontiguousarray(label, dtype=np.uint8)
        sample = np.ascontiguousarray(sample, dtype=np.uint32)
        grad = np.empty(RANKER_PARAMS, dtype=np.float32)
        loss = np.empty(sample.shape[0], dtype=np.float32)
        self._check(lib.cc_ranker_gradient(self.ctx, _ptr(soa), soa.shape[1], soa.shape[1], _ptr(label), _ptr(sample),
                                           sample.shape[0], ctypes.byref(model), _ptr(grad), _ptr(loss)))
        return grad, loss

    def ranker_rank(self, soa: np.ndarray, model: "Ranker", with_logits: bool = False):
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        n = soa.shape[1]
        top1 = np.empty(n, dtype=np.uint8)
        lg = np.empty((n, RANKER_CLASSES), dtype=np.float32) if with_logits else None
        self._check(lib.cc_ranker_rank(self.ctx, _ptr(soa), n, n, ctypes.byref(model), _ptr(top1),
                                       _ptr(lg) if with_logits else None))
        return (top1, lg) if with_logits else top1

    def eval_first(self, soa: np.ndarray, model: "Ranker" = None, cascade: bool = False):
        """First-success evaluation (plain order if model is None; cascade:
        criterion 0 on every box first, the rest on the boxes it leaves
        open): returns (first, evals, trap, FirstStats)."""
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        n = soa.shape[1]
        first = np.empty(n, dtype=np.uint8)
        evals = np.empty(n, dtype=np.uint8)
        trap = np.empty(n, dtype=np.uint8)
        st = FirstStats()
        self._check(lib.cc_c2_eval_first_ex(self.ctx, _ptr(soa), n, n,
                                            ctypes.byref(model) if model is not None else None, int(bool(cascade)),
                                            _ptr(first), _ptr(evals), _ptr(trap), ctypes.byref(st)))
        return first, evals, trap, st

    def eval_first_f32(self, soa: np.ndarray, with_widths: bool = False, cascade: bool = False):
        """FP32 pre-filter + FP64 plain first-success (cascade: in the
        cascade form, same outputs): returns (first, evals, trap,
        f32_decided, widths or None, FirstStats)."""
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        n = soa.shape[1]
        first, evals, trap, dec = (np.empty(n, dtype=np.uint8) for _ in range(4))
        w = np.empty((n, 4), dtype=np.float32) if with_widths else None
        st = FirstStats()
        self._check(lib.cc_c2_eval_first_f32_ex(self.ctx, _ptr(soa), n, n, int(bool(cascade)), _ptr(first),
                                                _ptr(evals), _ptr(trap), _ptr(dec), _ptr(w) if with_widths else None,
                                                ctypes.byref(st)))
        return first, evals, trap, dec, w, st

    def f32_report(self, soa: np.ndarray, with_inflation: bool = True):
        """Per-criterion FP32 shadow report: (mask (n,) uint32 with bit k =
        the shadow of criterion k proves the box, inflation (n, 10) float32
        FP32/FP64 weight-budget width ratios, -1 where undefined, or None)."""
        soa = np.ascontiguousarray(soa, dtype=np.float64)
        n = soa.shape[1]
        mask = np.empty(n, dtype=np.uint32)
        infl = np.empty((n, 10), dtype=np.float32) if with_inflation else None
        self._check(lib.cc_c2_f32_report(self.ctx, _ptr(soa), n, n, _ptr(mask),
                                         _ptr(infl) if with_inflation else None))
        return mask, infl

    def rsb_eval_masks(self, cells: np.ndarray) -> np.ndarray:
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 14)
        mask = np.empty(cells.shape[0], dtype=np.uint64)
        self._check(lib.cc_rsb_eval_masks(self.ctx, _ptr(cells), cells.shape[0], _ptr(mask)))
        return mask

    def prim_rn_selftest(self, a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
        a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (a, b, c))
        out = np.empty((a.size, 3), dtype=np.float64)
        self._check(lib.cc_prim_rn_selftest(self.ctx, _ptr(a), _ptr(b), _ptr(c), a.size, _ptr(out)))
        return out

    def prim_selftest(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.empty((a.size, 8), dtype=np.float64)
        self._check(lib.cc_prim_selftest(self.ctx, _ptr(a), _ptr(b), a.size, _ptr(out)))
        return out
