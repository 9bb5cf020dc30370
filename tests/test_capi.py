"""The C-ABI library loads on the CPU host and exports every function that
include/circlecover_amd.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "circlecover_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(cc_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    fns = declared_functions()
    for need in ("cc_init", "cc_destroy", "cc_eval_batch", "cc_eval_batch_async", "cc_gen_c2_boxes"):
        assert need in fns


def test_library_exports_every_declared_symbol():
    import circlecover_amd

    out = subprocess.run(["nm", "-D", "--defined-only", circlecover_amd.library_path], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\b[TW]\s+(cc_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, f"not exported: {missing}"
    for f in declared_functions():
        assert hasattr(circlecover_amd.lib, f)


def test_record_layouts_match_reference_sizes():
    # cc_tr_cell / cc_rsb_cell mirror the reference Variables structs (80 / 112 bytes)
    class IV(ctypes.Structure):
        _fields_ = [("lb", ctypes.c_double), ("ub", ctypes.c_double)]

    class TR(ctypes.Structure):
        _fields_ = [("la", IV), ("r", IV * 4)]

    class RSB(ctypes.Structure):
        _fields_ = [("la", IV), ("r", IV * 6)]

    assert ctypes.sizeof(TR) == 80 and ctypes.sizeof(RSB) == 112


def test_box_generator_is_deterministic_and_in_range():
    from circlecover_amd import gen_c2_boxes

    a = gen_c2_boxes(0xC1C1EC0000000002, 4096)
    b = gen_c2_boxes(0xC1C1EC0000000002, 4096)
    assert a.shape == (8, 4096) and (a.view(np.uint64) == b.view(np.uint64)).all()
    la_lb, r1, r2, r3 = a[0], a[2], a[4], a[6]
    assert ((la_lb >= 1.0) & (la_lb < 2.5)).all()
    assert ((r1 >= r2) & (r2 >= r3) & (r3 >= 0.02) & (r1 < 0.7)).all()
    for d in range(4):
        w = a[2 * d + 1] - a[2 * d]
        assert (w > 2.0 ** -21).all() and (w <= 2.0 ** -6 * 1.0000001).all()
    c = gen_c2_boxes(0xC1C1EC0000000003, 4096)
    assert not (a == c).all()
