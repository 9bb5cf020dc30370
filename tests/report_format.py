"""Python restatement of the reference's report formatting (test support):
output_case banners (Prover/output_case.hpp:35-54), the Interval stream
operator (algcuda/interval.hpp:294-303: right-aligned, width 26, precision
19) and the Variables operators (tight_rectangle.cu:539-545,
rectangle_size_bound/variables.hpp:112-118)."""
import ctypes

import numpy as np

import oracle

WIDTH = 120


def begin_case(msg):
    pre = (WIDTH - len(msg) - 2) // 2
    post = WIDTH - len(msg) - 2 - pre
    return "-" * pre + " " + msg + " " + "-" * post + "\n"


def end_case():
    return "-" * WIDTH + "\n"


def g6(x):  # std::ostream << double with the default precision 6
    return format(float(x), "g")


def iv(lo, hi):
    return "[" + format(float(lo), ">26.19g") + ", " + format(float(hi), ">26.19g") + "]"


def cell(v, nr):
    s = "λ = " + iv(v[0], v[1])
    for i in range(nr):
        s += f", r²_{i + 1} = " + iv(v[2 + 2 * i], v[3 + 2 * i])
    return s


class _Cell(ctypes.Structure):
    _fields_ = [("v", ctypes.c_double * 10)]


class MaxDist(ctypes.Structure):
    _fields_ = [("have_2d", ctypes.c_int), ("have_3d", ctypes.c_int), ("la_2d", ctypes.c_double * 2),
                ("la_3d", ctypes.c_double * 2), ("d_r1_2d", ctypes.c_double * 2), ("d_r1_2d_crit", _Cell * 2),
                ("d_r2_2d", ctypes.c_double * 2), ("d_r2_2d_crit", _Cell * 2), ("lb_h1_2d", ctypes.c_double),
                ("ub_la_by_S1", ctypes.c_double), ("ub_R3_2d", ctypes.c_double), ("d_r1_3d", ctypes.c_double * 2),
                ("d_r1_3d_crit", _Cell * 2), ("d_r2_3d", ctypes.c_double * 2), ("d_r2_3d_crit", _Cell * 2),
                ("d_r3_3d", ctypes.c_double * 2), ("d_r3_3d_crit", _Cell * 2), ("ub_R4_3d", ctypes.c_double),
                ("S1", ctypes.c_double * 2), ("h2", ctypes.c_double * 2), ("S3", ctypes.c_double * 2),
                ("pocket_3d", ctypes.c_double), ("trap", ctypes.c_int)]


def max_distances(cells):
    lib = oracle.lib
    lib.or_tr_max_distances_size.restype = ctypes.c_size_t
    assert lib.or_tr_max_distances_size() == ctypes.sizeof(MaxDist)
    lib.or_tr_max_distances.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.or_tr_max_distances.restype = None
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 10)
    r = MaxDist()
    lib.or_tr_max_distances(cells.ctypes.data, cells.shape[0], ctypes.byref(r))
    return r


def tr_report(criticals):
    """The stdout of tight_rectangle::find_critical_intervals for the final
    critical records (tight_rectangle.cu:459 and :498-536)."""
    out = begin_case("Critical intervals for rectangles (tight)")
    if len(criticals):
        r = max_distances(criticals)

        def crit(c):
            return "\t\t" + cell(list(c.v), 4) + "\n"

        if r.have_2d:
            out += ("Two disks: Range for lambda that contains 2-disk critical intervals: " + iv(*r.la_2d)
                    + "; maximal distances from theoretical values: \n")
            out += ("\tr_1^2: " + iv(*r.d_r1_2d) + ", r_2^2: " + iv(*r.d_r2_2d) + ", R_3 <= " + g6(r.ub_R3_2d)
                    + ", h_1 >= " + g6(r.lb_h1_2d) + ", lambda / S_1 <= " + g6(r.ub_la_by_S1) + "\n")
            out += "\tDominating criticals for r_1: \n" + crit(r.d_r1_2d_crit[0]) + crit(r.d_r1_2d_crit[1])
            out += "\tDominating criticals for r_2: \n" + crit(r.d_r2_2d_crit[0]) + crit(r.d_r2_2d_crit[1])
        if r.have_3d:
            out += ("Three disks: Range for lambda that contains 3-disk critical intervals: " + iv(*r.la_3d)
                    + "; maximal distances from theoretical values: \n")
            out += ("\tr_1^2: " + iv(*r.d_r1_3d) + ", r_2^2: " + iv(*r.d_r2_3d) + ", r_3^2: " + iv(*r.d_r3_3d)
                    + ", R_4 <= " + g6(r.ub_R4_3d) + "\n")
            out += ("\tS_1: " + iv(*r.S1) + ", h_2: " + iv(*r.h2) + ", S_3: " + iv(*r.S3) + ", pocket width <= "
                    + g6(r.pocket_3d) + "\n")
            out += "\tDominating criticals for r_1: \n" + crit(r.d_r1_3d_crit[0]) + crit(r.d_r1_3d_crit[1])
            out += "\tDominating criticals for r_2: \n" + crit(r.d_r2_3d_crit[0]) + crit(r.d_r2_3d_crit[1])
            out += "\tDominating criticals for r_3: \n" + crit(r.d_r3_3d_crit[0]) + crit(r.d_r3_3d_crit[1])
    return out + end_case()


def merge_critical(records, n_merged):
    """merge_critical_intervals<N> (rectangle_size_bound.cu:103-130)."""
    if len(records) == 0:
        return np.empty((0, 14))
    recs = sorted((tuple(r) for r in np.asarray(records)))
    res = [list(recs[0])]
    for v in recs:
        p = res[-1]
        if v[0] == p[0] and v[1] == p[1] and all(v[2 + k] == p[2 + k] for k in range(2 * n_merged)):
            for i in range(n_merged, 6):
                p[2 + 2 * i] = min(p[2 + 2 * i], v[2 + 2 * i])
                p[3 + 2 * i] = max(p[3 + 2 * i], v[3 + 2 * i])
        else:
            res.append(list(v))
    return np.array(res)


def rsb_report(criticals):
    out = begin_case("Critical intervals for rectangles with small disks")
    for c in criticals:
        out += "Critical: " + cell(c, 6) + "\n"
    return out + end_case()
