"""Reading a prover report back: the interval text of the reference's
writer and reader (Prover/algcuda/cuda-utils/cpp/include/algcuda/interval.hpp:
294-351, through the C ABI's cc_format_interval / cc_parse_interval) and the
output_case layout (Prover/output_case.hpp:35-54), so that two reports, for
instance this library's and an archived run of the reference, can be compared
interval by interval, bit for bit."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from .native import CC_OK, CCError, lib

WIDTH = 120  # output_case.hpp: detail::output_width


class _Iv(ctypes.Structure):
    _fields_ = [("lb", ctypes.c_double), ("ub", ctypes.c_double)]


def format_interval(lb: float, ub: float) -> str:
    """The reference's operator<< text of [lb, ub]."""
    x = _Iv(float(lb), float(ub))
    buf = ctypes.create_string_buffer(128)
    n = ctypes.c_size_t()
    rc = lib.cc_format_interval(ctypes.byref(x), buf, len(buf), ctypes.byref(n))
    if rc != CC_OK:
        raise CCError(rc, "cc_format_interval")
    return buf.value.decode()


def parse_interval(text: str, pos: int = 0):
    """Reads one interval from text[pos:] as the reference's operator>> does:
    returns (lb, ub, end), end = the position after it; ValueError where the
    reference's stream would fail."""
    raw = text[pos:].encode()
    x = _Iv()
    used = ctypes.c_size_t()
    if lib.cc_parse_interval(raw, len(raw), ctypes.byref(x), ctypes.byref(used)) != CC_OK:
        raise ValueError(f"no interval at {text[pos:pos + 40]!r}")
    return x.lb, x.ub, pos + len(raw[:used.value].decode())


@dataclass
class Line:
    template: str  # the line with every interval replaced by "[]"
    intervals: np.ndarray  # (k, 2) float64, in order of appearance


@dataclass
class Case:
    title: str
    lines: list = field(default_factory=list)


def _banner_title(line: str):
    """The message of an output_begin_case banner, or None."""
    if len(line) != WIDTH or not line.startswith("-") or not line.endswith("-"):
        return None
    body = line.strip("-")
    if len(body) < 2 or body[0] != " " or body[-1] != " ":
        return None
    return body[1:-1]


def read_line(line: str) -> Line:
    parts, ivs, pos = [], [], 0
    while True:
        k = line.find("[", pos)
        if k < 0:
            parts.append(line[pos:])
            break
        try:
            lb, ub, end = parse_interval(line, k)
        except ValueError:
            parts.append(line[pos:k + 1])
            pos = k + 1
            continue
        parts.append(line[pos:k] + "[]")
        ivs.append((lb, ub))
        pos = end
    return Line("".join(parts), np.array(ivs, dtype=np.float64).reshape(-1, 2))


def read_report(text: str) -> list:
    """The cases of a report: banner title and the lines up to the closing
    rule, each split into its text and its intervals.  Lines outside a case
    are ignored."""
    cases, cur = [], None
    for line in text.splitlines():
        title = _banner_title(line)
        if title is not None:
            cur = Case(title)
            cases.append(cur)
        elif line == "-" * WIDTH:
            cur = None
        elif cur is not None:
            cur.lines.append(read_line(line))
    return cases


def compare_reports(a: str, b: str) -> list:
    """Differences between two reports: titles, line texts, and intervals
    compared bit for bit; an empty list when they agree."""
    ca, cb = read_report(a), read_report(b)
    out = []
    if len(ca) != len(cb):
        out.append(f"{len(ca)} cases against {len(cb)}")
    for i, (x, y) in enumerate(zip(ca, cb)):
        if x.title != y.title:
            out.append(f"case {i}: title {x.title!r} against {y.title!r}")
        if len(x.lines) != len(y.lines):
            out.append(f"case {i}: {len(x.lines)} lines against {len(y.lines)}")
        for j, (p, q) in enumerate(zip(x.lines, y.lines)):
            if p.template != q.template:
                out.append(f"case {i} line {j}: text differs")
            elif p.intervals.shape != q.intervals.shape or \
                    (p.intervals.view(np.uint64) != q.intervals.view(np.uint64)).any():
                out.append(f"case {i} line {j}: intervals differ")
    return out
