"""Interval text and report reading (host only, no GPU): the C ABI's writer
is the reference's operator<< layout (checked against the Python
restatement of iostream formatting), its reader accepts exactly what the
reference's operator>> accepts, and writer -> reader is the identity on the
bits of every finite double; a report built from cells reads back to the
same cells."""
import struct

import numpy as np
import pytest

from circlecover_amd.report import compare_reports, format_interval, parse_interval, read_report
from report_format import iv, rsb_report


def random_doubles(n, seed=0):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 1 << 63, n, dtype=np.uint64) | (rng.integers(0, 2, n, dtype=np.uint64) << np.uint64(63))
    x = bits.view(np.float64)
    x = x[np.isfinite(x)]
    special = np.array([0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
                        -1.7976931348623157e308, 0.1, 1 / 3, 9 / 64, 0.61, 1.0, 2.5, 1e-5, 123456789012345678.0])
    return np.concatenate([special, x])


def test_format_is_the_reference_layout():
    x = random_doubles(4000, 1)
    for a, b in zip(x[0::2], x[1::2]):
        assert format_interval(a, b) == iv(a, b)


def test_write_then_read_is_bit_exact():
    x = random_doubles(40000, 2)
    for a, b in zip(x[0::2], x[1::2]):
        t = format_interval(a, b)
        lb, ub, end = parse_interval(t)
        assert struct.pack("<dd", lb, ub) == struct.pack("<dd", a, b)
        assert end == len(t) == 56


@pytest.mark.parametrize("text,expect", [
    ("[1, 2]", (1.0, 2.0)),
    ("[1; 2]", (1.0, 2.0)),
    ("  \t[ -1.5e-3 ,+2.]", (-1.5e-3, 2.0)),
    ("[.5,5.]", (0.5, 5.0)),
    ("3.25", (3.25, 3.25)),
    ("  -0", (-0.0, -0.0)),
    ("[1e2,1E+2]tail", (100.0, 100.0)),
    ("[1e-400, 4.940656458412465442e-324]", (0.0, 5e-324)),  # underflow reads as the rounded value
    ("0x10", (0.0, 0.0)),  # a scalar 0, the x left unread
])
def test_reader_accepts(text, expect):
    lb, ub, _ = parse_interval(text)
    assert struct.pack("<dd", lb, ub) == struct.pack("<dd", *expect)


@pytest.mark.parametrize("text", ["", "   ", "[1: 2]", "[1, 2)", "[1 2]", "[, 2]", "[1, ]", "[inf, 1]", "nan",
                                  "[0x1p3, 9]", "[1e, 2]", "[-, 2]", "[1, 2", "abc",
                                  "[1e400, 2]", "-1e400"])
def test_reader_rejects(text):
    with pytest.raises(ValueError):
        parse_interval(text)


def test_report_reads_back_to_its_cells():
    rng = np.random.default_rng(3)
    lo = rng.uniform(0.0, 2.0, (5, 7))
    cells = np.stack([lo, lo + rng.uniform(0, 1e-3, lo.shape)], axis=-1).reshape(5, 14)
    text = rsb_report(cells)
    cases = read_report(text)
    assert [c.title for c in cases] == ["Critical intervals for rectangles with small disks"]
    lines = [l for l in cases[0].lines if l.template.startswith("Critical: ")]
    assert len(lines) == 5
    for l, c in zip(lines, cells):
        assert l.template == "Critical: λ = [], r²_1 = [], r²_2 = [], r²_3 = [], r²_4 = [], r²_5 = [], r²_6 = []"
        assert (l.intervals.reshape(-1).view(np.uint64) == c.view(np.uint64)).all()
    assert compare_reports(text, text) == []
    other = rsb_report(np.concatenate([cells[:4], np.nextafter(cells[4:], 9.0)]))
    assert compare_reports(text, other) == ["case 0 line 4: intervals differ"]
