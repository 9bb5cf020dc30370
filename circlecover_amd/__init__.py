"""circlecover_amd — MI355X (gfx950) interval branch-and-bound prover.

Python host bindings over the C ABI declared in include/circlecover_amd.h
(libcirclecover_amd.so, built in-tree by circlecover_amd/build.py).  There
is no fallback: if the native library is missing or no gfx950 device is
present, the calls raise.
"""
from .native import (  # noqa: F401
    C2_CRITERIA,
    DBL_MAX,
    TR_LAMBDA_MAX,
    CCError,
    Prover,
    gen_c2_boxes,
    lib,
    library_path,
    rsb_grid,
    subintervals,
    tr_initial_cells,
)

__all__ = ["CCError", "Prover", "gen_c2_boxes", "lib", "library_path", "C2_CRITERIA", "DBL_MAX", "TR_LAMBDA_MAX",
           "subintervals", "tr_initial_cells", "rsb_grid"]
