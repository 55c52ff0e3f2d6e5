"""MI355X-native Semiconductor Bloch Equation propagator (Antelope's SBE loop).

The compute path is the HIP library built from antelope_amd/csrc (C ABI in
include/antelope/sbe_c.h); this package holds its Python binding, the
run/model parameter mirror and the build script.
"""
from .params import AU_ANGSTROM, AU_EV, AnalyticModel, HaldaneModel, Pulse, RunConfig, WannierModel, tmdc  # noqa: F401
from .solver import SBEBatch, SBEError, SBESolver, laser_eval, library_path, load_library, query_footprint  # noqa: F401
