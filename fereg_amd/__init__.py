"""fereg_amd: MI355X fixed-effect demeaning and estimation behind fereg-core's Absorber API.

The compute path is the HIP library libfereg_hip.so (fereg_amd/csrc, C ABI in
include/fereg_hip.h). This module is the host-side mirror of the reference's interface:
`Absorber` follows the `Absorber` trait (crates/fereg-core/src/absorber.rs:62-97),
`fit` / `feols` / `fepois` follow the fit entry points (crates/fereg-core/src/fit/mod.rs:591,
python/fereg/_binding.py:129), `reghdfe` the Session's reghdfe convention (session.rs, ols.rs),
and `Sample` the estimation sample (fit/input.rs, factor.rs, sample.rs). There is no CPU fallback: without the built library every call
raises.
"""
from ._lib import (FH_CONVERGENCE, FH_INTERNAL, FH_INTERRUPTED, FH_INVALID, FH_MIXED, FH_OK, FH_OOM,
                   FH_PLAIN_FP64, FH_TOO_FEW, FeregError, lib, library_path, memory_stats,
                   release_cached_memory)
from .absorber import Absorber, FeCoefs, SolveInfo, SolveOpts
from .fit import FitResult, WildBootResult, feols, fepois, fit, wildboot
from .reghdfe import ReghdfeResult, reghdfe
from .sample import Codes, Sample

__all__ = ["Absorber", "FeCoefs", "SolveInfo", "SolveOpts", "FitResult", "WildBootResult", "feols", "fepois", "fit",
           "wildboot", "reghdfe", "ReghdfeResult", "Sample", "Codes", "FeregError", "lib",
           "library_path", "FH_OK", "FH_INTERRUPTED", "FH_INVALID", "FH_CONVERGENCE", "FH_OOM", "FH_TOO_FEW",
           "FH_INTERNAL", "FH_PLAIN_FP64", "FH_MIXED", "memory_stats", "release_cached_memory"]
