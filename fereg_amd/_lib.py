"""ctypes declarations of the C ABI (include/fereg_hip.h)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(HERE, "libfereg_hip.so")

FH_OK = 0
FH_INTERRUPTED = 1
FH_INVALID = 198
FH_CONVERGENCE = 430
FH_OOM = 909
FH_TOO_FEW = 2001
FH_INTERNAL = 9999
FH_PLAIN_FP64 = 0
FH_MIXED = 1
FH_VCOV = {"iid": 0, "hetero": 1, "cluster": 2, "hc2": 3, "hc3": 4, "crv3": 5}
FH_KFIXEF = {"none": 0, "full": 1, "nonnested": 2}

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
i32p = ctypes.POINTER(ctypes.c_int32)
i64p = ctypes.POINTER(ctypes.c_int64)
f64p = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p


class FeregError(RuntimeError):
    """An error code of the reference's convention (error.rs:6-37) with its message."""

    def __init__(self, rc, msg):
        super().__init__(f"[{rc}] {msg}")
        self.rc = rc


class Term(ctypes.Structure):
    _fields_ = [("codes", u32p), ("g", ctypes.c_uint64), ("intercept", ctypes.c_int32), ("nslopes", ctypes.c_int32),
                ("slopes", ctypes.POINTER(f64p))]


class SolverConfig(ctypes.Structure):
    _fields_ = [("small_limit", ctypes.c_uint64), ("precision", ctypes.c_int32), ("bucket_log2", ctypes.c_int32),
                ("autotune", ctypes.c_int32), ("mixed_lead", ctypes.c_int32), ("sparse_chol", ctypes.c_int32),
                ("joint", ctypes.c_int32), ("dense_limit", ctypes.c_int32)]


class SpmvPlan(ctypes.Structure):
    _fields_ = [("dense", ctypes.c_int32), ("seg_lanes", ctypes.c_uint32), ("block_lanes", ctypes.c_uint32 * 9),
                ("block_mode", ctypes.c_int32 * 9), ("tuned", ctypes.c_int32)]


class SolveOptsC(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("maxiter", ctypes.c_uint64), ("want_coefs", ctypes.c_int32),
                ("precision", ctypes.c_int32)]


class SolveInfoC(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("escalated", ctypes.c_int32), ("exact", ctypes.c_int32),
                ("refinements", ctypes.c_uint64), ("seconds", ctypes.c_double)]


class Ctx(ctypes.Structure):
    _fields_ = [("interrupt", ctypes.CFUNCTYPE(ctypes.c_int, vp)), ("log", ctypes.CFUNCTYPE(None, vp, ctypes.c_char_p)),
                ("user", vp), ("verbose", ctypes.c_int32)]


class CallContext:
    """An fh_ctx (api.rs:198-229) from Python callables: interrupt() -> bool is polled between
    kernels and iterations (True stops the call with FH_INTERRUPTED); log(line) receives the
    library's progress lines at verbosity >= their level. Keeps the ctypes callbacks alive."""

    def __init__(self, interrupt=None, log=None, verbose=0):
        IFN = Ctx._fields_[0][1]
        LFN = Ctx._fields_[1][1]
        self._i = IFN((lambda _u: 1 if interrupt() else 0) if interrupt else 0)
        self._l = LFN((lambda _u, s: log(s.decode())) if log else 0)
        self.c = Ctx(self._i, self._l, None, int(verbose))

    def ref(self):
        return ctypes.byref(self.c)


def ctx_ref(ctx):
    return None if ctx is None else ctx.ref()


class StructureInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("nterms", ctypes.c_uint64), ("g1", ctypes.c_uint64),
                ("ncells", ctypes.c_uint64), ("nb", ctypes.c_uint64), ("m", ctypes.c_uint64),
                ("m_small", ctypes.c_uint64), ("small_work", ctypes.c_uint64), ("fe1", ctypes.c_int64),
                ("norm_fe1", ctypes.c_int64), ("has_cell_obs_off", ctypes.c_int32), ("dense_small", ctypes.c_int32)]


class Ssc(ctypes.Structure):
    _fields_ = [("k_adj", ctypes.c_int32), ("k_fixef", ctypes.c_int32), ("g_adj", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


FH_GDF = {"min": 0, "conventional": 1}


class Clusters(ctypes.Structure):
    _fields_ = [("codes", ctypes.POINTER(vp)), ("ndim", ctypes.c_size_t), ("g_df", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class FitStats(ctypes.Structure):
    _fields_ = [("ssr", ctypes.c_double), ("df_k", ctypes.c_double), ("nparams", ctypes.c_double),
                ("deviance", ctypes.c_double), ("iterations", ctypes.c_uint64), ("pcg_iterations", ctypes.c_uint64),
                ("n_clusters", ctypes.c_uint64), ("converged", ctypes.c_int32), ("reserved", ctypes.c_int32)]


FH_PVAL = {"two-tailed": 0, "equal-tailed": 1, "greater": 2, "less": 3}


class WbootSpec(ctypes.Structure):
    _fields_ = [("reps", ctypes.c_uint64), ("seed", ctypes.c_uint64), ("r_mat", ctypes.POINTER(ctypes.c_double)),
                ("r", ctypes.c_double), ("level", ctypes.c_double), ("tol", ctypes.c_double),
                ("maxiter", ctypes.c_uint64), ("pval_type", ctypes.c_int32), ("conf_int", ctypes.c_int32),
                ("boot_type", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class WbootResult(ctypes.Structure):
    _fields_ = [("stat", ctypes.c_double), ("p_val", ctypes.c_double), ("se", ctypes.c_double),
                ("point", ctypes.c_double), ("ci_lo", ctypes.c_double), ("ci_hi", ctypes.c_double),
                ("ssc", ctypes.c_double), ("reps", ctypes.c_uint64), ("reps_feasible", ctypes.c_uint64),
                ("n_clusters", ctypes.c_uint64), ("full_enumeration", ctypes.c_int32),
                ("fe_not_projected", ctypes.c_int32), ("seconds_bootstrap", ctypes.c_double),
                ("contraction_ms", ctypes.c_double), ("contraction_ops", ctypes.c_double)]


FH_MODEL = {"ols": 0, "ppml": 1}


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64), ("precision", ctypes.c_int32),
                ("pack", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int32), ("vcov_kind", ctypes.c_int32), ("n", ctypes.c_size_t),
                ("K", ctypes.c_size_t), ("nterms", ctypes.c_size_t), ("terms", ctypes.POINTER(Term)),
                ("y", ctypes.POINTER(ctypes.c_double)), ("X", ctypes.POINTER(ctypes.c_double)),
                ("cluster", ctypes.POINTER(ctypes.c_uint32)), ("nested", ctypes.POINTER(ctypes.c_int32)),
                ("ssc", Ssc), ("tol", ctypes.c_double), ("maxiter", ctypes.c_uint64),
                ("beta", ctypes.POINTER(ctypes.c_double)), ("vcov", ctypes.POINTER(ctypes.c_double)),
                ("kept", ctypes.POINTER(ctypes.c_int32))]


class Response(ctypes.Structure):
    _fields_ = [("rc", ctypes.c_int32), ("packed", ctypes.c_int32), ("queued_s", ctypes.c_double),
                ("service_s", ctypes.c_double), ("total_s", ctypes.c_double), ("iterations", ctypes.c_uint64),
                ("pcg_iterations", ctypes.c_uint64), ("start_s", ctypes.c_double), ("end_s", ctypes.c_double)]


FH_COL = {"f64": 0, "i32": 1, "i64": 2, "codes": 3}
FH_FAMILY = {"gaussian": 0, "poisson": 1, "logit": 2}
FH_WEIGHT = {None: 0, "aweight": 1, "fweight": 2, "pweight": 3}


class FactorCol(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32), ("data", vp), ("base", ctypes.c_int64),
                ("nlevels", ctypes.c_uint64)]


class SampleSpec(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("y", vp), ("x", ctypes.POINTER(vp)), ("nx", ctypes.c_size_t),
                ("weights", vp), ("offset", vp), ("drop", vp), ("factors", ctypes.POINTER(FactorCol)),
                ("nfactors", ctypes.c_size_t), ("term_factors", u32p), ("term_off", u32p), ("nterms", ctypes.c_size_t),
                ("cluster_factors", u32p), ("cluster_off", u32p), ("nclusters", ctypes.c_size_t),
                ("weight_kind", ctypes.c_int32), ("family", ctypes.c_int32), ("missing", ctypes.c_int32),
                ("drop_singletons", ctypes.c_int32), ("separation", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class SampleCounts(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("n_input", "nobs", "na", "zero_weight", "singletons", "perfect_fit",
                                               "na_y", "na_x", "na_fe", "na_cluster", "na_weights", "na_offset",
                                               "na_extra")]


FH_VCE = {"unadjusted": 0, "robust": 1, "cluster": 2}


class ReghdfeSpec(ctypes.Structure):
    _fields_ = [("vce", ctypes.c_int32), ("weight_kind", ctypes.c_int32), ("terms", ctypes.POINTER(Term)),
                ("nterms", ctypes.c_size_t), ("clusters", ctypes.POINTER(vp)), ("nclusters", ctypes.c_size_t),
                ("noconstant", ctypes.c_int32), ("dof_firstpair", ctypes.c_int32), ("dof_pairwise", ctypes.c_int32),
                ("dof_clusters", ctypes.c_int32), ("dof_continuous", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("tol", ctypes.c_double), ("maxiter", ctypes.c_uint64), ("num_singletons", ctypes.c_uint64)]


REGHDFE_SCALARS = ["N", "rank", "df_r", "df_m", "df_a", "df_a_initial", "df_a_redundant", "df_a_nested", "N_clust",
                   "N_clustervars"]
REGHDFE_FIT = ["tss", "tss_within", "rss", "mss", "rmse", "F", "ll", "ll_0", "r2", "r2_within", "r2_a",
               "r2_a_within", "sumweights", "num_singletons", "N_full", "report_constant", "converged", "y_absorbed"]


class ReghdfeStats(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_double) for k in REGHDFE_SCALARS] + [("N_clust_dim", ctypes.c_double * 4)]
                + [(k, ctypes.c_double) for k in REGHDFE_FIT]
                + [("iterations", ctypes.c_uint64), ("nslots", ctypes.c_uint64), ("psd_fixed", ctypes.c_int32),
                   ("reserved", ctypes.c_int32)])


# (name, restype, argtypes) of every exported entry point of include/fereg_hip.h
SIGNATURES = [
    ("fh_last_error", ctypes.c_char_p, []),
    ("fh_config_default", None, [ctypes.POINTER(SolverConfig)]),
    ("fh_absorber_get_plan", ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_size_t, ctypes.POINTER(SpmvPlan)]),
    ("fh_absorber_set_plan", ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_size_t, ctypes.POINTER(SpmvPlan)]),
    ("fh_tune_cache_clear", None, []),
    ("fh_absorber_create", ctypes.c_int, [ctypes.POINTER(Term), ctypes.c_size_t, f64p, ctypes.c_size_t,
                                          ctypes.POINTER(SolverConfig), ctypes.POINTER(Ctx), ctypes.POINTER(vp)]),
    ("fh_absorber_create_dev", ctypes.c_int, [ctypes.POINTER(Term), ctypes.c_size_t, vp, ctypes.c_size_t,
                                              ctypes.POINTER(SolverConfig), ctypes.POINTER(Ctx), ctypes.POINTER(vp)]),
    ("fh_absorber_destroy", None, [vp]),
    ("fh_absorber_n", ctypes.c_size_t, [vp]),
    ("fh_absorber_row_order", u32p, [vp]),
    ("fh_absorber_row_order_dev", vp, [vp]),
    ("fh_absorber_set_weights", ctypes.c_int, [vp, f64p, ctypes.POINTER(Ctx)]),
    ("fh_absorber_set_weights_dev", ctypes.c_int, [vp, vp, ctypes.POINTER(Ctx)]),
    ("fh_absorber_residualize", ctypes.c_int, [vp, f64p, ctypes.c_size_t, ctypes.POINTER(SolveOptsC), f64p, f64p,
                                               ctypes.POINTER(SolveInfoC), ctypes.POINTER(Ctx)]),
    ("fh_absorber_residualize_dev", ctypes.c_int, [vp, vp, ctypes.c_size_t, ctypes.POINTER(SolveOptsC), vp, vp,
                                                   ctypes.POINTER(SolveInfoC), ctypes.POINTER(Ctx)]),
    ("fh_absorber_coef_rows", ctypes.c_size_t, [vp]),
    ("fh_absorber_term_slots", ctypes.c_size_t, [vp]),
    ("fh_absorber_fit_rows", ctypes.c_int, [vp, f64p, ctypes.c_size_t, ctypes.c_size_t, f64p]),
    ("fh_absorber_term_values", ctypes.c_int, [vp, f64p, ctypes.c_size_t, ctypes.c_size_t, f64p, u64p]),
    ("fh_absorber_info", ctypes.c_int, [vp, ctypes.POINTER(StructureInfo)]),
    ("fh_absorber_copy_structure", ctypes.c_int, [vp, u32p, u32p, u32p, u32p]),
    ("fh_absorber_copy_blocks", ctypes.c_int, [vp, i64p, u64p, u64p, i32p]),
    ("fh_absorber_copy_level_cells", ctypes.c_int, [vp, ctypes.c_size_t, u32p, u32p]),
    ("fh_absorber_copy_weight_state", ctypes.c_int, [vp, f64p, f64p, f64p, ctypes.POINTER(ctypes.c_uint8), f64p]),
    ("fh_absorber_apply", ctypes.c_int, [vp, f64p, ctypes.c_size_t, f64p, ctypes.c_int32]),
    ("fh_absorber_apply_noalloc", ctypes.c_int, [vp, vp, ctypes.c_size_t, vp, ctypes.c_int32]),
    ("fh_noalloc_scratch_bytes", ctypes.c_size_t, []),
    ("fh_time_gather", ctypes.c_int, [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int32, ctypes.c_uint32, f64p]),
    ("fh_absorber_time_spmv", ctypes.c_int, [vp, ctypes.c_size_t, ctypes.c_int32, ctypes.c_uint32, f64p, f64p]),
    ("fh_ols_fit", ctypes.c_int, [vp, f64p, f64p, ctypes.c_size_t, ctypes.c_int32, u32p, i32p, ctypes.POINTER(Ssc),
                                  ctypes.POINTER(SolveOptsC), f64p, f64p, i32p, ctypes.POINTER(FitStats)]),
    ("fh_ols_fit_dev", ctypes.c_int, [vp, vp, vp, ctypes.c_size_t, ctypes.c_int32, vp, i32p, ctypes.POINTER(Ssc),
                                      ctypes.POINTER(SolveOptsC), f64p, f64p, i32p, ctypes.POINTER(FitStats)]),
    ("fh_ppml_fit", ctypes.c_int, [vp, f64p, f64p, ctypes.c_size_t, f64p, ctypes.c_int32, u32p, i32p,
                                   ctypes.POINTER(Ssc), ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(SolveOptsC),
                                   f64p, f64p, i32p, ctypes.POINTER(FitStats), ctypes.POINTER(Ctx)]),
    ("fh_ppml_fit_dev", ctypes.c_int, [vp, vp, vp, ctypes.c_size_t, vp, ctypes.c_int32, vp, i32p, ctypes.POINTER(Ssc),
                                       ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(SolveOptsC), f64p, f64p, i32p,
                                       ctypes.POINTER(FitStats), ctypes.POINTER(Ctx)]),
    ("fh_ols_fit_multi", ctypes.c_int, [vp, f64p, f64p, ctypes.c_size_t, ctypes.c_int32, ctypes.POINTER(Clusters), i32p,
                                        ctypes.POINTER(Ssc), ctypes.POINTER(SolveOptsC), f64p, f64p, i32p,
                                        ctypes.POINTER(FitStats)]),
    ("fh_ols_fit_multi_dev", ctypes.c_int, [vp, vp, vp, ctypes.c_size_t, ctypes.c_int32, ctypes.POINTER(Clusters),
                                            i32p, ctypes.POINTER(Ssc), ctypes.POINTER(SolveOptsC), f64p, f64p, i32p,
                                            ctypes.POINTER(FitStats)]),
    ("fh_ppml_fit_multi", ctypes.c_int, [vp, f64p, f64p, ctypes.c_size_t, f64p, ctypes.c_int32,
                                         ctypes.POINTER(Clusters), i32p, ctypes.POINTER(Ssc), ctypes.c_double,
                                         ctypes.c_uint64, ctypes.POINTER(SolveOptsC), f64p, f64p, i32p,
                                         ctypes.POINTER(FitStats), ctypes.POINTER(Ctx)]),
    ("fh_ppml_fit_multi_dev", ctypes.c_int, [vp, vp, vp, ctypes.c_size_t, vp, ctypes.c_int32,
                                             ctypes.POINTER(Clusters), i32p, ctypes.POINTER(Ssc), ctypes.c_double,
                                             ctypes.c_uint64, ctypes.POINTER(SolveOptsC), f64p, f64p, i32p,
                                             ctypes.POINTER(FitStats), ctypes.POINTER(Ctx)]),
    ("fh_ols_wildboot", ctypes.c_int, [vp, f64p, f64p, ctypes.c_size_t, u32p, i32p, ctypes.POINTER(Ssc),
                                       ctypes.POINTER(SolveOptsC), ctypes.POINTER(WbootSpec), f64p, f64p, i32p,
                                       ctypes.POINTER(WbootResult), f64p]),
    ("fh_ols_wildboot_dev", ctypes.c_int, [vp, vp, vp, ctypes.c_size_t, vp, i32p, ctypes.POINTER(Ssc),
                                           ctypes.POINTER(SolveOptsC), ctypes.POINTER(WbootSpec), f64p, f64p, i32p,
                                           ctypes.POINTER(WbootResult), f64p]),
    ("fh_factorize", ctypes.c_int, [ctypes.POINTER(FactorCol), ctypes.c_size_t, ctypes.c_int32, u32p, u64p, u64p]),
    ("fh_factorize_dev", ctypes.c_int, [ctypes.POINTER(FactorCol), ctypes.c_size_t, ctypes.c_int32, vp, u64p, u64p]),
    ("fh_combine", ctypes.c_int, [ctypes.POINTER(vp), u64p, ctypes.c_size_t, ctypes.c_size_t, u32p, u64p]),
    ("fh_combine_dev", ctypes.c_int, [ctypes.POINTER(vp), u64p, ctypes.c_size_t, ctypes.c_size_t, vp, u64p]),
    ("fh_restrict", ctypes.c_int, [u32p, ctypes.c_uint64, ctypes.c_size_t, u32p, ctypes.c_size_t, u32p, u64p]),
    ("fh_restrict_dev", ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_size_t, vp, ctypes.c_size_t, vp, u64p]),
    ("fh_drop_singletons", ctypes.c_int, [ctypes.POINTER(Term), ctypes.c_size_t, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_uint8), f64p, f64p, ctypes.c_int32, ctypes.c_int32,
                                          u64p, u64p]),
    ("fh_drop_singletons_dev", ctypes.c_int, [ctypes.POINTER(Term), ctypes.c_size_t, ctypes.c_size_t, vp, vp, vp,
                                              ctypes.c_int32, ctypes.c_int32, u64p, u64p]),
    ("fh_sample_build", ctypes.c_int, [ctypes.POINTER(SampleSpec), ctypes.POINTER(Ctx), ctypes.POINTER(vp)]),
    ("fh_sample_build_dev", ctypes.c_int, [ctypes.POINTER(SampleSpec), ctypes.POINTER(Ctx), ctypes.POINTER(vp)]),
    ("fh_sample_destroy", None, [vp]),
    ("fh_sample_get_counts", ctypes.c_int, [vp, ctypes.POINTER(SampleCounts)]),
    ("fh_sample_kept_dev", vp, [vp]),
    ("fh_sample_kept", ctypes.c_int, [vp, u32p]),
    ("fh_sample_term_dev", vp, [vp, ctypes.c_size_t, u64p]),
    ("fh_sample_term", ctypes.c_int, [vp, ctypes.c_size_t, u32p, u64p]),
    ("fh_sample_cluster_dev", vp, [vp, ctypes.c_size_t, u64p]),
    ("fh_sample_cluster", ctypes.c_int, [vp, ctypes.c_size_t, u32p, u64p]),
    ("fh_sample_gather", ctypes.c_int, [vp, f64p, f64p]),
    ("fh_sample_gather_dev", ctypes.c_int, [vp, vp, vp]),
    ("fh_reghdfe_spec_default", None, [ctypes.POINTER(ReghdfeSpec)]),
    ("fh_reghdfe_fit", ctypes.c_int, [vp, f64p, f64p, ctypes.c_size_t, ctypes.POINTER(ReghdfeSpec), f64p, f64p, i32p,
                                      f64p, ctypes.POINTER(ReghdfeStats), ctypes.POINTER(Ctx)]),
    ("fh_reghdfe_fit_dev", ctypes.c_int, [vp, vp, vp, ctypes.c_size_t, ctypes.POINTER(ReghdfeSpec), f64p, f64p, i32p,
                                          f64p, ctypes.POINTER(ReghdfeStats), ctypes.POINTER(Ctx)]),
    ("fh_server_config_default", None, [ctypes.POINTER(ServerConfig)]),
    ("fh_server_create", ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
    ("fh_server_submit", ctypes.c_int, [vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
    ("fh_server_wait", ctypes.c_int, [vp, ctypes.c_uint64, ctypes.POINTER(Response)]),
    ("fh_server_destroy", None, [vp]),
    ("fh_sparse_chol_factor", ctypes.c_int, [ctypes.c_size_t, u64p, u32p, f64p, ctypes.c_size_t, ctypes.POINTER(vp)]),
    ("fh_sparse_chol_info", ctypes.c_int, [vp, u64p, f64p, u64p, u64p]),
    ("fh_sparse_chol_solve", ctypes.c_int, [vp, f64p, ctypes.c_size_t]),
    ("fh_sparse_chol_inverse_entries", ctypes.c_int, [vp, ctypes.c_size_t, u32p, u32p, f64p]),
    ("fh_sparse_chol_destroy", None, [vp]),
    ("fh_device_alloc", ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_size_t]),
    ("fh_device_free", ctypes.c_int, [vp]),
    ("fh_memcpy_h2d", ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    ("fh_memcpy_d2h", ctypes.c_int, [vp, vp, ctypes.c_size_t]),
    ("fh_device_synchronize", ctypes.c_int, []),
    ("fh_release_cached_memory", ctypes.c_int, []),
    ("fh_memory_stats", ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
]

_lib = None


def lib():
    """The loaded HIP library; raises when it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise ImportError(f"{library_path} is missing: build it with `python -m fereg_amd.build` "
                              "(or __graft_entry__.build()); fereg_amd has no CPU fallback")
        # PyTorch ships its own HIP runtime under the same soname (libamdhip64.so.7): the first one
        # loaded serves the whole process. Load torch's first so that device tensors and this
        # library share one runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(library_path)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != FH_OK:
        msg = lib().fh_last_error()
        raise FeregError(rc, msg.decode() if msg else "")


def memory_stats():
    """(bytes in use, bytes cached) of the library's stream-ordered device pool."""
    a, b = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(lib().fh_memory_stats(ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def release_cached_memory():
    """Returns the pool's completed cached blocks to the device."""
    check(lib().fh_release_cached_memory())
