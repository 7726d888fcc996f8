"""ctypes bindings to the in-tree native libraries.

The HIP library is required: there is no CPU fallback for the search path, so
loading it raises NativeLibraryMissing when it has not been built.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
TOOLS_SO = os.path.join(_PKG, "libpanns_tools.so")
# PANNS_AMD_SO selects an instrumented build of the same library.
AMD_SO = os.environ.get("PANNS_AMD_SO", os.path.join(_PKG, "libpanns_amd.so"))

c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_int = ctypes.c_int
c_float = ctypes.c_float
c_double = ctypes.c_double
c_char_p = ctypes.c_char_p
P = ctypes.POINTER
vp = ctypes.c_void_p


class NativeLibraryMissing(RuntimeError):
    pass


class PannsParams(ctypes.Structure):
    _fields_ = [("K", c_u32), ("L", c_u32), ("M", c_u32), ("mode", c_u32),
                ("num_threads", c_u32), ("stop", c_u32), ("stop_threshold", c_float),
                ("hash_slots", c_u32), ("waves", c_u32), ("stop_min_iter", c_u32), ("stop_every", c_u32),
                ("bucket_slots", c_u32)]


class PannsCounters(ctypes.Structure):
    _fields_ = [("distance_computations", c_u64), ("merges", c_u64), ("expansions", c_u64),
                ("adjacency_slots", c_u64), ("visited_spills", c_u64), ("kernel_ms", c_double)]


class PannsOutputs(ctypes.Structure):
    _fields_ = [("ids", vp), ("dists", vp), ("queue_ids", vp), ("queue_dists", vp),
                ("queue_checked", vp), ("ndist", vp), ("nexp", vp), ("niter", vp)]


MAX_BUCKETS = 8


class PannsPlanOptions(ctypes.Structure):
    _fields_ = [("n_buckets", c_u32), ("bucket_max_L", c_u32 * MAX_BUCKETS), ("table_eighths", c_u32),
                ("use_graph", c_u32), ("concurrent", c_u32), ("want_queue", c_u32), ("waves", c_u32),
                ("rows_per_wave", c_u32), ("overflow_slots", c_u32), ("deferred_merge", c_u32),
                ("visited_table", c_u32)]


# name -> (restype, argtypes) for include/panns_tools.h
TOOLS_API = {
    "pt_last_error": (c_char_p, []),
    "pt_l2": (c_float, [vp, vp, c_u32]),
    "pt_gen_mixture": (c_int, [c_u64, c_u64, c_u64, c_u32, c_u32, c_u32, c_float, c_int, vp, c_int]),
    "pt_mixture_labels": (c_int, [c_u64, c_u64, c_u32, c_u32, vp, c_int]),
    "pt_fvecs_info": (c_int, [c_char_p, P(c_u32), P(c_u32)]),
    "pt_fvecs_read": (c_int, [c_char_p, vp, c_u32, c_u32]),
    "pt_fvecs_write": (c_int, [c_char_p, vp, c_u32, c_u32]),
    "pt_nsg_info": (c_int, [c_char_p, P(c_u32), P(c_u32), P(c_u32), P(c_u64), P(c_u32)]),
    "pt_nsg_read": (c_int, [c_char_p, c_u32, c_u32, vp, vp]),
    "pt_nsg_write": (c_int, [c_char_p, c_u32, c_u32, c_u32, c_u32, vp, vp]),
    "pt_truth_info": (c_int, [c_char_p, P(c_u32), P(c_u32)]),
    "pt_truth_read": (c_int, [c_char_p, c_u32, c_u32, vp, vp]),
    "pt_truth_write": (c_int, [c_char_p, c_u32, c_u32, vp, vp]),
    "pt_result_write": (c_int, [c_char_p, c_u32, c_u32, vp]),
    "pt_exact_knn": (c_int, [vp, c_u32, c_u32, vp, c_u32, c_u32, c_int, vp, vp, c_int]),
    "pt_recall": (c_double, [vp, c_u32, vp, c_u32, c_u32, c_u32, c_u32]),
    "pt_recall_per_query": (c_int, [vp, c_u32, vp, c_u32, c_u32, c_u32, c_u32, vp]),
    "pt_navigating_node": (c_int, [vp, c_u32, c_u32, P(c_u32), c_int]),
    "pt_build_graph": (c_int, [vp, c_u32, c_u32, vp, c_u32, c_u32, c_u32, c_u32, c_u32, c_u32, vp, vp, c_int]),
    "pt_reachable": (c_int, [vp, vp, c_u32, c_u32, c_u32, P(c_u32)]),
}

# name -> (restype, argtypes) for include/panns_amd.h
AMD_API = {
    "panns_last_error": (c_char_p, []),
    "panns_create": (c_int, [c_int, P(vp)]),
    "panns_destroy": (c_int, [vp]),
    "panns_load_data_load": (c_int, [vp, c_char_p]),
    "panns_load_queries_load": (c_int, [vp, c_char_p]),
    "panns_set_base": (c_int, [vp, vp, c_u32, c_u32]),
    "panns_set_queries": (c_int, [vp, vp, c_u32, c_u32]),
    "panns_load_common_nsg_graph": (c_int, [vp, c_char_p]),
    "panns_set_graph": (c_int, [vp, c_u32, c_u32, c_u32, c_u32, vp, vp]),
    "panns_load_nsg_graph": (c_int, [vp, c_char_p]),
    "panns_set_distance": (c_int, [vp, c_int]),
    "panns_get_distance": (c_int, [vp, P(c_int)]),
    "panns_get_norms": (c_int, [vp, vp]),
    "panns_get_info": (c_int, [vp, P(c_u32), P(c_u32), P(c_u32), P(c_u32), P(c_u32), P(c_u64)]),
    "panns_prepare_init_ids": (c_int, [vp, c_u32, vp]),
    "panns_search": (c_int, [vp, P(PannsParams), vp, c_u32, vp, c_u32, P(PannsOutputs), P(PannsCounters)]),
    "panns_search_noalloc": (c_int, [vp, P(PannsParams), vp, c_u32, vp, c_u32, P(PannsOutputs),
                                     P(PannsCounters)]),
    "panns_subsearch_with_top_m": (c_int, [vp, c_u32, c_u32, c_u32, vp, vp, vp, P(c_u32), vp, P(c_u64)]),
    "panns_plan_create": (c_int, [vp, P(PannsParams), vp, c_u32, vp, vp, vp, c_u32, P(vp)]),
    "panns_plan_run": (c_int, [vp, c_int, P(c_float)]),
    "panns_plan_fetch": (c_int, [vp, P(PannsOutputs), P(PannsCounters)]),
    "panns_plan_destroy": (c_int, [vp]),
    "panns_plan_describe": (c_int, [vp, P(c_u32), P(c_u32), P(c_u32), P(c_u32)]),
    "panns_plan_create_ex": (c_int, [vp, P(PannsParams), vp, c_u32, vp, vp, vp, c_u32, P(PannsPlanOptions),
                                     P(vp)]),
    "panns_plan_describe_buckets": (c_int, [vp, P(c_u32), vp, vp, vp, vp, vp]),
    "panns_plan_autotune": (c_int, [vp, P(PannsParams), vp, c_u32, vp, vp, vp, c_u32, c_u32,
                                    P(PannsPlanOptions), P(c_float)]),
    "panns_exact_knn": (c_int, [vp, c_int, c_u32, vp, vp]),
    "panns_exact_knn_ex": (c_int, [vp, c_int, c_u32, c_int, vp, vp, P(c_u32)]),
    "panns_mfma_f32_probe": (c_int, [vp, vp, vp, vp]),
    "panns_search_trace": (c_int, [vp, P(PannsParams), vp, c_u32, vp, c_u32, c_u32, vp, vp, vp, vp]),
    "panns_search_trace_hits": (c_int, [vp, P(PannsParams), vp, c_u32, vp, c_u32, c_u32, vp, c_u32, vp, vp, vp,
                                        vp, vp, vp]),
    "panns_set_stop_model": (c_int, [vp, vp, c_u32]),
    "panns_mlp_train": (c_int, [vp, vp, vp, c_u64, vp, c_u32, c_u32, c_float, c_u64, vp]),
    "panns_mlp_grad": (c_int, [vp, vp, vp, vp, c_u32, vp, P(c_float)]),
    "panns_mfma_probe": (c_int, [vp, vp, vp, vp]),
    "panns_plan_phase_cycles": (c_int, [vp, vp]),
    "panns_device_copy": (c_int, [vp, c_u64, P(c_float)]),
}

_libs = {}


def _bind(lib, api):
    for name, (res, args) in api.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def tools():
    if "tools" not in _libs:
        if not os.path.exists(TOOLS_SO):
            raise NativeLibraryMissing(
                f"{TOOLS_SO} is not built; run `python -m parallel_anns_amd.build`")
        _libs["tools"] = _bind(ctypes.CDLL(TOOLS_SO), TOOLS_API)
    return _libs["tools"]


def amd():
    if "amd" not in _libs:
        tools()
        if not os.path.exists(AMD_SO):
            raise NativeLibraryMissing(
                f"{AMD_SO} is not built; run `python -m parallel_anns_amd.build` "
                "(the search path has no CPU fallback)")
        _libs["amd"] = _bind(ctypes.CDLL(AMD_SO), AMD_API)
    return _libs["amd"]


def ptr(a):
    """Data pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(vp)
