"""ctypes binding of libpt_gravity.so (the C ABI declared in include/pt_gravity.h).

The shared library is built in-tree by paratreet_amd/build.py.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpt_gravity.so")

PT_OK, PT_ERR_ARG, PT_ERR_HIP, PT_ERR_UNSUPPORTED, PT_ERR_CAPACITY, PT_ERR_STATE = 0, -1, -2, -3, -4, -5
PT_TREE_OCT, PT_TREE_BINARY_OCT, PT_TREE_KD, PT_TREE_LONGEST = 0, 1, 2, 3
PT_DECOMP_OCT, PT_DECOMP_BINARY_OCT, PT_DECOMP_SFC, PT_DECOMP_KD, PT_DECOMP_LONGEST = 0, 1, 2, 3, 4
PT_FP64, PT_FP32C = 0, 1
PT_HOST, PT_DEVICE = 0, 1
PT_FORCE_TREE, PT_FORCE_DIRECT = 0, 1
TIMERS = ("universe", "keys", "sort", "build", "moments", "traverse", "force", "integrate",
          "adjoint")

# Every symbol the header declares (checked by the CPU test suite).
EXPORTS = (
    "pt_config_default", "pt_ctx_create", "pt_ctx_destroy", "pt_last_error",
    "pt_get_device_view", "pt_set_particles", "pt_remake_universe", "pt_assign_keys",
    "pt_sort", "pt_build_tree", "pt_upward", "pt_traverse", "pt_interact", "pt_start_dual",
    "pt_compute_forces", "pt_kick_drift", "pt_kick", "pt_perturb", "pt_run", "pt_direct_sum", "pt_get_universe",
    "pt_get_keys", "pt_get_sorted", "pt_get_tree_size", "pt_get_nodes", "pt_get_node_data",
    "pt_get_buckets", "pt_get_lists_size", "pt_get_lists", "pt_get_accelerations",
    "pt_get_state", "pt_get_counters", "pt_set_timing", "pt_get_timings",
    "pt_force_vjp", "pt_leapfrog_forward", "pt_leapfrog_adjoint", "pt_get_node_hexadecapole",
    "pt_set_adjoint_precision", "pt_get_group_lists", "pt_sort_noalloc",
    "pt_set_fused_grouping",
)


class PtConfig(ctypes.Structure):
    _fields_ = [
        ("min_n_subtrees", ctypes.c_int32),
        ("min_n_partitions", ctypes.c_int32),
        ("max_particles_per_leaf", ctypes.c_int32),
        ("decomp_type", ctypes.c_int32),
        ("tree_type", ctypes.c_int32),
        ("num_iterations", ctypes.c_int32),
        ("dsoft", ctypes.c_double),
        ("has_dsoft", ctypes.c_int32),
        ("periodic", ctypes.c_int32),
        ("hexadecapole", ctypes.c_int32),
        ("period", ctypes.c_double * 3),
        ("n_replicas", ctypes.c_int32),
        ("dual_tree", ctypes.c_int32),
    ]


class PtCounters(ctypes.Structure):
    _fields_ = [
        ("n_node_ints", ctypes.c_uint64),
        ("n_part_ints", ctypes.c_uint64),
        ("n_opens", ctypes.c_uint64),
        ("n_closes", ctypes.c_uint64),
    ]

    def as_tuple(self):
        return (self.n_node_ints, self.n_part_ints, self.n_opens, self.n_closes)


class PtDeviceView(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64)] + [
        (name, ctypes.c_void_p) for name in
        ("x", "y", "z", "vx", "vy", "vz", "mass", "soft", "ax", "ay", "az", "pot",
         "keys", "sorted_keys", "perm", "stream")
    ]


class PtError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("pt status %d: %s" % (status, msg))
        self.status = status


_lib = None


def _init_torch_first():
    """When torch is in the process, let it bring up its HIP state before the
    library does: torch reports no GPU if the runtime was initialised by
    another user first.  Both resolve to the same libamdhip64.so.7."""
    import sys
    torch = sys.modules.get("torch")
    if torch is not None:
        try:
            torch.cuda.is_available()
        except Exception:
            pass


def load():
    """Loads libpt_gravity.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libpt_gravity.so is not built (%s); run "
                           "`python -c 'import __graft_entry__ as g; g.build()'` or "
                           "`python paratreet_amd/build.py`" % LIB_PATH)
    _init_torch_first()
    L = ctypes.CDLL(LIB_PATH)
    P, I64, I32, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    sig = {
        "pt_config_default": (None, [P]),
        "pt_ctx_create": (I32, [P, ctypes.c_int, P, P]),
        "pt_ctx_destroy": (None, [P]),
        "pt_last_error": (ctypes.c_char_p, [P]),
        "pt_get_device_view": (I32, [P, P]),
        "pt_set_particles": (I32, [P, I64, P, P, P, P, P, P, P, P, ctypes.c_int]),
        "pt_remake_universe": (I32, [P, P]),
        "pt_assign_keys": (I32, [P]),
        "pt_sort": (I32, [P]),
        "pt_sort_noalloc": (I32, [P, I64, P, P, P]),
        "pt_build_tree": (I32, [P]),
        "pt_upward": (I32, [P]),
        "pt_traverse": (I32, [P, D, P, P]),
        "pt_interact": (I32, [P, ctypes.c_int]),
        "pt_start_dual": (I32, [P, D, P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, P]),
        "pt_compute_forces": (I32, [P, D, P, ctypes.c_int, P]),
        "pt_kick_drift": (I32, [P, D]),
        "pt_kick": (I32, [P, D]),
        "pt_perturb": (I32, [P, D]),
        "pt_run": (I32, [P, ctypes.c_int, D, D, ctypes.c_int, P]),
        "pt_direct_sum": (I32, [P, P, I64, P]),
        "pt_get_universe": (I32, [P, P]),
        "pt_get_keys": (I32, [P, P]),
        "pt_get_sorted": (I32, [P, P, P]),
        "pt_get_tree_size": (I32, [P, P, P, P]),
        "pt_get_nodes": (I32, [P, P, P, P, P, P, P, P]),
        "pt_get_node_data": (I32, [P, P]),
        "pt_get_node_hexadecapole": (I32, [P, P]),
        "pt_get_buckets": (I32, [P, P]),
        "pt_get_lists_size": (I32, [P, P, P]),
        "pt_get_lists": (I32, [P, P, P, P, P]),
        "pt_get_group_lists": (I32, [P, D, P, P, P, P, P, P, P, P, P]),
        "pt_get_accelerations": (I32, [P, P, P]),
        "pt_get_state": (I32, [P, P, P]),
        "pt_get_counters": (I32, [P, P]),
        "pt_set_timing": (I32, [P, ctypes.c_int]),
        "pt_set_fused_grouping": (I32, [P, ctypes.c_int]),
        "pt_get_timings": (I32, [P, P]),
        "pt_force_vjp": (I32, [P, D, ctypes.c_int, P, P, P, ctypes.c_int]),
        "pt_set_adjoint_precision": (I32, [P, ctypes.c_int]),
        "pt_leapfrog_forward": (I32, [P, I32, D, D, ctypes.c_int, ctypes.c_int, I32, P]),
        "pt_leapfrog_adjoint": (I32, [P, P, P, P, P, P, ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
