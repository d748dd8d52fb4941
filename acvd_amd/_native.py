"""ctypes binding of libacvd_amd.so (the C-ABI declared in include/acvd_c.h)
and of the host-only mesh generator libacvd_meshgen.so.

The product path has no CPU fallback: if the HIP library is missing or no
device is visible, calls fail with an explicit error.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libacvd_amd.so")
MESHGEN_PATH = os.path.join(_HERE, "libacvd_meshgen.so")

ACVD_OK = 0
METRIC_ISOTROPIC = 0
METRIC_QEM = 1


class AcvdError(RuntimeError):
    """A C-ABI call returned a negative status."""


class Params(ctypes.Structure):
    _fields_ = [
        ("number_of_clusters", ctypes.c_int32),
        ("metric", ctypes.c_int32),
        ("quadrics_optimization_level", ctypes.c_int32),
        ("unconstrained_initialization", ctypes.c_int32),
        ("max_number_of_loops", ctypes.c_int32),
        ("max_number_of_convergences", ctypes.c_int32),
        ("console_output", ctypes.c_int32),
        ("seed", ctypes.c_uint32),
        ("initial_sampling", ctypes.c_int32),
        ("curvature_indicator", ctypes.c_int32),
        ("first_phase_rounds", ctypes.c_int32),
        ("write_energy_log", ctypes.c_int32),
        ("gradation", ctypes.c_double),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("sweep_segments", ctypes.c_int32),
        ("tile_on_divisor", ctypes.c_int32),
        ("chunk_passes", ctypes.c_int32),
        ("short_chunk_passes", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("loops", ctypes.c_int64),
        ("convergences", ctypes.c_int64),
        ("total_moves", ctypes.c_int64),
        ("last_disconnected", ctypes.c_int32),
        ("finished", ctypes.c_int32),
        ("ctx_passes", ctypes.c_int64),
        ("ctx_evaluated", ctypes.c_int64),
        ("ctx_proposed", ctypes.c_int64),
        ("ctx_accepted", ctypes.c_int64),
        ("converge_ms", ctypes.c_double),
        ("build_ms", ctypes.c_double),
        ("n_items", ctypes.c_int64),
        ("n_adjacency", ctypes.c_int64),
        ("kernel_ms", ctypes.c_double * 4),
        ("kernel_launches", ctypes.c_int64 * 4),
        ("ctx_rounds", ctypes.c_int64),
        ("ctx_walked", ctypes.c_int64),
        ("ctx_reeval_live", ctypes.c_int64),
        ("ctx_reevaluated", ctypes.c_int64),
        ("ctx_accepted_constrained", ctypes.c_int64),
        ("ctx_rechecked", ctypes.c_int64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["kernel_ms"] = list(self.kernel_ms)
        d["kernel_launches"] = list(self.kernel_launches)
        return d


# every symbol declared in include/acvd_c.h
EXPORTS = [
    "acvd_version", "acvd_last_error", "acvd_device_count", "acvd_params_init", "acvd_create", "acvd_destroy",
    "acvd_set_input_f64", "acvd_set_input_f32", "acvd_add_input_f64", "acvd_add_input_f32", "acvd_num_meshes",
    "acvd_set_params", "acvd_get_params", "acvd_set_indicator", "acvd_get_indicator", "acvd_set_initial_labels",
    "acvd_process_clustering", "acvd_get_labels", "acvd_get_cluster_sizes", "acvd_get_cluster_points",
    "acvd_get_energy", "acvd_get_stats", "acvd_get_adjacency_size", "acvd_get_adjacency", "acvd_get_items",
    "acvd_get_scales", "acvd_build", "acvd_initial_sampling", "acvd_set_timing", "acvd_build_output",
    "acvd_get_output_sizes", "acvd_get_output", "acvd_energy_create", "acvd_energy_destroy", "acvd_energy_add_mesh",
    "acvd_energy_num_items", "acvd_energy_forward", "acvd_energy_backward", "acvd_reference_initial_sampling",
    "acvd_set_fixed_clusters", "acvd_subdivide", "acvd_subdivide_input", "acvd_get_input_sizes", "acvd_get_input",
    "acvd_get_energy_log", "acvd_force_manifold",
    "acvd_set_frozen_clusters", "acvd_vertex_manifold", "acvd_force_manifold_step", "acvd_fix_mesh_boundaries",
    "acvd_mesh_boundary_fix", "acvd_reference_seed_sampling", "acvd_set_tuning", "acvd_get_tuning", "acvd_autotune",
    "acvd_energy_forward_cvt_noalloc",
]

_lib = None
_gen = None


def lib():
    """Loads libacvd_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AcvdError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (make -C acvd_amd/csrc)")
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # soname as ROCm's).  Loading torch first makes this library bind to that
    # runtime, so device pointers and streams from torch tensors are valid
    # here; loading ROCm's copy first would leave torch without a device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64 = ctypes.c_int32, ctypes.c_int64
    L.acvd_version.restype = ctypes.c_char_p
    L.acvd_last_error.restype = ctypes.c_char_p
    L.acvd_device_count.restype = ctypes.c_int
    L.acvd_params_init.argtypes = [ctypes.POINTER(Params)]
    L.acvd_params_init.restype = None
    L.acvd_create.argtypes = [ctypes.POINTER(P), ctypes.c_int]
    L.acvd_destroy.argtypes = [P]
    for f in ("acvd_set_input_f64", "acvd_set_input_f32"):
        getattr(L, f).argtypes = [P, i64, P, i64, P]
    for f in ("acvd_add_input_f64", "acvd_add_input_f32"):
        getattr(L, f).argtypes = [P, i64, P, i64, P, ctypes.POINTER(i32)]
    L.acvd_num_meshes.argtypes = [P]
    L.acvd_set_params.argtypes = [P, i32, ctypes.POINTER(Params)]
    L.acvd_get_params.argtypes = [P, i32, ctypes.POINTER(Params)]
    L.acvd_set_indicator.argtypes = [P, i32, P]
    L.acvd_get_indicator.argtypes = [P, i32, P]
    L.acvd_set_initial_labels.argtypes = [P, i32, P]
    L.acvd_process_clustering.argtypes = [P]
    L.acvd_build.argtypes = [P]
    L.acvd_initial_sampling.argtypes = [P]
    L.acvd_set_timing.argtypes = [P, ctypes.c_int]
    L.acvd_build_output.argtypes = [P, i32, i32]
    L.acvd_get_output_sizes.argtypes = [P, i32, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.acvd_get_output.argtypes = [P, i32, P, P]
    for f in ("acvd_get_labels", "acvd_get_cluster_sizes", "acvd_get_cluster_points", "acvd_get_energy",
              "acvd_get_scales"):
        getattr(L, f).argtypes = [P, i32, P]
    L.acvd_get_stats.argtypes = [P, i32, ctypes.POINTER(Stats)]
    L.acvd_get_adjacency_size.argtypes = [P, i32, ctypes.POINTER(i64)]
    L.acvd_get_adjacency.argtypes = [P, i32, P, P]
    L.acvd_get_items.argtypes = [P, i32, P, P]
    L.acvd_reference_initial_sampling.argtypes = [i64, P, i64, P, i32, P, ctypes.c_double, i32, P]
    L.acvd_energy_create.argtypes = [ctypes.POINTER(P), ctypes.c_int]
    L.acvd_energy_destroy.argtypes = [P]
    L.acvd_energy_add_mesh.argtypes = [P, i64, i64, P, i32, i32, i32, ctypes.POINTER(i32)]
    L.acvd_energy_num_items.argtypes = [P, ctypes.POINTER(i64), ctypes.POINTER(i32)]
    L.acvd_energy_forward.argtypes = [P, P, P, P, P, P]
    L.acvd_energy_backward.argtypes = [P, P, P, P, P, P, P, P]
    L.acvd_set_fixed_clusters.argtypes = [P, i32, i64, P]
    L.acvd_subdivide.argtypes = [i64, P, i64, P, ctypes.POINTER(i64), ctypes.POINTER(i64), P, P, P]
    L.acvd_subdivide_input.argtypes = [P, i32, i64, ctypes.POINTER(i32)]
    L.acvd_get_input_sizes.argtypes = [P, i32, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.acvd_get_input.argtypes = [P, i32, P, P]
    L.acvd_get_energy_log.argtypes = [P, i32, i64, ctypes.POINTER(i64), P, P, P]
    L.acvd_force_manifold.argtypes = [P, i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i64)]
    L.acvd_set_frozen_clusters.argtypes = [P, i32, i64, P]
    L.acvd_fix_mesh_boundaries.argtypes = [P, i32, ctypes.POINTER(i32)]
    L.acvd_reference_seed_sampling.argtypes = [i64, i32, P]
    L.acvd_set_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
    L.acvd_get_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
    L.acvd_autotune.argtypes = [P, i32, ctypes.POINTER(Tuning), ctypes.POINTER(ctypes.c_double)]
    L.acvd_mesh_boundary_fix.argtypes = [i64, P, i64, P, P, i32, ctypes.POINTER(i64), ctypes.POINTER(i64), P, P]
    L.acvd_vertex_manifold.argtypes = [i64, i64, P, P]
    L.acvd_force_manifold_step.argtypes = [i64, i64, P, i64, P, ctypes.POINTER(i32), P, P, i64,
                                           ctypes.POINTER(i64)]
    L.acvd_energy_forward_cvt_noalloc.argtypes = [ctypes.c_int, i64, i32, P, P, P, P, P]
    _lib = L
    return L


def check(rc):
    if rc != ACVD_OK:
        msg = lib().acvd_last_error().decode(errors="replace")
        raise AcvdError(f"acvd call failed ({rc}): {msg}")
    return rc


def meshgen():
    global _gen
    if _gen is not None:
        return _gen
    if not os.path.exists(MESHGEN_PATH):
        raise AcvdError(f"{MESHGEN_PATH} is missing: build it with __graft_entry__.build()")
    G = ctypes.CDLL(MESHGEN_PATH)
    P = ctypes.c_void_p
    i64p = ctypes.POINTER(ctypes.c_int64)
    G.acvd_gen_icosphere_sizes.argtypes = [ctypes.c_int, i64p, i64p]
    G.acvd_gen_icosphere.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, P, P]
    G.acvd_gen_star_icosphere.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, P, P]
    G.acvd_gen_cube_sizes.argtypes = [ctypes.c_int, i64p, i64p]
    G.acvd_gen_cube.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, P, P]
    G.acvd_gen_torus_sizes.argtypes = [ctypes.c_int, ctypes.c_int, i64p, i64p]
    G.acvd_gen_torus.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_uint64, P, P]
    G.acvd_mean_curvature.argtypes = [ctypes.c_int64, P, P, ctypes.c_int64, P, P]
    G.acvd_gen_seed_labels.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64, P]
    G.acvd_gen_batch_levels.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, P]
    _gen = G
    return G
