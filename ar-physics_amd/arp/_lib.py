"""ctypes binding of libarp.so (the C ABI declared in include/arp.h).

The library must have been built (``python ar-physics_amd/build.py``); loading fails
loudly otherwise -- there is no fallback implementation."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libarp.so")

SYMBOLS = [
    "arp_create", "arp_destroy", "arp_last_error", "arp_set_stream", "arp_synchronize",
    "arp_norm_const", "arp_default_interaction", "arp_resolve_interactions", "arp_tableau_by_name",
    "arp_scene_upload", "arp_set_integrator", "arp_update", "arp_step", "arp_time", "arp_download",
    "arp_upload_state", "arp_compute_forces", "arp_lincomb", "arp_velocity", "arp_neighbors_update",
    "arp_grid_info", "arp_neighbor_dump", "arp_cell_counts", "arp_profile", "arp_profile_query",
    "arp_profile_reset", "arp_adjoint_record", "arp_adjoint_backward",
    "arp_profile_select", "arp_batch_create", "arp_batch_destroy", "arp_batch_last_error", "arp_batch_step",
    "arp_batch_tune", "arp_set_stiffness_field", "arp_adjoint_stiffness_grad",
    "arp_set_neighbor_skin", "arp_neighbor_stats", "arp_speculation_stats", "arp_set_neighbor_max_age",
    "arp_scene_parse", "arp_scene_free", "arp_scene_error", "arp_scene_settings", "arp_scene_object",
    "arp_scene_generate", "arp_scene_arrays", "arp_scene_contains", "arp_scene_apply", "arp_toml_to_json",
    "arp_kernel_radius", "arp_num_materials", "arp_add_particles", "arp_scene_spawn", "arp_num_particles",
    "arp_num_solids", "arp_neighbor_lists_fit", "arp_lincomb_noalloc", "arp_batch_adjoint_record",
    "arp_batch_adjoint_backward", "arp_batch_tuning", "arp_set_neighbor_short_skin", "arp_neighbor_refreshes",
    "arp_profile_durations", "arp_tune_lists", "arp_batch_tune_lists",
]


class ArpError(RuntimeError):
    pass


_lib = None
vp = C.c_void_p


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ArpError("libarp.so is not built: run python ar-physics_amd/build.py (%s)" % LIBPATH)
    # One HIP runtime per process: torch ships its own libamdhip64, and a second copy loaded
    # first (the one libarp.so links) leaves torch without devices.  Loading torch first makes
    # libarp.so bind to torch's copy.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIBPATH)
    L.arp_create.argtypes = [C.c_int, C.c_uint32, C.POINTER(vp)]
    L.arp_destroy.argtypes = [vp]
    L.arp_destroy.restype = None
    L.arp_last_error.argtypes = [vp]
    L.arp_last_error.restype = C.c_char_p
    L.arp_set_stream.argtypes = [vp, vp]
    L.arp_synchronize.argtypes = [vp]
    L.arp_norm_const.argtypes = [C.c_uint32, C.c_float]
    L.arp_norm_const.restype = C.c_float
    L.arp_default_interaction.argtypes = [vp, vp, C.c_float, vp]
    L.arp_default_interaction.restype = None
    L.arp_resolve_interactions.argtypes = [C.c_uint32, vp, vp, vp, C.c_float, vp]
    L.arp_resolve_interactions.restype = None
    L.arp_tableau_by_name.argtypes = [C.c_char_p, vp]
    L.arp_scene_upload.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32,
                                   vp, vp, C.c_float, C.c_float, C.c_float]
    L.arp_set_integrator.argtypes = [vp, vp, C.c_float, C.c_uint32]
    L.arp_update.argtypes = [vp, C.c_uint32]
    L.arp_step.argtypes = [vp, C.c_float, C.c_uint32]
    L.arp_time.argtypes = [vp]
    L.arp_time.restype = C.c_double
    L.arp_download.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32]
    L.arp_upload_state.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32]
    L.arp_compute_forces.argtypes = [vp, vp, vp, vp, vp]
    L.arp_lincomb.argtypes = [vp, C.c_uint32, vp, vp, vp, vp, vp]
    L.arp_lincomb_noalloc.argtypes = [vp, C.c_uint32, vp, vp, vp, vp, vp]
    L.arp_batch_adjoint_record.argtypes = [vp, C.c_float, C.c_uint32, C.c_uint32]
    L.arp_batch_adjoint_backward.argtypes = [vp, vp, vp, vp]
    L.arp_batch_tuning.argtypes = [vp, vp, vp]
    L.arp_velocity.argtypes = [vp, C.c_float, vp, vp, vp, vp]
    L.arp_neighbors_update.argtypes = [vp, vp]
    L.arp_grid_info.argtypes = [vp, vp]
    L.arp_neighbor_dump.argtypes = [vp, vp, vp, vp, C.c_uint64]
    L.arp_cell_counts.argtypes = [vp, C.c_uint64, C.c_uint64, vp]
    L.arp_profile.argtypes = [vp, C.c_int]
    L.arp_profile_query.argtypes = [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    L.arp_profile_reset.argtypes = [vp]
    L.arp_profile_select.argtypes = [vp, C.c_char_p]
    L.arp_batch_create.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(vp)]
    L.arp_batch_destroy.argtypes = [vp]
    L.arp_batch_destroy.restype = None
    L.arp_batch_last_error.argtypes = [vp]
    L.arp_batch_last_error.restype = C.c_char_p
    L.arp_batch_step.argtypes = [vp, C.c_float, C.c_uint32]
    L.arp_batch_tune.argtypes = [vp, C.c_float, C.c_uint32, C.POINTER(C.c_uint32)]
    L.arp_set_stiffness_field.argtypes = [vp, vp, C.c_uint32]
    L.arp_adjoint_stiffness_grad.argtypes = [vp, vp, C.c_uint32]
    L.arp_set_neighbor_skin.argtypes = [vp, C.c_float, C.c_uint32, C.c_uint32]
    L.arp_set_neighbor_max_age.argtypes = [vp, C.c_uint32]
    L.arp_neighbor_lists_fit.argtypes = [C.c_uint32, C.c_uint32]
    L.arp_neighbor_stats.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.arp_set_neighbor_short_skin.argtypes = [vp, C.c_float]
    L.arp_profile_durations.argtypes = [vp, C.c_char_p, C.POINTER(C.c_float), C.c_uint64, C.POINTER(C.c_uint64)]
    L.arp_neighbor_refreshes.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.arp_tune_lists.argtypes = [vp, C.c_float, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.arp_batch_tune_lists.argtypes = [vp, C.c_float, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.arp_speculation_stats.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.arp_adjoint_record.argtypes = [vp, C.c_float, C.c_uint32, C.c_uint32]
    L.arp_adjoint_backward.argtypes = [vp, vp, vp, vp]
    L.arp_scene_parse.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_size_t]
    L.arp_scene_free.argtypes = [vp]
    L.arp_scene_free.restype = None
    L.arp_scene_error.argtypes = [vp]
    L.arp_scene_error.restype = C.c_char_p
    L.arp_scene_settings.argtypes = [vp, vp]
    L.arp_scene_object.argtypes = [vp, C.c_uint32, vp]
    L.arp_scene_generate.argtypes = [vp, C.c_uint32]
    L.arp_scene_arrays.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp]
    L.arp_scene_contains.argtypes = [vp, C.c_uint32, vp, C.c_uint32, vp]
    L.arp_scene_apply.argtypes = [vp, vp]
    L.arp_toml_to_json.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.arp_kernel_radius.argtypes = [vp]
    L.arp_kernel_radius.restype = C.c_float
    L.arp_num_materials.argtypes = [vp]
    L.arp_num_materials.restype = C.c_uint32
    for nm in ("arp_num_particles", "arp_num_solids"):
        getattr(L, nm).argtypes = [vp]
        getattr(L, nm).restype = C.c_uint32
    L.arp_add_particles.argtypes = [vp, vp, vp, C.c_uint32, C.POINTER(C.c_uint32)]
    L.arp_scene_spawn.argtypes = [vp, vp, C.c_uint32, vp, C.POINTER(C.c_uint32)]
    _lib = L
    return L


def check(ctx, rc):
    if rc != 0:
        msg = lib().arp_last_error(ctx)
        raise ArpError("arp error %d: %s" % (rc, msg.decode() if msg else ""))
