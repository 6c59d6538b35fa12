"""ctypes binding of libanneal_hip.so (the C ABI declared in include/anneal_hip.h).

The library is required: there is no CPU fallback.  Importing this module
without the built library, or calling into it without a HIP device, raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libanneal_hip.so")

c_int, c_int64, c_double, c_void_p = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_intp = ctypes.POINTER(ctypes.c_int)
c_i64p = ctypes.POINTER(ctypes.c_int64)

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "anneal_last_error": [],
    "anneal_abi_version": [],
    "anneal_ke_host": [c_double, c_double, c_void_p, c_void_p],
    "anneal_fem_create": [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_double, c_double, c_void_p, c_void_p,
                          c_int, c_int, c_void_p],
    "anneal_fem_destroy": [c_void_p],
    "anneal_fem_info": [c_void_p, c_i64p, c_intp, c_void_p],
    "anneal_fem_level_info": [c_void_p, c_int, c_intp, c_intp, c_intp, c_intp],
    "anneal_fem_set_Ee": [c_void_p, c_void_p],
    "anneal_fem_rebuild": [c_void_p],
    "anneal_fem_apply_f32": [c_void_p, c_void_p, c_void_p, c_int],
    "anneal_fem_apply_f64": [c_void_p, c_void_p, c_void_p, c_int],
    "anneal_fem_solve": [c_void_p, c_void_p, c_void_p, c_int, c_double, c_int, c_int, c_intp],
    "anneal_fem_autotune": [c_void_p, c_int, c_intp, c_void_p, c_intp],
    "anneal_fem_set_level_tile": [c_void_p, c_int, c_int],
    "anneal_fem_set_option": [c_void_p, ctypes.c_char_p, c_double],
    "anneal_fem_vcycle": [c_void_p, c_void_p, c_void_p, c_int],
    "anneal_fem_smooth_step_f32": [c_void_p, c_void_p, c_void_p, c_void_p, c_int],
    "anneal_fem_element_energy": [c_void_p, c_void_p, c_int, c_void_p],
    "anneal_fem_sensitivity": [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_double, c_double, c_double, c_void_p, c_void_p],
    "anneal_filter_create": [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_double, c_void_p, c_void_p],
    "anneal_filter_destroy": [c_void_p],
    "anneal_filter_forward": [c_void_p, c_void_p, c_void_p],
    "anneal_filter_backward": [c_void_p, c_void_p, c_void_p],
    "anneal_filter_backward_pair": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "anneal_filter_physical": [c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p],
    "anneal_oc_bisect": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_double, c_double,
                         c_double, c_double, c_double, c_void_p, c_intp, c_intp, ctypes.POINTER(c_double)],
    "anneal_oc_begin": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_double, c_double,
                        c_double, c_double, c_double, c_void_p],
    "anneal_oc_begin_stats": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_void_p,
                              c_double, c_double, c_double, c_void_p],
    "anneal_oc_poll": [c_void_p, c_intp],
    "anneal_oc_end": [c_void_p, c_intp, c_intp, ctypes.POINTER(c_double)],
    "anneal_stiffness_law": [c_void_p, c_void_p, c_void_p, c_double, c_double, c_double, c_int64, c_void_p,
                             c_void_p],
    "anneal_design_stats": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                            c_void_p, c_void_p, c_void_p],
    "anneal_dot_f64": [c_void_p, c_void_p, c_int64, c_int64, c_int, c_void_p, c_void_p, c_void_p],
    "anneal_dot_f64_noalloc": [c_void_p, c_void_p, c_int64, c_int64, c_int, c_void_p, c_void_p],
    "anneal_unet_create": [ctypes.POINTER(c_void_p), c_int, c_int, c_intp, c_int, c_void_p],
    "anneal_unet_destroy": [c_void_p],
    "anneal_unet_nparams": [c_void_p, c_i64p],
    "anneal_unet_get_params": [c_void_p, c_void_p],
    "anneal_unet_set_params": [c_void_p, c_void_p],
    "anneal_unet_get_grads": [c_void_p, c_void_p],
    "anneal_unet_forward": [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p],
    "anneal_unet_backward": [c_void_p, c_void_p],
    "anneal_unet_zero_grad": [c_void_p],
    "anneal_unet_adam": [c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, c_int,
                         ctypes.c_float],
    "anneal_unet_rel_mse": [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                            c_void_p],
    "anneal_unet_restrict": [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, ctypes.c_float, c_void_p],
    "anneal_unet_prolong": [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_int, c_int, c_int,
                            ctypes.c_float, c_int, c_void_p],
    "anneal_fem_set_learned": [c_void_p, c_void_p, c_int, c_double, c_int],
    "anneal_fem_learned_stage": [c_void_p, c_void_p, c_void_p, c_int],
    "anneal_fem_learned_pair": [c_void_p, c_void_p, c_void_p, c_int, c_int, c_double, c_void_p, c_void_p],
    "anneal_batch_create": [ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), c_int],
    "anneal_fem_group_create": [ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), c_int, c_int, c_void_p],
    "anneal_batch_destroy": [c_void_p],
    "anneal_batch_solve": [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), c_intp, c_double, c_int, c_int,
                           c_intp, c_intp],
    "anneal_fem_stress": [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                          c_void_p],
    "anneal_fields_corner_values": [c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_void_p, c_void_p,
                                    c_void_p, c_int64, c_void_p, c_void_p],
    "anneal_region_select": [c_int, c_int, c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                             c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "anneal_region_sdf": [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p],
    "anneal_boundary_fields": [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                               c_void_p],
    "anneal_gather_node_bits": [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "anneal_voxel_components": [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_i64p, c_void_p],
    "anneal_voxel_link": [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, ctypes.c_uint32, c_void_p, c_void_p,
                          c_void_p, c_i64p, ctypes.POINTER(ctypes.c_uint32), c_void_p],
    "anneal_voxel_threshold": [c_int64, c_void_p, c_double, c_void_p, c_void_p],
    "anneal_voxel_count_above": [c_int64, c_void_p, c_double, c_void_p, c_i64p, c_void_p],
    "anneal_output_field": [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "anneal_choose_level_work_bytes": [],
    "anneal_choose_level": [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_double),
                            ctypes.POINTER(c_double), c_void_p],
    "anneal_paths_sssp": [c_int, c_int, c_int, c_void_p, c_void_p, c_double, c_void_p, ctypes.c_uint32, c_void_p,
                          c_void_p, c_void_p, c_void_p, c_intp, c_void_p],
    "anneal_paths_trace": [c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                           c_i64p, c_i64p, ctypes.POINTER(c_double), c_void_p],
    "anneal_paths_raise": [c_int, c_int, c_int, c_void_p, c_int64, c_void_p, c_void_p, c_double, c_void_p, c_void_p],
}
_RESTYPE = {"anneal_last_error": ctypes.c_char_p, "anneal_choose_level_work_bytes": c_int64}


class AnnealLibError(RuntimeError):
    """A call into libanneal_hip.so failed; the message is the library's."""

    def __init__(self, name, status, msg):
        super().__init__(f"{name} failed (status {status}): {msg}")
        self.status = status
        self.msg = msg


_lib = None


def load():
    """Load the shared library once (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m anneal_amd.build` "
                          f"(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, c_int)
    _lib = lib
    return lib


def call(name, *args):
    """Call an entry point and raise AnnealLibError on a non-zero status."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != 0:
        raise AnnealLibError(name, st, lib.anneal_last_error().decode())
    return st


def ptr(t):
    """Device (or host) address of a contiguous torch tensor / numpy array as c_void_p."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        if not t.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return c_void_p(t.data_ptr())
    if not t.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return c_void_p(t.ctypes.data)


def stream_ptr():
    """The current torch HIP stream as a raw hipStream_t."""
    import torch
    return c_void_p(torch.cuda.current_stream().cuda_stream)
