"""ctypes binding of libspecfwat_hip.so (the C ABI declared in include/specfwat_hip.h).

Every entry point is Fortran-callable: all arguments by reference.  `call()`
converts Python scalars to by-reference C scalars and numpy arrays to
pointers according to a per-function signature string:

    L  long* Mesh_pointer (a HandleBox)      i / f / d / l  int / float / double / long scalar
    I  int32 array   F  float32 array   D  float64 array   (None -> NULL)
    P  device address (a Python int from the caller's device allocation; None -> NULL)

There is no fallback: if the HIP library is missing the import-time load
raises, so a GPU run can never silently take another path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspecfwat_hip.so")

SIGNATURES = {
    "initialize_gpu_device_": "iii",
    "prepare_constants_device_": "LiiiiFFFFFFFFFFffIIFFFFFiiIIFFiFIiIFiii",
    "prepare_fields_elastic_device_": "LFFFFFFFi" + "F" * 21 + "iIiii",
    "prepare_fields_elastic_adj_dev_": "LiiFii",
    "prepare_sim2_or_3_const_device_": "Lii",
    "update_displacement_cuda_": "Lfffi",
    "compute_forces_viscoelastic_cuda_": "Lifiiiii",
    "compute_stacey_viscoelastic_cuda_": "Liii",
    "compute_add_sources_viscoelastic_cuda_": "LiDi",
    "add_sources_el_sim_type_2_or_3_": "LiFi",
    "kernel_3_a_cuda_": "Lfi",
    "compute_seismograms_cuda_": "Li",
    "transfer_seismograms_from_device_": "LF",
    "compute_kernels_elastic_cuda_": "Lf",
    "transfer_kernels_el_to_host_": "LFFFi",
    "transfer_kernels_el_to_host_dp_": "LDDDi",
    "transfer_kernels_ani_to_host_dp_": "LDi",
    "transfer_kernels_hess_el_tohost_dp_": "LDi",
    "transfer_fields_el_to_device_": "iFFFL",
    "transfer_fields_el_from_device_": "iFFFL",
    "transfer_b_fields_to_device_": "iFFFL",
    "transfer_b_fields_from_device_": "iFFFL",
    "transfer_stacey_store_from_device_": "LF",
    "transfer_stacey_store_to_device_": "LF",
    "prepare_cleanup_device_": "Liii",
    "fwat_set_stf_": "LDi",
    "fwat_set_adjoint_sources_": "LFi",
    "fwat_set_fk_injection_": "LiiiFF",
    "fwat_iterate_time_": "Lifii",
    "fwat_last_force_time_": "LDI",
    "fwat_last_assembly_time_": "LDI",
    "fwat_synchronize_": "L",
    "fwat_copy_probe_": "liD",
    "fwat_copy_probe_unrolled_": "liiiD",
    "fwat_copy_forward_to_backward_": "LL",
    "fwat_move_stacey_store_": "LL",
    "fwat_set_static_precision_": "Li",
    "fwat_set_tuning_": "Liii",
    "fwat_set_lattice_": "Liii",
    "fwat_column_block_order_": "iiiiI",
    "fwat_set_interior_": "Li",
    "fwat_interior_info_": "LIIII",
    "fwat_set_brick_": "Liiii",
    "fwat_brick_info_": "Lill",
    "fwat_set_kernel_block_": "Li",
    "fwat_set_assembly_lanes_": "Li",
    "fwat_set_assembly_block_": "Li",
    "fwat_set_floc_layout_": "Li",
    "fwat_floc_row_layout_": "I",
    "fwat_set_graph_": "Li",
    "fwat_assembly_csr_": "iiIiIIiIiIiIiiII",
    "fwat_set_slabs_": "Li",
    "fwat_batch_create_": "LLiIfFID",
    "fwat_batch_set_tuning_": "Li",
    "fwat_batch_set_interior_": "Li",
    "fwat_batch_set_store_precision_": "Li",
    "fwat_batch_forward_": "Li",
    "fwat_batch_set_adjoint_sources_": "LF",
    "fwat_batch_adjoint_": "Li",
    "fwat_batch_seismograms_": "LF",
    "fwat_batch_kernels_": "LFDDD",
    "fwat_batch_kernels_ani_": "LFDD",
    "fwat_batch_last_time_": "LDIDI",
    "fwat_batch_destroy_": "L",
    "smooth_pde_cuda_": "iiiiiiIii" + "F" * 10 + "F" + "FFFFF" + "I" + "I" + "ff" + "iiIII" + "FFFFF" + "Ifi",
    "fwat_smooth_set_faithful_": "i",
    "fwat_smooth_pde_batch_": "iiiii" + "F" * 10 + "F" + "FFFFF" + "I" + "ff" + "FF" + "Ifii",
    "fwat_smooth_last_time_": "DD",
    "fwat_post_create_": "Liii",
    "fwat_post_set_model_": "LFFFFFF",
    "fwat_post_add_event_": "LLi",
    "fwat_post_add_event_files_": "LFFi",
    "fwat_post_get_kernels_": "LD",
    "fwat_post_get_hess_": "LD",
    "fwat_post_set_hess_": "LD",
    "fwat_post_invert_hess_": "Li",
    "fwat_post_apply_precond_": "L",
    "fwat_post_info_": "Liii",
    "fwat_post_last_time_": "Ld",
    "fwat_post_destroy_": "L",
    "fwat_gll2grid_prepare_": "LiiIFFF" + "f" * 6 + "iiii" + "l",
    "fwat_gll2grid_": "LiFF",
    "fwat_gll2grid_owner_": "LID",
    "fwat_gll2grid_last_time_": "Ld",
    "fwat_gll2grid_destroy_": "L",
    "fwat_taper_grid_": "f" * 18 + "iiii" + "FF",
    "fwat_opt_sd_direction_": "lFFF",
    "fwat_opt_lbfgs_direction_": "lliFFFfffFd",
    "fwat_opt_cg_direction_": "llFFFFd",
    "fwat_opt_model_update_": "liiiifFFF",
    "fwat_opt_last_time_": "d",
    "fwat_waveform_adjoint_batch_": "iidFFiDdiiFDDF",
    "fwat_waveform_adjoint_last_time_": "d",
    "fwat_waveform_adjoint_batch_noalloc_": "iidPPiDdiiFPPPl",
    "fwat_cc_adjoint_batch_": "iidFFiDdiiDiFDDF",
    "fwat_mt_adjoint_batch_": "iidFFiDdiiDiFDDF",
    "fwat_ep_adjoint_batch_": "iidFFiDdidiFDDF",
    "fwat_win_sel_batch_": "iidFFdDDdDiID",
    "fwat_flexwin_batch_": "iidFFdDDddDiIDD",
    "fwat_conv_adjoint_batch_": "iidFFFFDdiDDFF",
}

_lib = None


class HandleBox:
    """Holds the opaque `long Mesh_pointer` the Fortran side would keep."""

    def __init__(self):
        self.value = C.c_long(0)

    def __bool__(self):
        return self.value.value != 0


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"specfwat_amd: HIP library {LIB_PATH} is not built; run __graft_entry__.build() "
                "(hipcc --offload-arch=gfx950). There is no CPU fallback.")
        _lib = C.CDLL(LIB_PATH)
        for name in SIGNATURES:
            fn = getattr(_lib, name)
            fn.restype = None
            fn.argtypes = [C.c_void_p] * len(SIGNATURES[name])
    return _lib


_SCALAR = {"i": C.c_int, "f": C.c_float, "d": C.c_double, "l": C.c_long}
_ARRAY = {"I": np.int32, "F": np.float32, "D": np.float64}


def call(name, *args):
    """Call a C-ABI entry point; returns the list of by-reference scalars."""
    sig = SIGNATURES[name]
    if len(sig) != len(args):
        raise TypeError(f"{name}: expected {len(sig)} arguments, got {len(args)}")
    fn = getattr(lib(), name)
    cargs, keep = [], []
    for code, a in zip(sig, args):
        if code == "L":
            if not isinstance(a, HandleBox):
                raise TypeError(f"{name}: Mesh_pointer must be a HandleBox")
            cargs.append(C.cast(C.pointer(a.value), C.c_void_p))
        elif code == "P":
            cargs.append(None if a is None else C.c_void_p(int(a)))
        elif code in _SCALAR:
            v = _SCALAR[code](a)
            keep.append(v)
            cargs.append(C.cast(C.pointer(v), C.c_void_p))
        else:
            if a is None:
                cargs.append(None)
                continue
            if not isinstance(a, np.ndarray) or a.dtype != _ARRAY[code]:
                raise TypeError(f"{name}: argument {len(cargs)} must be a {_ARRAY[code].__name__} array")
            if not (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]):
                raise TypeError(f"{name}: argument {len(cargs)} must be contiguous")
            keep.append(a)
            cargs.append(C.c_void_p(a.ctypes.data))
    fn(*cargs)
    return [k.value for k in keep if isinstance(k, (C.c_int, C.c_float, C.c_double, C.c_long))]


def exported_symbols():
    """Names the shared library exports (checked against the header by the tests)."""
    return sorted(SIGNATURES)
