"""ctypes binding of libmsl.so (the C-ABI in include/msl.h).

The engine is the HIP library; this module only declares its signatures.
There is no fallback: if libmsl.so is missing or cannot be loaded, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsl.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "msl.h")


class MslError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"msl status {code}: {msg}")
        self.code = code


class msl_info(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("MaximumX", "MinimumX", "MaximumY", "MinimumY", "MaximumZ", "MinimumZ")] + [
        ("numberOfSlices", C.c_int), ("kmax", C.c_float), ("conventionaldz", C.c_float),
        ("maxposIm", C.c_float), ("maxposEW", C.c_float), ("minposEW", C.c_float), ("minposIm", C.c_float),
        ("imagecontrast", C.c_float), ("wavelength", C.c_float), ("sigma2", C.c_float), ("dzFD", C.c_float),
        ("natoms", C.c_int), ("nslab", C.c_int), ("resX", C.c_int), ("resY", C.c_int), ("band_tiles", C.c_int)]


class msl_stem_batch_desc(C.Structure):
    _fields_ = [("nprobe", C.c_int), ("posx", C.c_void_p), ("posy", C.c_void_p), ("nconfig", C.c_int),
                ("seed", C.c_uint64), ("sigma", C.c_float), ("dz", C.c_float), ("nslices", C.c_int),
                ("adf_inner", C.c_float), ("adf_outer", C.c_float), ("batch", C.c_int), ("use_graph", C.c_int),
                ("adf_out", C.c_void_p), ("pacbed_out", C.c_void_p), ("exit_out", C.c_void_p)]


class msl_job(C.Structure):
    _fields_ = [("n", C.c_int), ("method", C.c_int), ("nslices", C.c_int), ("natoms", C.c_int),
                ("Z", C.c_void_p), ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("nprobe", C.c_int), ("posx", C.c_void_p), ("posy", C.c_void_p),
                ("exit_out", C.c_void_p), ("adf_out", C.c_void_p)]


class msl_serve_params(C.Structure):
    _fields_ = [("kV", C.c_float), ("ps", C.c_float), ("dz", C.c_float), ("aperture", C.c_float),
                ("adf_inner", C.c_float), ("adf_outer", C.c_float), ("potential_mode", C.c_int),
                ("budget_bytes", C.c_longlong), ("group_sizes", C.c_int)]


_vp, _i, _f, _cp = C.c_void_p, C.c_int, C.c_float, C.c_char_p
_SIGS = {
    "msl_create": ([C.POINTER(_vp), _i], _i), "msl_destroy": ([_vp], None),
    "msl_last_error": ([_vp], _cp), "msl_status_string": ([_i], _cp),
    "msl_set_compat": ([_vp, C.c_uint], _i), "msl_set_potential_mode": ([_vp, _i], _i), "msl_cancel": ([_vp], _i),
    "msl_load_atom_file": ([_vp, _cp], _i),
    "msl_set_atoms": ([_vp, _i, _vp, _vp, _vp, _vp], _i),
    "msl_upload_binned_atoms": ([_vp], _i),
    "msl_set_tds": ([_vp, C.c_uint64, _f], _i),
    "msl_upload_binned_atoms_tds": ([_vp], _i),
    "msl_upload_binned_atoms_displaced": ([_vp, _vp, _vp, _vp], _i),
    "msl_tds_displacements": ([C.c_uint64, _i, _i, _f, _vp, _vp, _vp], _i),
    "msl_get_parameterisation": ([_vp, _cp], _i), "msl_get_fparams": ([_vp, _vp], _i),
    "msl_apply_microscope_parameters": ([_vp, _f, _f, _f, _f, _f, _f, _f, _f], _i),
    "msl_set_calculation_variables": ([_vp, _f, _f, _i, _i, _f, _f, _f, _f, _f], _i),
    "msl_stem_set_calculation_variables": ([_vp, _f, _f, _i, _i, _f, _f, _f], _i),
    "msl_precalculate_frequencies": ([_vp], _i),
    "msl_initialise_wavefunctions": ([_vp], _i),
    "msl_initialise_wavefunctions_dz": ([_vp, _f], _i),
    "msl_stem_initialise_wavefunctions": ([_vp, _i, _i], _i),
    "msl_set_conventional_dz": ([_vp, _f], _i),
    "msl_multislice_step": ([_vp, _i], _i), "msl_multislice_step_conv": ([_vp, _i], _i),
    "msl_multislice_step_rs": ([_vp, _i], _i),
    "msl_get_slices": ([_vp, _i, C.POINTER(_i)], _i),
    "msl_get_exit_wave": ([_vp, _vp], _i), "msl_get_stem_exit_wave": ([_vp, _vp], _i),
    "msl_get_diffraction_display": ([_vp, _vp], _i),
    "msl_alloc_tds": ([_vp], _i), "msl_add_tds_waves": ([_vp], _i), "msl_get_tds": ([_vp, _vp], _i),
    "msl_simulate_image": ([_vp, _f], _i), "msl_get_image": ([_vp, _vp], _i),
    "msl_ctf_sweep": ([_vp, _i, _vp, _vp], _i),
    "msl_adf": ([_vp, _f, _f, C.POINTER(_f)], _i),
    "msl_get_info": ([_vp, C.POINTER(msl_info)], _i),
    "msl_stem_batched": ([_vp, C.POINTER(msl_stem_batch_desc)], _i),
    "msl_last_timings": ([_vp, _vp], _i),
    "msl_set_profiling": ([_vp, _i], _i), "msl_kernel_stats": ([_vp, _vp], _i),
    "msl_diff_create": ([C.POINTER(_vp), _i, _i, _i, _f, _f, _f, _i, _i, _i], _i),
    "msl_diff_destroy": ([_vp], None), "msl_diff_last_error": ([_vp], _cp),
    "msl_diff_forward": ([_vp, _vp, _vp, _vp, _vp], _i),
    "msl_diff_forward_only": ([_vp, _vp, _vp, _vp, _vp], _i),
    "msl_diff_backward": ([_vp, _vp, _vp, _vp], _i),
    "msl_diff_info": ([_vp, C.POINTER(_f), C.POINTER(_f), C.POINTER(_f), C.POINTER(C.c_longlong)], _i),
    "msl_diff_set_chunk": ([_vp, _i], _i),
    "msl_fft2_device": ([_i, _i, _i, _vp, _i, _vp], _i),
    "msl_set_potential_tile": ([_vp, _i], _i),
    "msl_autotune_potential": ([_vp, _f, _i, C.POINTER(_i), _vp], _i),
    "msl_fft2_device_noalloc": ([_i, _i, _i, _vp, _i, _vp], _i),
    "msl_set_pass_variant": ([_i, _i, _i], _i), "msl_get_pass_variant": ([_i, _i, C.POINTER(_i)], _i),
    "msl_bench_pass": ([_i, _i, _i, _i, _i, C.POINTER(_f), C.POINTER(_f)], _i),
    "msl_autotune_passes": ([_i, _i, _i, C.POINTER(_i), C.POINTER(_i), _vp], _i),
    "msl_autotune_batch": ([_vp, C.POINTER(msl_stem_batch_desc), _i, C.POINTER(_i), _vp, _vp, C.POINTER(_i)], _i),
    "msl_copy_device": ([_vp, _vp, C.c_longlong, _vp], _i),
    "msl_get_binned": ([_vp, _vp, _vp, _vp, _vp, _vp], _i),
    "msl_get_bins": ([_vp, _vp], _i),
    "msl_potential_slice": ([_vp, _i, _f, _i, _vp], _i),
    "msl_set_wave": ([_vp, _i, _vp], _i), "msl_get_wave": ([_vp, _i, _vp], _i),
    "msl_set_detector": ([_vp, _i, C.c_uint64], _i), "msl_get_detector_image": ([_vp, _vp], _i),
    "msl_write_tiff": ([_cp, _i, _i, _vp], _i), "msl_save_tiff": ([_vp, _i, _cp], _i),
    "msl_cs_df_sweep": ([_vp, _f, _f, _i, _f, _f, _i, _vp, _cp], _i),
    "msl_readout_noise": ([_vp, C.c_longlong, C.POINTER(_f)], _i),
    "msl_server_create": ([C.POINTER(_vp), _i], _i), "msl_server_destroy": ([_vp], None),
    "msl_server_last_error": ([_vp], _cp),
    "msl_server_run": ([_vp, C.POINTER(msl_serve_params), C.POINTER(msl_job), _i, _vp], _i),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        # torch bundles its own HIP runtime with the same soname
        # (libamdhip64.so.7); loading torch first makes libmsl bind to that
        # copy, so a process holds a single HIP/HSA runtime
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise MslError(-1, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    """Function names declared in include/msl.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(msl_[a-z0-9_]+)\s*\(", txt)))
