"""ctypes bindings of ``libpsvwave_amd.so`` (declared in include/psvwave.h).

The library is the only implementation: if it is missing this module raises
ImportError instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsvwave_amd.so")

PSV_OK, PSV_EVALUE, PSV_ERUNTIME = 0, 1, 2


class Params(C.Structure):
    """psv_params: the reference's explicit-argument constructor values."""
    _fields_ = [
        ("nt", C.c_int), ("nx_inner", C.c_int), ("nz_inner", C.c_int),
        ("nx_inner_boundary", C.c_int), ("nz_inner_boundary", C.c_int),
        ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("np_boundary", C.c_int), ("np_factor", C.c_double),
        ("scalar_rho", C.c_double), ("scalar_vp", C.c_double),
        ("scalar_vs", C.c_double), ("npx", C.c_int), ("npz", C.c_int),
        ("peak_frequency", C.c_double), ("source_timeshift", C.c_double),
        ("delay_cycles_per_shot", C.c_double), ("n_sources", C.c_int),
        ("n_shots", C.c_int), ("ix_sources", C.POINTER(C.c_int)),
        ("iz_sources", C.POINTER(C.c_int)),
        ("moment_angles", C.POINTER(C.c_double)),
        ("which_offsets", C.POINTER(C.c_int)),
        ("which_sources", C.POINTER(C.c_int)), ("nr", C.c_int),
        ("ix_receivers", C.POINTER(C.c_int)),
        ("iz_receivers", C.POINTER(C.c_int)),
        ("snapshot_interval", C.c_int),
        ("observed_data_folder", C.c_char_p), ("stf_folder", C.c_char_p),
    ]


class Info(C.Structure):
    """psv_info: the read-only attributes of a model."""
    _fields_ = [(n, C.c_int) for n in (
        "nt", "nx", "nz", "nx_inner", "nz_inner", "nx_inner_boundary",
        "nz_inner_boundary", "np_boundary", "nx_free_parameters",
        "nz_free_parameters", "free_parameters", "snapshot_interval",
        "snapshots", "n_shots", "n_sources", "nr", "npx", "npz")] + [
        (n, C.c_double) for n in (
            "dx", "dz", "dt", "np_factor", "peak_frequency",
            "source_timeshift", "delay_cycles_per_shot", "scalar_rho",
            "scalar_vp", "scalar_vs", "misfit")]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_S = C.c_char_p
_L = C.c_int64
_i = C.c_int

# name -> argtypes (all return int unless listed in _RESTYPES)
_SIGNATURES = {
    "psv_last_error": [],
    "psv_config_from_ini": [_S, C.POINTER(_P)],
    "psv_config_from_params": [C.POINTER(Params), C.POINTER(_P)],
    "psv_config_destroy": [_P],
    "psv_config_get_info": [_P, C.POINTER(Info)],
    "psv_config_get_lists": [_P, _I, _I, _I, _I, _I, _I],
    "psv_config_get_setup": [_P, _D, _D, _D],
    "psv_create_from_config": [_P, C.POINTER(_P)],
    "psv_create_from_ini": [_S, C.POINTER(_P)],
    "psv_create": [C.POINTER(Params), C.POINTER(_P)],
    "psv_copy": [_P, C.POINTER(_P)],
    "psv_destroy": [_P],
    "psv_get_info": [_P, C.POINTER(Info)],
    "psv_get_which": [_P, _I, _I],
    "psv_get_source_indices": [_P, _I, _I],
    "psv_get_receiver_indices": [_P, _I, _I],
    "psv_get_folders": [_P, C.c_char_p, C.c_char_p, _i],
    "psv_forward": [_P, _i, _i, _i, _i],
    "psv_adjoint": [_P, _i, _i],
    "psv_calculate_l2_misfit": [_P, _D],
    "psv_calculate_l2_adjoint_sources": [_P],
    "psv_reset_kernels": [_P],
    "psv_map_kernels_to_velocity": [_P],
    "psv_run_model": [_P, _i, _i],
    "psv_get_parameter_fields": [_P, _D, _D, _D],
    "psv_set_parameter_fields": [_P, _D, _D, _D],
    "psv_get_kernels": [_P, _D, _D, _D],
    "psv_get_lame_kernels": [_P, _D, _D, _D],
    "psv_get_synthetic_data": [_P, _D, _D],
    "psv_set_synthetic_data": [_P, _D, _D],
    "psv_get_observed_data": [_P, _D, _D],
    "psv_set_observed_data": [_P, _D, _D],
    "psv_get_adjoint_sources": [_P, _D, _D],
    "psv_get_snapshots": [_P, _D, _D, _D, _D, _D],
    "psv_get_snapshot": [_P, _i, _i, _D, _D, _D, _D, _D],
    "psv_get_model_vector": [_P, _D, _L],
    "psv_set_model_vector": [_P, _D, _L],
    "psv_get_gradient_vector": [_P, _D, _L],
    "psv_load_vector": [_P, _S, _i, _D, _L],
    "psv_get_setup": [_P, _D, _D, _D],
    "psv_write_receivers": [_P, _S],
    "psv_write_sources": [_P],
    "psv_write_kernels": [_P],
    "psv_load_receivers": [_P, _i],
    "psv_load_model": [_P, _S, _S, _S, _i],
    "psv_set_step_variant": [_P, _i, _i],
    "psv_set_timing": [_P, _i],
    "psv_set_graph_mode": [_P, _i],
    "psv_set_checkpointing": [_P, _i],
    "psv_set_async": [_P, _i],
    "psv_synchronize": [_P],
    "psv_get_step_timing": [_P, _D, C.POINTER(C.c_int64),
                            C.POINTER(C.c_int64)],
    "psv_reset_step_timing": [_P],
    "psv_device_bytes": [_P, C.POINTER(C.c_int64)],
    "psv_calibrate_copy": [C.c_int64, _i, _i, _D],
    "psv_stream_pressure": [_I, _I],
    "psv_l2_misfit_noalloc": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                              C.c_double, C.c_void_p, C.c_void_p],
    "psv_batch_create": [_i, C.POINTER(Params), C.POINTER(_P)],
    "psv_batch_destroy": [_P],
    "psv_batch_size": [_P, C.POINTER(C.c_int)],
    "psv_batch_get_info": [_P, _i, C.POINTER(Info)],
    "psv_batch_set_parameter_fields": [_P, _i, _D, _D, _D],
    "psv_batch_set_observed_data": [_P, _i, _D, _D],
    "psv_batch_get_synthetic_data": [_P, _i, _D, _D],
    "psv_batch_forward": [_P, _i],
    "psv_batch_forward_all": [_P],
    "psv_batch_lockstep": [_P, C.POINTER(_i)],
    "psv_batch_misfits": [_P, _D],
    "psv_batch_gradients": [_P],
    "psv_batch_set_persistent": [_P, _i],
    "psv_batch_get_kernels": [_P, _i, _D, _D, _D],
    "psv_batch_set_variant": [_P, _i, _i],
    "psv_batch_set_graph_mode": [_P, _i],
    "psv_batch_autotune": [_P, _i, C.POINTER(C.c_int), _D],
    "psv_batch_timing": [_P, _D, C.POINTER(C.c_int64), _i],
    "psv_conv2d_bf16": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _i, _i, _i, _i,
                        _i, _i, _i, _i, _i, _i, C.c_void_p],
    "psv_maxpool2_bf16": [C.c_void_p, C.c_void_p, _i, _i, _i, C.c_void_p],
    "psv_conv_bf16": [C.c_void_p, _i, C.c_void_p, C.c_void_p, C.c_void_p, _i, _i, _i,
                      C.c_void_p, C.c_void_p, _i, _i, _i, _i, _i, _i, _i, _i, _i, C.c_void_p],
    "psv_conv_wgrad_bf16": [C.c_void_p, _i, _i, _i, C.c_void_p, _i, _i, _i, _i, _i, _i,
                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p],
    "psv_channel_sum_bf16": [C.c_void_p, _i, _i, _i, C.c_int64, C.c_void_p, C.c_void_p,
                             C.c_int64, C.c_void_p],
    "psv_maxpool2_backward_bf16": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _i, _i,
                                   C.c_void_p, _i, _i, _i, _i, C.c_void_p],
}
_RESTYPES = {"psv_last_error": C.c_char_p, "psv_config_destroy": None,
             "psv_destroy": None, "psv_batch_destroy": None}

_lib = None


def load():
    """Load the HIP library; ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m psvwave_amd.build`"
            " (hipcc, gfx950). There is no non-HIP implementation.")
    # One HIP runtime per process: torch ships its own copy of the HIP and
    # HSA runtimes.  Loaded first, it is the one the library binds to (same
    # soname); loaded after the library's /opt/rocm copy had initialised the
    # GPU, torch would find no GPU.  So torch, when installed, comes first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def check(status):
    """Raise the Python exception matching a C-ABI status."""
    if status == PSV_OK:
        return
    msg = load().psv_last_error().decode(errors="replace")
    if status == PSV_EVALUE:
        raise ValueError(msg)
    raise RuntimeError(msg)


def dptr(a):
    return a.ctypes.data_as(_D)


def iptr(a):
    return a.ctypes.data_as(_I)


def make_params(cfg):
    """psv_params from a configuration dict (psvWave/__init__.py:34-97
    layout).  Returns (params, keepalive) -- keep the second alive while the
    first is in use."""
    d, b, med = cfg["domain"], cfg["boundary"], cfg["medium"]
    s, r = cfg["sources"], cfg["receivers"]
    which = [list(w) for w in s["which_source_to_fire_in_which_shot"]]
    # The C-ABI reads exactly the declared counts, so list lengths are
    # checked here with the reference's messages (src/fdModel.cpp:414-452).
    n_src = int(s["n_sources"])
    if not (len(s["ix_sources"]) == len(s["iz_sources"]) ==
            len(s["moment_angles"]) == n_src):
        raise ValueError(
            "Dimension mismatch between n_sources and sources.ix_sources, "
            "sources.iz_sources or sources.moment_angles")
    if len(which) != int(s["n_shots"]):
        raise ValueError("Mismatch between n_shots and "
                         "sources.which_source_to_fire_in_which_shot")
    if sum(len(w) for w in which) != n_src:
        raise ValueError("Mismatch between n_sources and "
                         "sources.which_source_to_fire_in_which_shot.")
    if not (len(r["ix_receivers"]) == len(r["iz_receivers"]) == int(r["nr"])):
        raise ValueError("Mismatch between nr and receivers.ix_receivers or "
                         "receivers.iz_receivers")
    offs = np.zeros(len(which) + 1, dtype=np.int32)
    for k, w in enumerate(which):
        offs[k + 1] = offs[k] + len(w)
    keep = [
        np.ascontiguousarray(s["ix_sources"], dtype=np.int32).reshape(-1),
        np.ascontiguousarray(s["iz_sources"], dtype=np.int32).reshape(-1),
        np.ascontiguousarray(s["moment_angles"], dtype=np.float64).reshape(-1),
        offs,
        np.array([int(v) for w in which for v in w] or [0], dtype=np.int32),
        np.ascontiguousarray(r["ix_receivers"], dtype=np.int32).reshape(-1),
        np.ascontiguousarray(r["iz_receivers"], dtype=np.int32).reshape(-1),
    ]
    out = cfg.get("output", {})
    folders = [str(out.get("observed_data_folder", ".")).encode(),
               str(out.get("stf_folder", ".")).encode()]
    keep += folders
    p = Params(
        int(d["nt"]), int(d["nx_inner"]), int(d["nz_inner"]),
        int(d["nx_inner_boundary"]), int(d["nz_inner_boundary"]),
        float(d["dx"]), float(d["dz"]), float(d["dt"]),
        int(b["np_boundary"]), float(b["np_factor"]),
        float(med["scalar_rho"]), float(med["scalar_vp"]),
        float(med["scalar_vs"]), int(cfg["basis"]["npx"]),
        int(cfg["basis"]["npz"]), float(s["peak_frequency"]),
        float(s["source_timeshift"]), float(s["delay_cycles_per_shot"]),
        int(s["n_sources"]), int(s["n_shots"]), iptr(keep[0]),
        iptr(keep[1]), dptr(keep[2]), iptr(keep[3]), iptr(keep[4]),
        int(r["nr"]), iptr(keep[5]), iptr(keep[6]),
        int(cfg["inversion"]["snapshot_interval"]), folders[0], folders[1],
    )
    return p, keep
