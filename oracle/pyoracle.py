"""ctypes wrapper around the CPU oracle (``oracle/libpsv_oracle.so``).

TEST INFRASTRUCTURE ONLY.  The oracle restates psvWave's ``fdModel`` hot path
(src/fdModel.cpp) on the CPU so that the HIP product path can be checked
against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module.

Configurations are plain dicts with the sections and keys of the reference's
INI schema (psvWave/__init__.py:34-97), so the same dict can build both an
oracle model and a product model.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpsv_oracle.so")
_lib = None


def build():
    """Compile the oracle library with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    lib.oracle_last_error.restype = C.c_char_p
    lib.oracle_create.restype = P
    lib.oracle_create.argtypes = [C.c_void_p]
    lib.oracle_copy.restype = P
    lib.oracle_copy.argtypes = [P]
    lib.oracle_destroy.argtypes = [P]
    lib.oracle_dims.argtypes = [P, C.POINTER(C.c_int)]
    lib.oracle_set_threads.argtypes = [C.c_int]
    lib.oracle_forward.argtypes = [P, C.c_int, C.c_int]
    lib.oracle_adjoint.argtypes = [P, C.c_int]
    lib.oracle_l2_misfit.restype = C.c_double
    lib.oracle_l2_misfit.argtypes = [P]
    lib.oracle_get_snapshot.argtypes = [P, C.c_int, C.c_int] + [dp] * 5
    lib.oracle_set_snapshot.argtypes = [P, C.c_int, C.c_int] + [dp] * 5
    lib.oracle_correlate_step.argtypes = [P, C.c_int, C.c_int]
    for name in ["oracle_l2_adjoint_sources", "oracle_reset_kernels",
                 "oracle_map_kernels"]:
        getattr(lib, name).argtypes = [P]
    for name in ["oracle_stress_phase", "oracle_velocity_phase"]:
        getattr(lib, name).argtypes = [P, C.c_int]
    _lib = lib
    return lib


class _Params(C.Structure):
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
    ]


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleModel:
    """CPU oracle model; arrays use the reference layout ([ix][iz])."""

    def __init__(self, cfg=None, _handle=None):
        lib = _load()
        self._lib = lib
        if _handle is not None:
            self._h = C.c_void_p(_handle)
        else:
            d, b, med = cfg["domain"], cfg["boundary"], cfg["medium"]
            s, r = cfg["sources"], cfg["receivers"]
            which = s["which_source_to_fire_in_which_shot"]
            offs = np.zeros(len(which) + 1, dtype=np.int32)
            for k, w in enumerate(which):
                offs[k + 1] = offs[k] + len(w)
            flat = np.array([int(v) for w in which for v in w] or [0],
                            dtype=np.int32)
            self._keep = [
                np.asarray(s["ix_sources"], dtype=np.int32),
                np.asarray(s["iz_sources"], dtype=np.int32),
                np.asarray(s["moment_angles"], dtype=np.float64),
                offs, flat,
                np.asarray(r["ix_receivers"], dtype=np.int32),
                np.asarray(r["iz_receivers"], dtype=np.int32),
            ]
            ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
            p = _Params(
                int(d["nt"]), int(d["nx_inner"]), int(d["nz_inner"]),
                int(d["nx_inner_boundary"]), int(d["nz_inner_boundary"]),
                float(d["dx"]), float(d["dz"]), float(d["dt"]),
                int(b["np_boundary"]), float(b["np_factor"]),
                float(med["scalar_rho"]), float(med["scalar_vp"]),
                float(med["scalar_vs"]), int(cfg["basis"]["npx"]),
                int(cfg["basis"]["npz"]), float(s["peak_frequency"]),
                float(s["source_timeshift"]),
                float(s["delay_cycles_per_shot"]), int(s["n_sources"]),
                int(s["n_shots"]), ip(self._keep[0]), ip(self._keep[1]),
                _dptr(self._keep[2]), ip(self._keep[3]), ip(self._keep[4]),
                int(r["nr"]), ip(self._keep[5]), ip(self._keep[6]),
                int(cfg["inversion"]["snapshot_interval"]),
            )
            h = lib.oracle_create(C.byref(p))
            if not h:
                raise ValueError(lib.oracle_last_error().decode())
            self._h = C.c_void_p(h)
        dims = (C.c_int * 8)()
        lib.oracle_dims(self._h, dims)
        (self.nx, self.nz, self.nt, self.snapshots, self.free_parameters,
         self.n_shots, self.nr, self.n_sources) = list(dims)
        self.misfit = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.oracle_destroy(h)
            self._h = None

    @staticmethod
    def set_threads(n):
        _load().oracle_set_threads(int(n))

    def copy(self):
        return OracleModel(_handle=self._lib.oracle_copy(self._h))

    def _grid(self, k=1):
        return [np.empty((self.nx, self.nz)) for _ in range(k)]

    def _rec(self, k=2):
        return [np.empty((self.n_shots, self.nr, self.nt)) for _ in range(k)]

    def set_parameter_fields(self, vp, vs, rho):
        vp, vs, rho = _f64(vp), _f64(vs), _f64(rho)
        self._lib.oracle_set_parameter_fields(self._h, _dptr(vp), _dptr(vs),
                                              _dptr(rho))

    def get_parameter_fields(self):
        a = self._grid(3)
        self._lib.oracle_get_parameter_fields(self._h, *map(_dptr, a))
        return tuple(a)

    def forward_simulate(self, i_shot, store_fields=True):
        if self._lib.oracle_forward(self._h, int(i_shot), int(bool(store_fields))):
            raise ValueError(self._lib.oracle_last_error().decode())

    def adjoint_simulate(self, i_shot):
        if self._lib.oracle_adjoint(self._h, int(i_shot)):
            raise ValueError(self._lib.oracle_last_error().decode())

    def calculate_l2_misfit(self):
        self.misfit = self._lib.oracle_l2_misfit(self._h)
        return self.misfit

    def calculate_l2_adjoint_sources(self):
        self._lib.oracle_l2_adjoint_sources(self._h)

    def reset_kernels(self):
        self._lib.oracle_reset_kernels(self._h)

    def map_kernels_to_velocity(self):
        self._lib.oracle_map_kernels(self._h)

    def get_kernels(self):
        a = self._grid(3)
        self._lib.oracle_get_kernels(self._h, *map(_dptr, a))
        return tuple(a)

    def get_lame_kernels(self):
        a = self._grid(3)
        self._lib.oracle_get_lame_kernels(self._h, *map(_dptr, a))
        return tuple(a)

    def set_lame_kernels(self, lambda_k, mu_k, rho_l_k):
        arrs = [_f64(a) for a in (lambda_k, mu_k, rho_l_k)]
        self._lib.oracle_set_lame_kernels(self._h, *map(_dptr, arrs))

    def get_synthetic_data(self):
        a = self._rec()
        self._lib.oracle_get_synthetic(self._h, *map(_dptr, a))
        return tuple(a)

    def set_synthetic_data(self, ux, uz):
        ux, uz = _f64(ux), _f64(uz)
        self._lib.oracle_set_synthetic(self._h, _dptr(ux), _dptr(uz))

    def get_observed_data(self):
        a = self._rec()
        self._lib.oracle_get_observed(self._h, *map(_dptr, a))
        return tuple(a)

    def set_observed_data(self, ux, uz):
        ux, uz = _f64(ux), _f64(uz)
        self._lib.oracle_set_observed(self._h, _dptr(ux), _dptr(uz))

    def get_adjoint_sources(self):
        a = self._rec()
        self._lib.oracle_get_adjoint_sources(self._h, *map(_dptr, a))
        return tuple(a)

    def get_model_vector(self):
        out = np.empty(self.free_parameters)
        self._lib.oracle_get_model_vector(self._h, _dptr(out))
        return out

    def set_model_vector(self, m):
        m = _f64(m)
        assert m.size == self.free_parameters
        self._lib.oracle_set_model_vector(self._h, _dptr(m))

    def get_gradient_vector(self):
        out = np.empty(self.free_parameters)
        self._lib.oracle_get_gradient_vector(self._h, _dptr(out))
        return out

    def get_taper(self):
        (a,) = self._grid(1)
        self._lib.oracle_get_taper(self._h, _dptr(a))
        return a

    def get_stf(self):
        a = np.empty((self.n_sources, self.nt))
        self._lib.oracle_get_stf(self._h, _dptr(a))
        return a

    def get_moment(self):
        a = np.empty((self.n_sources, 2, 2))
        self._lib.oracle_get_moment(self._h, _dptr(a))
        return a

    def get_lame(self):
        a = self._grid(4)
        self._lib.oracle_get_lame(self._h, *map(_dptr, a))
        return tuple(a)  # lm, la, mu, b

    def get_state(self):
        a = self._grid(5)
        self._lib.oracle_get_state(self._h, *map(_dptr, a))
        return tuple(a)

    def set_state(self, vx, vz, txx, tzz, txz):
        arrs = [_f64(a) for a in (vx, vz, txx, tzz, txz)]
        self._lib.oracle_set_state(self._h, *map(_dptr, arrs))

    def stress_phase(self, adjoint=False):
        self._lib.oracle_stress_phase(self._h, int(adjoint))

    def velocity_phase(self, adjoint=False):
        self._lib.oracle_velocity_phase(self._h, int(adjoint))

    def set_snapshot(self, i_shot, i_snap, vx, vz, txx, tzz, txz):
        arrs = [_f64(a) for a in (vx, vz, txx, tzz, txz)]
        if self._lib.oracle_set_snapshot(self._h, int(i_shot), int(i_snap),
                                         *map(_dptr, arrs)):
            raise ValueError("no such snapshot")

    def correlate_step(self, i_shot, it):
        """The kernel correlation of adjoint time step `it` (a multiple of
        the snapshot interval) between the current state and the stored
        snapshot of that step."""
        if self._lib.oracle_correlate_step(self._h, int(i_shot), int(it)):
            raise ValueError("no stored snapshot for this step")

    def get_snapshot(self, i_shot, i_snap):
        a = self._grid(5)
        if self._lib.oracle_get_snapshot(self._h, int(i_shot), int(i_snap),
                                         *map(_dptr, a)):
            raise ValueError("snapshot not available")
        return tuple(a)
