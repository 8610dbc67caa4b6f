"""``fdModel``: the reference's Python class, backed by the HIP library.

Method names, arguments, defaults, return shapes and exception types follow
the pybind11 class of psvWave (src/psvWave.cpp:465-655, fdModelExtended at
:85-458).  Every method calls the C-ABI of ``libpsvwave_amd.so``; all arrays
are copies in the reference layout (grids (nx, nz) with iz fastest).
"""
import ctypes as C
import warnings

import numpy as np

from . import _lib

# set once the queue-sharing warning has been given in this process
_queue_warning_issued = False


def _check_queue_sharing(lib):
    """Warn (once per process) when the library holds more HIP streams, with
    the default stream, than the runtime has hardware queues: a model's two
    lanes may then share a queue and run one after the other.  Results are
    unchanged; only the overlap of the two time loops is lost."""
    global _queue_warning_issued
    if _queue_warning_issued:
        return
    live, queues = C.c_int(), C.c_int()
    _lib.check(lib.psv_stream_pressure(C.byref(live), C.byref(queues)))
    if live.value + 1 > queues.value:
        _queue_warning_issued = True
        warnings.warn(
            f"psvwave_amd: {live.value} HIP streams of this library are live, "
            f"with the default stream more than the {queues.value} hardware "
            "queues of GPU_MAX_HW_QUEUES; a model's two lanes may share a queue "
            "and lose their overlap (results are unchanged).  Release models "
            "no longer used or raise GPU_MAX_HW_QUEUES.", RuntimeWarning,
            stacklevel=3)


def _f64(a, shape, name):
    """C-contiguous float64 copy of `a`, or ValueError if its shape is wrong
    (the check of src/psvWave.cpp:266-280, 397-407, 435-445)."""
    arr = np.asarray(a, dtype=np.float64)
    if tuple(arr.shape) != tuple(shape):
        raise ValueError(f"The input ndarray {name} does not have the right shape.")
    return np.ascontiguousarray(arr)


class Config:
    """Parsed, validated configuration (host only, no GPU needed)."""

    def __init__(self, source):
        lib = _lib.load()
        self._lib = lib
        h = C.c_void_p()
        if isinstance(source, dict):
            params, self._keep = _lib.make_params(source)
            _lib.check(lib.psv_config_from_params(C.byref(params), C.byref(h)))
        else:
            _lib.check(lib.psv_config_from_ini(str(source).encode(), C.byref(h)))
        self._h = h
        info = _lib.Info()
        _lib.check(lib.psv_config_get_info(h, C.byref(info)))
        self.info = {n: getattr(info, n) for n, _ in _lib.Info._fields_}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.psv_config_destroy(h)
            self._h = None

    def lists(self):
        """(which, (ix_src, iz_src), (ix_rec, iz_rec)); indices include the
        absorbing boundary."""
        i = self.info
        offs = np.zeros(i["n_shots"] + 1, np.int32)
        srcs = np.zeros(max(i["n_sources"], 1), np.int32)
        ixs = np.zeros(max(i["n_sources"], 1), np.int32)
        izs = np.zeros_like(ixs)
        ixr = np.zeros(max(i["nr"], 1), np.int32)
        izr = np.zeros_like(ixr)
        _lib.check(self._lib.psv_config_get_lists(
            self._h, _lib.iptr(offs), _lib.iptr(srcs), _lib.iptr(ixs),
            _lib.iptr(izs), _lib.iptr(ixr), _lib.iptr(izr)))
        which = [srcs[offs[k]:offs[k + 1]].tolist() for k in range(i["n_shots"])]
        n, r = i["n_sources"], i["nr"]
        return which, (ixs[:n], izs[:n]), (ixr[:r], izr[:r])

    def setup(self):
        """(stf (n_sources, nt), moment (n_sources, 2, 2), taper (nx, nz))."""
        i = self.info
        stf = np.zeros((i["n_sources"], i["nt"]))
        mom = np.zeros((i["n_sources"], 2, 2))
        taper = np.zeros((i["nx"], i["nz"]))
        _lib.check(self._lib.psv_config_get_setup(
            self._h, _lib.dptr(stf), _lib.dptr(mom), _lib.dptr(taper)))
        return stf, mom, taper


class fdModel:
    """fdModel(configuration_file_path: str)

    Class to simulate P-SV wave physics and its adjoint state on one MI355X.
    A configuration dict (the layout of :func:`get_dictionary`) is accepted
    in place of the path, mirroring the reference's explicit-argument C++
    constructor (src/fdModel.h:50-64).
    """

    units = [r"$m/s$", r"$m/s$", r"$kg/m^3$"]
    parameters = [r"$v_p$", r"$v_s$", r"$\rho$"]

    def __init__(self, configuration_file_path, _handle=None):
        lib = _lib.load()
        self._lib = lib
        if _handle is not None:
            self._h = _handle
        else:
            h = C.c_void_p()
            if isinstance(configuration_file_path, dict):
                params, keep = _lib.make_params(configuration_file_path)
                _lib.check(lib.psv_create(C.byref(params), C.byref(h)))
            else:
                path = str(configuration_file_path)
                _lib.check(lib.psv_create_from_ini(path.encode(), C.byref(h)))
            self._h = h
        info = _lib.Info()
        _lib.check(lib.psv_get_info(self._h, C.byref(info)))
        self._info = {n: getattr(info, n) for n, _ in _lib.Info._fields_}
        # change counters: parameters, synthetics or snapshots; observed data
        self._generation = 0
        self._obs_generation = 0
        offs = np.zeros(self.n_shots + 1, np.int32)
        srcs = np.zeros(max(self.n_sources, 1), np.int32)
        _lib.check(lib.psv_get_which(self._h, _lib.iptr(offs), _lib.iptr(srcs)))
        self._which = [srcs[offs[k]:offs[k + 1]].tolist()
                       for k in range(self.n_shots)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.psv_destroy(h)
            self._h = None

    # -- read-only attributes (src/psvWave.cpp:499-588, 621-626) ----------
    def _get(name):
        return property(lambda self: self._info[name])

    nt = _get("nt")
    nx = _get("nx")
    nz = _get("nz")
    dx = _get("dx")
    dz = _get("dz")
    dt = _get("dt")
    nx_inner = _get("nx_inner")
    nz_inner = _get("nz_inner")
    nx_inner_boundary = _get("nx_inner_boundary")
    nz_inner_boundary = _get("nz_inner_boundary")
    np_boundary = _get("np_boundary")
    nx_free_parameters = _get("nx_free_parameters")
    nz_free_parameters = _get("nz_free_parameters")
    free_parameters = _get("free_parameters")
    snapshot_interval = _get("snapshot_interval")
    snapshots = _get("snapshots")
    n_shots = _get("n_shots")
    n_sources = _get("n_sources")
    nr = _get("nr")
    del _get

    @property
    def misfit(self):
        """Current misfit, updated by calculate_l2_misfit()."""
        return self._info["misfit"]

    @property
    def which_source_to_fire_in_which_shot(self):
        return [list(w) for w in self._which]

    # -- construction ------------------------------------------------------
    def copy(self):
        """copy() -> fdModel: a deep copy, device state included."""
        h = C.c_void_p()
        _lib.check(self._lib.psv_copy(self._h, C.byref(h)))
        return fdModel(None, _handle=h)

    # -- the hot path --------------------------------------------------------
    def forward_simulate(self, i_shot, store_fields=True, verbose=False,
                         output_wavefields=False, omp_threads_override=0):
        """Run the forward simulation of shot `i_shot` (src/fdModel.cpp:526-787).

        `omp_threads_override` is accepted for signature compatibility; the
        simulation runs on the GPU and does not use host threads.  The call
        returns once the time loop is enqueued (see set_async); every other
        method waits for it.
        """
        self._generation += 1
        _lib.check(self._lib.psv_forward(self._h, int(i_shot),
                                         int(bool(store_fields)),
                                         int(bool(verbose)),
                                         int(bool(output_wavefields))))
        _check_queue_sharing(self._lib)

    def adjoint_simulate(self, i_shot, verbose=False, omp_threads_override=0):
        """Adjoint simulation of shot `i_shot`, correlating with the stored
        forward snapshots into the Lame kernels (src/fdModel.cpp:789-969).
        Returns once the time loop is enqueued (see set_async)."""
        _lib.check(self._lib.psv_adjoint(self._h, int(i_shot),
                                         int(bool(verbose))))
        _check_queue_sharing(self._lib)

    def calculate_l2_misfit(self):
        """Calculate the L2 misfit of synthetic w.r.t. observed data."""
        v = C.c_double()
        _lib.check(self._lib.psv_calculate_l2_misfit(self._h, C.byref(v)))
        self._info["misfit"] = v.value

    def calculate_l2_adjoint_sources(self):
        """Adjoint sources of the L2 misfit: synthetic minus observed."""
        _lib.check(self._lib.psv_calculate_l2_adjoint_sources(self._h))

    def reset_kernels(self):
        """Reset the kernel accumulators to zero."""
        _lib.check(self._lib.psv_reset_kernels(self._h))

    def map_kernels_to_velocity(self):
        """Map the Lame kernels (lambda, mu, rho) to (vp, vs, rho)."""
        _lib.check(self._lib.psv_map_kernels_to_velocity(self._h))

    def run_model(self, verbose=False, simulate_adjoint=True):
        """Forward all shots, misfit, and optionally adjoint sources, adjoint
        of all shots and kernel mapping (src/fdModel.cpp:1291-1308)."""
        self._generation += 1
        _lib.check(self._lib.psv_run_model(self._h, int(bool(verbose)),
                                           int(bool(simulate_adjoint))))
        _check_queue_sharing(self._lib)
        self._refresh_misfit()

    def _refresh_misfit(self):
        info = _lib.Info()
        _lib.check(self._lib.psv_get_info(self._h, C.byref(info)))
        self._info["misfit"] = info.misfit

    # -- parameters and kernels -----------------------------------------------
    def _grids(self, k):
        return [np.empty((self.nx, self.nz)) for _ in range(k)]

    def get_parameter_fields(self):
        """(vp, vs, rho), each (nx, nz)."""
        a = self._grids(3)
        _lib.check(self._lib.psv_get_parameter_fields(self._h, *map(_lib.dptr, a)))
        return tuple(a)

    def set_parameter_fields(self, vp, vs, rho):
        """Set (vp, vs, rho), each (nx, nz), and update the Lame fields."""
        self._generation += 1
        shape = (self.nx, self.nz)
        a = [_f64(vp, shape, "_vp"), _f64(vs, shape, "_vs"),
             _f64(rho, shape, "_rho")]
        _lib.check(self._lib.psv_set_parameter_fields(self._h, *map(_lib.dptr, a)))

    def get_kernels(self):
        """(vp_kernel, vs_kernel, rho_kernel), each (nx, nz)."""
        a = self._grids(3)
        _lib.check(self._lib.psv_get_kernels(self._h, *map(_lib.dptr, a)))
        return tuple(a)

    def get_lame_kernels(self):
        """(lambda_kernel, mu_kernel, density_l_kernel), each (nx, nz)."""
        a = self._grids(3)
        _lib.check(self._lib.psv_get_lame_kernels(self._h, *map(_lib.dptr, a)))
        return tuple(a)

    def get_model_vector(self):
        """Model vector (free_parameters,): cell averages of vp, vs, rho."""
        v = np.empty(self.free_parameters)
        _lib.check(self._lib.psv_get_model_vector(self._h, _lib.dptr(v), v.size))
        return v

    def set_model_vector(self, m):
        """Update vp, vs, rho of the free region from a model vector."""
        self._generation += 1
        v = np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(-1))
        _lib.check(self._lib.psv_set_model_vector(self._h, _lib.dptr(v), v.size))

    def get_gradient_vector(self):
        """Gradient vector (free_parameters,): cell sums of the kernels."""
        v = np.empty(self.free_parameters)
        _lib.check(self._lib.psv_get_gradient_vector(self._h, _lib.dptr(v), v.size))
        return v

    def load_vector(self, relative_path, verbose):
        """Load a vector of 3 * nx_free * nz_free values from a text file."""
        n = 3 * self.nx_free_parameters * self.nz_free_parameters
        v = np.empty(n)
        _lib.check(self._lib.psv_load_vector(self._h, str(relative_path).encode(),
                                             int(bool(verbose)), _lib.dptr(v), n))
        return v

    # -- data ------------------------------------------------------------------
    def _rec(self):
        return [np.empty((self.n_shots, self.nr, self.nt)) for _ in range(2)]

    def get_synthetic_data(self):
        """(ux, uz), each (n_shots, nr, nt)."""
        a = self._rec()
        _lib.check(self._lib.psv_get_synthetic_data(self._h, *map(_lib.dptr, a)))
        return tuple(a)

    def set_synthetic_data(self, ux, uz):
        self._generation += 1
        shape = (self.n_shots, self.nr, self.nt)
        a = [_f64(ux, shape, "ux"), _f64(uz, shape, "uz")]
        _lib.check(self._lib.psv_set_synthetic_data(self._h, *map(_lib.dptr, a)))

    def get_observed_data(self):
        """(ux, uz), each (n_shots, nr, nt)."""
        a = self._rec()
        _lib.check(self._lib.psv_get_observed_data(self._h, *map(_lib.dptr, a)))
        return tuple(a)

    def set_observed_data(self, ux, uz):
        self._obs_generation += 1
        shape = (self.n_shots, self.nr, self.nt)
        a = [_f64(ux, shape, "ux"), _f64(uz, shape, "uz")]
        _lib.check(self._lib.psv_set_observed_data(self._h, *map(_lib.dptr, a)))

    def get_adjoint_sources(self):
        """(a_ux, a_uz), each (n_shots, nr, nt)."""
        a = self._rec()
        _lib.check(self._lib.psv_get_adjoint_sources(self._h, *map(_lib.dptr, a)))
        return tuple(a)

    def get_snapshots(self):
        """Snapshots of (vx, vz, txx, tzz, txz), each
        (n_shots, snapshots, nx, nz)."""
        shape = (self.n_shots, self.snapshots, self.nx, self.nz)
        a = [np.empty(shape) for _ in range(5)]
        _lib.check(self._lib.psv_get_snapshots(self._h, *map(_lib.dptr, a)))
        return tuple(a)

    def get_snapshot(self, i_shot, i_snapshot):
        """One snapshot of (vx, vz, txx, tzz, txz), each (nx, nz)."""
        a = self._grids(5)
        _lib.check(self._lib.psv_get_snapshot(self._h, int(i_shot),
                                              int(i_snapshot), *map(_lib.dptr, a)))
        return tuple(a)

    # -- geometry (src/psvWave.cpp:99-110, 207-246, 324-387) ----------------
    def get_extent(self, include_absorbing_boundary=True):
        if include_absorbing_boundary:
            return (-self.np_boundary, self.dx * (self.nx_inner + self.np_boundary),
                    -self.np_boundary, self.dz * (self.nz_inner + self.np_boundary))
        return (0, self.dx * self.nx_inner, 0, self.dz * self.nz_inner)

    def get_coordinates(self, in_units):
        ix = np.arange(self.nx, dtype=np.float64)[:, None] * np.ones((1, self.nz))
        iz = np.ones((self.nx, 1)) * np.arange(self.nz, dtype=np.float64)[None, :]
        if in_units:
            return (ix - self.np_boundary) * self.dx, (iz - self.np_boundary) * self.dz
        return ix, iz

    def _positions(self, fn, n, in_units, include_abs):
        ix = np.zeros(max(n, 1), np.int32)
        iz = np.zeros(max(n, 1), np.int32)
        _lib.check(fn(self._h, _lib.iptr(ix), _lib.iptr(iz)))
        x = ix[:n].astype(np.float64)
        z = iz[:n].astype(np.float64)
        if not include_abs or in_units:
            x -= self.np_boundary
            z -= self.np_boundary
            if in_units:
                x *= self.dx
                z *= self.dz
        return x, z

    def get_sources(self, in_units=True, include_absorbing_boundary_as_index=True):
        return self._positions(self._lib.psv_get_source_indices, self.n_sources,
                               in_units, include_absorbing_boundary_as_index)

    def get_receivers(self, in_units=True, include_absorbing_boundary_as_index=True):
        return self._positions(self._lib.psv_get_receiver_indices, self.nr,
                               in_units, include_absorbing_boundary_as_index)

    # -- text formats (src/fdModel.cpp:971-1126, 1185-1289, 1326-1356) --------
    def write_receivers(self, prefix=""):
        _lib.check(self._lib.psv_write_receivers(self._h, str(prefix).encode()))

    def write_sources(self):
        _lib.check(self._lib.psv_write_sources(self._h))

    def write_kernels(self):
        _lib.check(self._lib.psv_write_kernels(self._h))

    def load_receivers(self, verbose=False):
        self._obs_generation += 1
        _lib.check(self._lib.psv_load_receivers(self._h, int(bool(verbose))))

    def load_model(self, de_path, vp_path, vs_path, verbose=False):
        self._generation += 1
        _lib.check(self._lib.psv_load_model(self._h, str(de_path).encode(),
                                            str(vp_path).encode(),
                                            str(vs_path).encode(),
                                            int(bool(verbose))))

    # -- setup products, tuning and measurement --------------------------------
    def get_setup(self):
        """(stf (n_sources, nt), moment (n_sources, 2, 2), taper (nx, nz))."""
        stf = np.zeros((self.n_sources, self.nt))
        mom = np.zeros((self.n_sources, 2, 2))
        taper = np.zeros((self.nx, self.nz))
        _lib.check(self._lib.psv_get_setup(self._h, _lib.dptr(stf),
                                           _lib.dptr(mom), _lib.dptr(taper)))
        return stf, mom, taper

    def set_step_variant(self, variant, xcd_remap=True):
        """Step kernel (bit-identical results): 0 fused 64x8 tiles, 1 fused
        64x16, 2 split (two launches per step); 3-5 the same with
        reciprocal-multiply division by the grid spacings; 6, 7 two steps per
        launch; 8-13 lean kernels with 64 x 16, 8, 12, 20, 24, 32 tiles; 14
        lean two steps per launch; 15-17 lean strips of 2-4
        tiles per block; 18 lean 64x4; 19, 20 lean 64x8 / 64x12 at 6 / 5
        waves per SIMD; 21 lean 64x8 with halo columns by DMA; 22, 23 lean
        64x8 / 64x4 with separate stress stages; 24 lean 64x8
        with halo columns and ring operands by DMA (6 waves per SIMD); 25
        the same without the halo columns; 26 lean 64x16 with ring operands
        by DMA (4 waves per SIMD); 27 (the default), 28 variants 24, 26 with
        their loads capped to the grid (edge tiles read no padding)."""
        _lib.check(self._lib.psv_set_step_variant(self._h, int(variant),
                                                  int(bool(xcd_remap))))

    def set_graph_mode(self, on=True):
        """Replay each shot's time loop as a captured HIP graph."""
        _lib.check(self._lib.psv_set_graph_mode(self._h, int(bool(on))))

    def set_async(self, on):
        """on (the default): forward_simulate and adjoint_simulate return as
        soon as their time loop is enqueued on the GPU -- forwards of
        consecutive calls run two at a time, adjoints in pairs with the
        second one's kernel terms added afterwards in its own order -- and
        every other method waits for them first; results are bitwise those
        of synchronous calls.  off: each call returns when its loop has
        finished, as in the reference."""
        _lib.check(self._lib.psv_set_async(self._h, int(bool(on))))

    def synchronize(self):
        """Wait for the time loops enqueued by forward_simulate and
        adjoint_simulate."""
        _lib.check(self._lib.psv_synchronize(self._h))

    def set_checkpointing(self, every):
        """Keep every `every`-th forward snapshot only and recompute the
        rest during adjoint_simulate (bit-identical kernels, 1/every of the
        snapshot memory, one extra forward per adjoint)."""
        self._generation += 1
        _lib.check(self._lib.psv_set_checkpointing(self._h, int(every)))

    def set_timing(self, stride):
        """Step timing: stride k > 0 records the kernel start and end of every
        k-th step launch; k < 0 brackets whole time loops with HIP events
        (launches plus gaps); 0 turns timing off."""
        _lib.check(self._lib.psv_set_timing(self._h, int(stride)))

    def step_timing(self):
        """(total_ms, timed_launches, launches) since the last reset."""
        ms = C.c_double()
        timed = C.c_int64()
        launches = C.c_int64()
        _lib.check(self._lib.psv_get_step_timing(self._h, C.byref(ms),
                                                 C.byref(timed), C.byref(launches)))
        return ms.value, timed.value, launches.value

    def reset_step_timing(self):
        _lib.check(self._lib.psv_reset_step_timing(self._h))

    def device_bytes(self):
        b = C.c_int64()
        _lib.check(self._lib.psv_device_bytes(self._h, C.byref(b)))
        return b.value
