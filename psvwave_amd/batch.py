"""``Batch``: many independent fdModels of uneven size advanced together.

The reference runs one ``fdModel`` at a time (its stated purpose is
Monte-Carlo simulation, README.md:16-17).  A ``Batch`` holds n members that
share ``nt`` and ``n_shots``; ``forward(i_shot)`` advances every member with
one kernel launch per time step over the tiles of all members.  Each
member's results equal those of ``fdModel(config)`` run alone, bit for bit.
"""
import ctypes as C

import numpy as np

from . import _lib
from .model import _f64

# step variants Batch.autotune times, in the library's order
AUTOTUNE_VARIANTS = (0, 1, 6, 7, 8, 9, 10, 11, 12, 13, 18, 19, 20, 22, 23, 24, 26, 27,
                     28)


class Batch:
    def __init__(self, configs):
        lib = _lib.load()
        self._lib = lib
        built = [_lib.make_params(c) for c in configs]
        self._keep = [k for _, k in built]
        arr = (_lib.Params * len(built))(*[p for p, _ in built])
        h = C.c_void_p()
        _lib.check(lib.psv_batch_create(len(built), arr, C.byref(h)))
        self._h = h
        self.infos = []
        for m in range(len(built)):
            info = _lib.Info()
            _lib.check(lib.psv_batch_get_info(h, m, C.byref(info)))
            self.infos.append({n: getattr(info, n) for n, _ in _lib.Info._fields_})

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.psv_batch_destroy(h)
            self._h = None

    def __len__(self):
        return len(self.infos)

    @property
    def nt(self):
        return self.infos[0]["nt"]

    @property
    def n_shots(self):
        return self.infos[0]["n_shots"]

    def set_parameter_fields(self, m, vp, vs, rho):
        i = self.infos[m]
        shape = (i["nx"], i["nz"])
        a = [_f64(vp, shape, "_vp"), _f64(vs, shape, "_vs"), _f64(rho, shape, "_rho")]
        _lib.check(self._lib.psv_batch_set_parameter_fields(self._h, int(m),
                                                            *map(_lib.dptr, a)))

    def set_observed_data(self, m, ux, uz):
        i = self.infos[m]
        shape = (i["n_shots"], i["nr"], i["nt"])
        a = [_f64(ux, shape, "ux"), _f64(uz, shape, "uz")]
        _lib.check(self._lib.psv_batch_set_observed_data(self._h, int(m),
                                                         *map(_lib.dptr, a)))

    def get_synthetic_data(self, m):
        i = self.infos[m]
        a = [np.empty((i["n_shots"], i["nr"], i["nt"])) for _ in range(2)]
        _lib.check(self._lib.psv_batch_get_synthetic_data(self._h, int(m),
                                                          *map(_lib.dptr, a)))
        return tuple(a)

    def forward(self, i_shot):
        """Forward simulation of shot `i_shot` for every member."""
        _lib.check(self._lib.psv_batch_forward(self._h, int(i_shot)))

    def forward_all(self):
        """Forward simulation of every shot for every member: bitwise what
        forward(0), forward(1), ... give.  With 2 to 8 shots and a lean step
        variant the shots advance in the same launches (each tile's shots
        side by side, sharing their material reads), shot s in state buffer
        set s of each member; in graph mode the whole loop is captured once
        as one HIP graph and replayed."""
        _lib.check(self._lib.psv_batch_forward_all(self._h))

    def lockstep(self):
        """Whether forward_all advances the shots together with the current
        variant and modes (allocates the extra state sets if needed)."""
        on = C.c_int()
        _lib.check(self._lib.psv_batch_lockstep(self._h, C.byref(on)))
        return bool(on.value)

    def misfits(self):
        """L2 misfit of every member (calculate_l2_misfit of each)."""
        out = np.empty(len(self))
        _lib.check(self._lib.psv_batch_misfits(self._h, _lib.dptr(out)))
        return out

    def gradients(self):
        """FWI gradient of every member against its observed data (each
        member's run_model: forward with snapshots, misfit, adjoint sources,
        kernel reset, adjoint with correlation, kernel mapping), shot by
        shot with one launch per time step over all members.  Returns the
        misfits; the velocity kernels come from get_kernels(m)."""
        _lib.check(self._lib.psv_batch_gradients(self._h))
        out = np.empty(len(self))
        _lib.check(self._lib.psv_batch_misfits(self._h, _lib.dptr(out)))
        return out

    def get_kernels(self, m):
        """Velocity kernels (vp, vs, rho) of member m after gradients()."""
        i = self.infos[m]
        a = [np.empty((i["nx"], i["nz"])) for _ in range(3)]
        _lib.check(self._lib.psv_batch_get_kernels(self._h, int(m), *map(_lib.dptr, a)))
        return tuple(a)

    def set_variant(self, variant, xcd_remap=True):
        _lib.check(self._lib.psv_batch_set_variant(self._h, int(variant),
                                                   int(bool(xcd_remap))))

    def set_persistent(self, on=True):
        """One cooperative launch per shot's forward loop (blocks stay
        resident, a grid barrier between steps) instead of one launch per
        step; step variant 9 only."""
        _lib.check(self._lib.psv_batch_set_persistent(self._h, int(bool(on))))

    def set_graph_mode(self, on=True):
        _lib.check(self._lib.psv_batch_set_graph_mode(self._h, int(bool(on))))

    def autotune(self, steps=20):
        """Time variants 0, 1, 6, 7, 8-13, 18-20, 22-24, 26-28 over `steps` steps,
        each in the loop forward_all runs with it (all shots in lockstep
        where that applies), and keep the fastest.  Returns (best variant,
        {variant: ms per time step of one shot})."""
        best = C.c_int()
        ms = np.zeros(len(AUTOTUNE_VARIANTS))
        _lib.check(self._lib.psv_batch_autotune(self._h, int(steps), C.byref(best),
                                                _lib.dptr(ms)))
        return best.value, {v: t for v, t in zip(AUTOTUNE_VARIANTS, ms.tolist())
                            if t >= 0}

    def timing(self, reset=False):
        """(summed forward-loop milliseconds, grid points of all members)."""
        ms = C.c_double()
        pts = C.c_int64()
        _lib.check(self._lib.psv_batch_timing(self._h, C.byref(ms), C.byref(pts),
                                              int(bool(reset))))
        return ms.value, pts.value
