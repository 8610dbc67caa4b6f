"""The benchmark workloads are generated, not stored: these pins fix the
generator's output (SHA-256 of the float64 arrays) and its stated
properties, so a changed generator cannot silently change the benchmark."""
import hashlib

import numpy as np

from psvwave_amd.synthetic import (WATER_ROWS, layered_faulted_config,
                                   layered_faulted_model, monte_carlo_configs,
                                   monte_carlo_model)


def _sha(arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype="<f8").tobytes())
    return h.hexdigest()


def test_layered_faulted_model_is_pinned():
    cfg = layered_faulted_config()
    start = layered_faulted_model(cfg, seed=2, faulted=True)
    true = layered_faulted_model(cfg, seed=2, faulted=False)
    assert start[0].shape == (1800, 450)
    assert _sha(start) == "905e34d7dab3ae74f2e06daaaeea5fc22f4d2a1a6bddeecd72e88bcbaf81d236"
    assert _sha(true) == "08208e50f2028f1528569481d68e7a35bda86aedadb96c20e2914b9d159b1a92"
    vp, vs, rho = start
    npb = cfg["boundary"]["np_boundary"]
    water = slice(npb, npb + WATER_ROWS)
    assert np.all(vp[:, water] == 1500.0) and np.all(vs[:, water] == 0.0)
    below = vp[npb:-npb, npb + WATER_ROWS:-npb]
    assert below.min() >= 1500.0 and below.max() <= 5500.0
    assert np.array_equal(rho, 310.0 * vp ** 0.25)
    # the absorbing zone repeats the edge values
    assert np.array_equal(vp[0], vp[npb]) and np.array_equal(vp[:, -1], vp[:, -npb - 1])


def test_monte_carlo_batch_is_pinned():
    cfgs = monte_carlo_configs()
    assert len(cfgs) == 512
    npts = sum((c["domain"]["nx_inner"] + 60) * (c["domain"]["nz_inner"] + 60) for c in cfgs)
    assert npts == 87067595
    for c in cfgs:
        assert 200 <= c["domain"]["nx_inner"] <= 800
        assert 100 <= c["domain"]["nz_inner"] <= 400
    rng = np.random.default_rng(3)
    arrays = []
    for c in cfgs[:8]:
        vp, vs, rho = monte_carlo_model(c, rng)
        assert vp.min() >= 1800.0 and vp.max() <= 3000.0
        assert np.all(vs >= 0.45 * vp - 1e-9) and np.all(vs <= 0.55 * vp + 1e-9)
        assert rho.min() >= 1400.0 and rho.max() <= 2200.0
        arrays += [vp, vs, rho]
    assert _sha(arrays) == "18617b9134806070cd67fc738b6949562130c5b0c144cf7fd994c41cd7242d07"
