"""Failure isolation in the batched and served forms.  A box that fails
numerically is flagged by the kernels (the health bits of sph_common.hpp:
non-finite state or rates, the vdW pole rho b >= 1, a neighbour list over
capacity), evicted by the serving loop with its flag, and the boxes beside it
in the same launches are bitwise what they are alone.  The reference's
drivers stop a run on NaN positions (toybox.py:104-106) and its tests stop
on the first non-finite state (test/pbc_vtest.py:65-77); here one job's
failure must not touch the others.  A box too dense for the arena's lists is
refused when it is added.  Every failure is a flagged numerical condition
the kernels handle (no GPU fault): non-finite coordinates map to clamped
cells, and list appends past the capacity are dropped and flagged."""
import math

import numpy as np
import pytest

from pyticles_amd import _lib, configs
from pyticles_amd.arena import Arena
from pyticles_amd.device import DeviceBox
from pyticles_amd.serve import Server

pytestmark = pytest.mark.gpu

NONFINITE, EOS_POLE, SKIN, LISTCAP = 1, 2, 4, 8


def alone(sysm, precision):
    prm = dict(sysm.prm, precision=precision)
    with DeviceBox(sysm.n, prm) as b:
        b.upload(r=sysm.r, v=sysm.v, m=sysm.m, u=np.zeros(sysm.n))
        b.init_thermal(sysm.T)
        b.step(sysm.steps)
        return b.download("r", "v", "u", "rho")


def _good(k):
    return [configs.lognormal_box(700 + j, 300, 3000, 0.4, 1.0, steps=25 + 5 * j) for j in range(k)]


def _nan_box():
    s = configs.lognormal_box(801, 400, 900, 0.5, 0.8, steps=30)
    v = s.v.copy()
    v[s.n // 2] = (math.nan, 0.0)  # one particle's velocity
    return configs.System(s.r, v, s.m, s.T, s.prm, s.steps, s.seed)


def _pole_box():
    s = configs.lognormal_box(802, 400, 900, 0.5, 0.8, steps=30)
    return configs.System(s.r, s.v, s.m, s.T, dict(s.prm, b=1.0), s.steps, s.seed)  # rho b ~ 1.02


def _clump_box():
    # mean density 1 (admitted), but every particle inside a 6 x 6 corner at
    # spacing 0.3: ~640 neighbours within the list radius against 192 entries
    L, k = 20.0, 20
    g = (np.arange(k) + 0.5) * 0.3
    r = np.stack(np.meshgrid(g, g, indexing="ij"), axis=-1).reshape(-1, 2)
    n = r.shape[0]
    rng = np.random.Generator(np.random.PCG64(803))
    v = configs.velocities(n, 0.5, rng)
    return configs.System(np.ascontiguousarray(r), v, np.ones(n), 0.5, configs.params(L), 20, 803)


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_failing_boxes_evicted_neighbours_unchanged(precision):
    good = _good(5)
    bad = {"nan": _nan_box(), "pole": _pole_box(), "clump": _clump_box()}
    expect = {"nan": NONFINITE, "pole": EOS_POLE, "clump": LISTCAP}
    # bad boxes interleaved with good ones, so they share waves and launches
    jobs = [good[0], bad["nan"], good[1], bad["pole"], good[2], bad["clump"], good[3], good[4]]
    names = ["g0", "nan", "g1", "pole", "g2", "clump", "g3", "g4"]
    arr = np.zeros(len(jobs))
    srv = Server(2e9, precision=precision, max_resident=50000, max_job=4000, chunk=2, keep_results=True)
    try:
        recs = srv.run(jobs, arr)
    finally:
        srv.close()
    assert srv.max_running == len(jobs)  # all in the same launches
    for rec in recs:
        nm = names[rec.job_id]
        print("[failure] %s health %d" % (nm, rec.health), flush=True)
        if nm in expect:
            assert rec.health & expect[nm], (nm, rec.health)
        else:
            assert rec.health == 0, (nm, rec.health)
            ref = alone(jobs[rec.job_id], precision)
            for k in ("r", "v", "u", "rho"):
                assert np.array_equal(rec.result[k], ref[k]), (nm, k)


def test_box_too_dense_for_the_lists_is_refused():
    # three times the unit density: ~174 mean neighbours within the list
    # radius, over half of the arena's 192 entries
    s = configs.lognormal_box(804, 900, 1200, 0.5, 0.8, steps=10)
    dense = configs.System(s.r / math.sqrt(3.0), s.v, s.m, s.T,
                           dict(s.prm, Lx=s.prm["Lx"] / math.sqrt(3.0), Ly=s.prm["Ly"] / math.sqrt(3.0)),
                           s.steps, s.seed)
    with Arena(10000, precision=_lib.FP32C) as ar:
        with pytest.raises(_lib.SphError) as ei:
            ar.add(dense)
        assert ei.value.code == _lib.SPH_ERR_ARG and "too dense" in str(ei.value)
        ar.add(s)  # the arena stays usable
    # the serving loop refuses that job alone and serves the rest
    good = _good(2)
    srv = Server(2e9, precision=_lib.FP32C, max_resident=20000, max_job=4000, chunk=2, keep_results=True)
    try:
        recs = srv.run([good[0], dense, good[1]], np.zeros(3))
    finally:
        srv.close()
    assert recs[1].error is not None and recs[1].result is None
    for i in (0, 2):
        assert recs[i].error is None and recs[i].health == 0
        ref = alone([good[0], dense, good[1]][i], _lib.FP32C)
        assert np.array_equal(recs[i].result["r"], ref["r"])
