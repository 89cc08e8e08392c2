"""Trajectory snapshots from a device-resident box: asynchronous snapshots
equal synchronous downloads at the same steps, and a run resumed from the
last frame continues the original one."""
import numpy as np
import pytest

from pyticles_amd import _lib, configs, spam_nc
from pyticles_amd.device import DeviceBox, PinnedState, from_system

pytestmark = pytest.mark.gpu


def test_async_snapshot_equals_download():
    s = configs.square_system(20, 0.9, 21, 0)
    with from_system(s) as b:
        buf = PinnedState(s.n)
        b.step(7)
        b.snapshot(buf)
        b.step(3)  # overlaps the copy; must not disturb it
        b.snapshot_wait()
        ref_after = b.download("r", "v", "u")
        with from_system(s) as c:
            c.step(7)
            ref = c.download("r", "v", "u")
        assert np.array_equal(buf.r, ref["r"]) and np.array_equal(buf.v, ref["v"])
        assert np.array_equal(buf.u, ref["u"]) and np.array_equal(buf.m, s.m)
        assert not np.array_equal(buf.r, ref_after["r"])
        buf.close()


def test_device_trajectory_frames_and_resume(tmp_path):
    f = str(tmp_path / "traj.sph")
    s = configs.square_system(24, 0.9, 22, 0)
    with from_system(s) as b:
        tr = spam_nc.DeviceTrajectory(b, f, every=10, attribs={"workload": "test"})
        assert tr.run(30) == 4  # the start and three chunks
        tr.close()
        final = b.download("r", "v", "u")
        b.step(10)
        cont = b.download("r", "v", "u")
    assert spam_nc.nframes(f) == 4
    assert [spam_nc.read_frame(f, i)["timestep"] for i in range(4)] == [0.0, 10.0, 20.0, 30.0]
    fr = spam_nc.read_frame(f)
    assert np.array_equal(fr["position"], final["r"]) and np.array_equal(fr["internal_energy"], final["u"])
    # chunked stepping gives the same frames as plain steps of the same sizes
    with from_system(s) as c:
        c.step(10)
        assert np.array_equal(spam_nc.read_frame(f, 1)["velocity"], c.download("v"))
    # resume from the last frame in a new box
    with DeviceBox(s.n, s.prm) as r:
        r.upload(r=fr["position"], v=fr["velocity"], m=fr["mass"], u=fr["internal_energy"])
        r.step(10)
        got = r.download("r", "v", "u")
    L = s.prm["Lx"]
    dr = got["r"] - cont["r"]
    dr = np.where(dr > L / 2, dr - L, np.where(dr < -L / 2, dr + L, dr))
    # same state, different cell order inside the new box: equal to rounding
    assert np.max(np.abs(dr)) < 1e-12
    assert np.max(np.abs(got["v"] - cont["v"])) < 1e-11 * np.max(np.abs(cont["v"]))
    assert np.max(np.abs(got["u"] - cont["u"])) < 1e-11 * np.max(np.abs(cont["u"]))
