"""Trajectory files (the reference's spam_nc schema) on the CPU: round trip,
frame selection, a torn last frame, NumPy export."""
import numpy as np
import pytest

from pyticles_amd import spam_nc


class P:
    def __init__(self, n, maxn, seed):
        rng = np.random.default_rng(seed)
        self.n = n
        self.r = rng.uniform(0, 10, size=(maxn, 2))
        self.v = rng.normal(size=(maxn, 2))
        self.u = rng.normal(size=maxn)
        self.m = rng.uniform(0.5, 1.5, size=maxn)


def test_write_read_round_trip(tmp_path):
    f = str(tmp_path / "run.sph")
    spam_nc.create_sph_ncfile(f, {"creator": "test", "log": "functional test"}, 5, 2)
    hdr, _ = spam_nc.read_header(f)
    assert hdr["dimensions"] == {"timestep": None, "particle": 5, "spatial": 2}
    assert set(hdr["variables"]) == {"timestep", "position", "velocity", "internal_energy", "mass"}
    assert hdr["attributes"]["creator"] == "test" and hdr["attributes"]["Creator"] == "pyticles_amd"
    states = [P(5, 8, s) for s in range(3)]
    for p in states:
        spam_nc.write_step(f, p)
    assert spam_nc.nframes(f) == 3
    q = P(5, 8, 99)
    spam_nc.read_step(f, q)  # last frame
    for name in ("r", "v"):
        assert np.array_equal(getattr(q, name)[:5], getattr(states[2], name)[:5])
    assert np.array_equal(q.u[:5], states[2].u[:5]) and np.array_equal(q.m[:5], states[2].m[:5])
    spam_nc.read_step(f, q, step=1)
    assert np.array_equal(q.r[:5], states[1].r[:5])
    assert [spam_nc.read_frame(f, i)["timestep"] for i in range(3)] == [1.0, 2.0, 3.0]  # i+1, spam_nc.py:131
    with pytest.raises(IndexError):
        spam_nc.read_frame(f, 3)


def test_torn_last_frame_is_ignored_and_overwritten(tmp_path):
    f = str(tmp_path / "run.sph")
    spam_nc.create_sph_ncfile(f, {}, 4, 2)
    a, b = P(4, 4, 1), P(4, 4, 2)
    spam_nc.write_step(f, a)
    spam_nc.write_step(f, b)
    with open(f, "r+b") as fh:  # a crash halfway through the second frame
        fh.seek(0, 2)
        fh.truncate(fh.tell() - 30)
    assert spam_nc.nframes(f) == 1
    assert np.array_equal(spam_nc.read_frame(f)["position"], a.r[:4])
    spam_nc.write_step(f, b)
    assert spam_nc.nframes(f) == 2
    assert np.array_equal(spam_nc.read_frame(f)["velocity"], b.v[:4])


def test_npz_export(tmp_path):
    f = str(tmp_path / "run.sph")
    spam_nc.create_sph_ncfile(f, {"k": 1}, 3, 2)
    ps = [P(3, 3, s) for s in range(4)]
    for p in ps:
        spam_nc.write_step(f, p)
    out = str(tmp_path / "run.npz")
    spam_nc.to_npz(f, out)
    z = spam_nc.read_npz(out)
    assert z["position"].shape == (4, 3, 2) and z["internal_energy"].shape == (4, 3)
    assert np.array_equal(z["mass"][2], ps[2].m)
    assert z["attributes"]["k"] == 1 and list(z["particle"]) == [0, 1, 2]


def test_rejects_wrong_shapes_and_files(tmp_path):
    f = str(tmp_path / "run.sph")
    spam_nc.create_sph_ncfile(f, {}, 3, 2)
    with pytest.raises(ValueError):
        spam_nc.append_frame(f, np.zeros((4, 2)), np.zeros((3, 2)), np.zeros(3), np.zeros(3))
    with pytest.raises(ValueError):
        spam_nc.read_frame(f)  # no frame yet
    g = tmp_path / "other.bin"
    g.write_bytes(b"not a trajectory")
    with pytest.raises(ValueError):
        spam_nc.read_header(str(g))
