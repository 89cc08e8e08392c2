"""Neighbour-list entry forms: 16-bit slot offsets (the default) and 32-bit
slots hold the same entries in the same order, so every result is the same
bits in either form; a box whose offsets do not fit 16 bits holds the step
of that build and continues with 32-bit entries, in a single box (stand-alone
evaluation, sph_step, open runs) and in an arena beside other boxes."""
import numpy as np
import pytest

from pyticles_amd import _lib, configs
from pyticles_amd.arena import Arena
from pyticles_amd.device import DeviceBox

pytestmark = pytest.mark.gpu

FIELDS = ("r", "v", "u", "rho")


def _box(s, bits):
    b = DeviceBox(s.n, s.prm)
    b.set_list_bits(bits)
    b.upload(r=s.r, v=s.v, m=s.m, u=np.zeros(s.n))
    b.init_thermal(s.T)
    return b


def _same(a, b, tag):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), (tag, k)


def wide_system(precision, Lx=8192, Ly=16, T=0.8, seed=41, steps=10):
    # one particle per unit area on a jittered Lx x Ly lattice: 1,820 x 3
    # cells of ~24 particles, so a row of cells spans ~44k slots and the
    # listed neighbours one row away (or across the periodic seam) lie
    # beyond 16-bit offsets either way
    rng = np.random.Generator(np.random.PCG64(seed))
    n = Lx * Ly
    k = np.arange(n)
    r = np.empty((n, 2))
    r[:, 0] = k % Lx + 0.5
    r[:, 1] = k // Lx + 0.5
    r += rng.uniform(-0.1, 0.1, size=(n, 2))
    r[:, 0] = np.mod(r[:, 0], Lx)
    r[:, 1] = np.mod(r[:, 1], Ly)
    v = configs.velocities(n, T, rng)
    prm = configs.params(Lx, Ly=Ly, precision=precision)
    return configs.System(np.ascontiguousarray(r), v, np.ones(n), T, prm, steps, seed)


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
@pytest.mark.parametrize("nside", [48, 512])
def test_16_and_32_bit_entries_give_the_same_bits(precision, nside):
    # 512 x 512 (C2's box, 262,144 slots): the periodic neighbours across the
    # top and bottom rows are stored modulo the slot count
    s = configs.square_system(nside, 0.9, 51, 0, precision=precision)
    out = {}
    for bits in (16, 32):
        with _box(s, bits) as b:
            b.step(30)
            assert b.list_bits() == bits
            out[bits] = b.download(*FIELDS)
    _same(out[16], out[32], (precision, nside))


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_wide_box_falls_back_to_32_bit_entries(precision):
    s = wide_system(precision)
    with _box(s, 32) as a:
        a.step(s.steps)
        ref = a.download(*FIELDS)
    # the stand-alone evaluation of init_thermal meets the overflow first
    with _box(s, 16) as b:
        assert b.list_bits() == 32
        b.step(s.steps)
        _same(b.download(*FIELDS), ref, "init")


@pytest.mark.parametrize("precision", [_lib.FP64, _lib.FP32C])
def test_held_build_inside_a_run(precision):
    # lists built in 32 bits, then 16 asked for: the first build inside the
    # next run overflows, the box holds that step and continues in 32 bits
    s = wide_system(precision, steps=40)
    with _box(s, 32) as a:
        a.step(5)
        a.step(s.steps)
        ref = a.download(*FIELDS)
        builds = a.run_stats()["builds"]
    assert builds >= 2
    with _box(s, 32) as b:  # sph_step
        b.step(5)
        b.set_list_bits(16)
        b.step(s.steps)
        assert b.list_bits() == 32
        _same(b.download(*FIELDS), ref, "step")
    with _box(s, 32) as c:  # open run, one step per call
        c.step(5)
        c.set_list_bits(16)
        for _ in range(s.steps):
            c.advance(1)
        c.sync()
        assert c.list_bits() == 32
        _same(c.download(*FIELDS), ref, "advance")


def test_arena_box_falls_back_beside_others():
    prec = _lib.FP32C
    wide = wide_system(prec, steps=12)
    small = [configs.lognormal_box(seed, 256, 2048, 0.4, 1.0, steps=st) for seed, st in ((61, 14), (62, 9))]
    ref = []
    for s in [wide] + small:
        with _box(s, 32) as b:
            b.step(s.steps)
            ref.append(b.download(*FIELDS))
    with Arena(wide.n + sum(s.n for s in small) + 256, precision=prec) as ar:
        ids = [ar.add(s) for s in small[:1] + [wide] + small[1:]]
        order = [1, 0, 2]  # ids[i] holds system order[i] of [wide] + small
        for _ in range(200):
            st, _, hl = ar.poll_all()  # releases a held box
            assert not np.any(hl[ids])
            if all(st[i] == _lib.BOX_DONE for i in ids):
                break
            ar.step(1)
        else:
            raise AssertionError("boxes did not finish")
        for i, j in zip(ids, order):
            got = ar.result(i)
            for k in FIELDS:
                assert np.array_equal(got[k], ref[j][k]), (j, k)
        assert ar.list_bits(ids[1]) == 32
        assert ar.list_bits(ids[0]) == 16 and ar.list_bits(ids[2]) == 16
