"""SPH rendering onto a grid (interpolate.splash_map_2d) on the device
against the oracle's restatement of the reference loop."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from pyticles_amd import _lib, configs, interpolate

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_lattice_sites_render_one():
    ns, L = 32, 32.0
    r = configs.lattice(ns * ns, L, None, jitter=0)
    rho0 = json.load(open(os.path.join(GOLD, "lattice.json")))["rho_h2"]
    x = np.arange(64) * 0.5
    Z = interpolate.splash_map_2d(x, x, r, 1.0, rho0, 1.0, 2.0, (0, L, 0, L), 2.6)
    np.testing.assert_allclose(Z[1::2, 1::2], 1.0, rtol=1e-14)


@pytest.mark.parametrize("L,cut,res", [(40.0, 2.6, 0.5), (10.0, 4.5, 0.25), (64.0, 5.0, 1.0)])
def test_matches_oracle(L, cut, res):
    # cutoff >= h + pixel width: the reference's half-open pixel ranges drop
    # only zero-weight pixels, so the two agree to summation order
    rng = np.random.default_rng(int(L))
    n = int(L * L)
    r = configs.lattice(n, L, rng, jitter=0.3)
    m = rng.uniform(0.8, 1.2, n)
    rho = rng.uniform(0.9, 1.1, n)
    a = rng.normal(1.0, 0.3, n)
    h = 2.0 if cut >= 2.0 + res else 1.0
    x = np.arange(int(L / res)) * res
    y = np.arange(int(L / res)) * res
    Z = interpolate.splash_map_2d(x, y, r, m, rho, a, h, (0, L, 0, L), cut)
    Zo = O.splash_map_2d(x, y, r, m, rho, a, h, (0, L, 0, L), cut)
    assert rel(Z, Zo) < 1e-12
    Z3 = interpolate.splash_map_2d(x, y, r, m, rho, 3.0 * a, h, (0, L, 0, L), cut)
    assert rel(Z3, 3.0 * Z) < 1e-13


def test_per_particle_h_and_empty():
    rng = np.random.default_rng(3)
    L, n = 20.0, 300
    r = rng.uniform(0, L, size=(n, 2))
    h = rng.uniform(1.0, 2.0, n)
    x = np.arange(40) * 0.5
    Z = interpolate.splash_map_2d(x, x, r, 1.0, 1.0, 1.0, h, (0, L, 0, L), 2.6)
    Zo = O.splash_map_2d(x, x, r, np.ones(n), np.ones(n), np.ones(n), h, (0, L, 0, L), 2.6)
    assert rel(Z, Zo) < 1e-12
    Ze = interpolate.splash_map_2d(x, x, np.zeros((0, 2)), [], [], [], [], (0, L, 0, L), 2.6)
    assert Ze.shape == (40, 40) and np.all(Ze == 0.0)
    with pytest.raises(_lib.SphError):
        interpolate.splash_map_2d(x, x, r, 1.0, 1.0, 1.0, h, (0, L, 0, L), 0.0)
