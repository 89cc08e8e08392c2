"""Writes the golden vectors of this folder (run from the repo root):

    python tests/golden/make_golden.py

literal_n64.npz  -- a jittered 8x8 periodic system (L = 8) with velocities
                    and internal energies, and the outputs of the line-by-line
                    restatement in literal.py: pair list, separations,
                    densities, EOS, pressure + cohesion rates.
lattice.json     -- closed finite sums over an unjittered unit lattice:
                    density at h = 2 and h_l = 4 and the pair counts per
                    particle at r^2 < 17 and r^2 < 16, from the Lucy kernel
                    formula and integer lattice vectors.
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import literal  # noqa: E402


def lattice_sums():
    out = {}
    for h in (2.0, 4.0):
        terms = [literal.lucy_kernel(0.0, (0.0, 0.0), h)[0]]
        R = int(h) + 1
        for i in range(-R, R + 1):
            for j in range(-R, R + 1):
                if (i, j) != (0, 0):
                    r = math.sqrt(i * i + j * j)
                    terms.append(literal.lucy_kernel(r, (float(i), float(j)), h)[0])
        out["rho_h%g" % h] = math.fsum(terms)
    for thr in (17, 16):
        c = sum(1 for i in range(-5, 6) for j in range(-5, 6) if (i, j) != (0, 0) and i * i + j * j < thr)
        out["neighbours_rsq_lt_%d" % thr] = c
    return out


def literal_system():
    rng = np.random.Generator(np.random.PCG64(64))
    n, L = 64, 8.0
    k = np.arange(n)
    r = np.stack([(k % 8 + 0.5), (k // 8 + 0.5)], axis=1) + rng.uniform(-0.1, 0.1, size=(n, 2))
    r = np.mod(r, L)
    v = rng.normal(0.0, math.sqrt(0.9), size=(n, 2))
    v -= v.mean(axis=0)
    u = rng.uniform(-1.5, -0.5, size=n)
    m = np.ones(n)
    rl, vl = r.tolist(), v.tolist()
    iap = literal.verlet_build(rl, L, L, 4.0, 1.0)
    drij, rij, dv = literal.separations(rl, vl, iap, L, L)
    rho, rho_lr, T, p, pco, dwij, dwij_lr = literal.properties(m.tolist(), iap, rij, drij, u.tolist(),
                                                               2.0, 4.0, 2.0, 0.5, 1.0)
    vdot, udot = literal.spam_forces(m.tolist(), iap, dv, rho, rho_lr, p, pco, dwij, dwij_lr)
    return dict(r=r, v=v, u=u, m=m, L=L, iap=np.array(iap, dtype=np.int64), drij=np.array(drij),
                rij=np.array(rij), dv=np.array(dv), rho=np.array(rho), rho_lr=np.array(rho_lr),
                T=np.array(T), p=np.array(p), pco=np.array(pco), vdot=np.array(vdot), udot=np.array(udot))


if __name__ == "__main__":
    with open(os.path.join(HERE, "lattice.json"), "w") as f:
        json.dump(lattice_sums(), f, indent=1)
    np.savez(os.path.join(HERE, "literal_n64.npz"), **literal_system())
    print("wrote lattice.json and literal_n64.npz")
