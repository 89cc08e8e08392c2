"""Simulation boxes with the reference's surface (box.py).

    Box(p, xmax, ymax, zmax)   box.py:16-26
    PeriodicBox                box.py:28-47
    MirrorBox                  box.py:49-74

The device leapfrog wraps positions periodically with
r - L floor(r / L) on the GPU; PeriodicBox.apply is kept for host-side
callers and translates an escaped coordinate by the box side (the
reference sets it to 0 or xmax, box.py:36-47, which is not a translation).
MirrorBox reflects at the walls as the reference does; the device path is
periodic only, so a system with a MirrorBox cannot be stepped on the GPU.
"""
import numpy as np


class Box:
    periodic = False

    def __init__(self, p="none", xmax=64, ymax=48, zmax=100):
        self.p = p
        self.xmax = xmax
        self.ymax = ymax
        self.zmax = zmax

    def apply(self, p=None):
        print("Do nothing")


class PeriodicBox(Box):
    periodic = True

    def apply(self, p):
        n = p.n
        for c, L in ((0, self.xmax), (1, self.ymax)):
            x = p.r[:n, c]
            x[:] = np.mod(x, L)
            x[x >= L] -= L


class MirrorBox(Box):
    def apply(self, p):
        n = p.n
        for c, L in ((0, self.xmax), (1, self.ymax)):
            x = p.r[:n, c]
            v = p.v[:n, c]
            hi = x > L
            lo = x < 0
            x[hi] = L
            x[lo] = 0
            v[hi | lo] = -v[hi | lo]
