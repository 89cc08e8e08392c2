"""Neighbour lists with the reference's surface (neighbour_list.py), built on
the GPU.

    NeighbourList(particle)                          neighbour_list.py:16-124
    VerletList(particle, cutoff=2.0, tolerance=1.0)  neighbour_list.py:127-252
    SortedVerletList(...)                            neighbour_list.py:262-301

build() runs the periodic cell-list pair build on the device (every pair i<j
with rsq < cutoff^2 + tolerance^2, lexicographic, bit-exact against the
reference's brute-force loop); separations() computes dr = r_j - r_i with
the minimum image, r, r^2 and dv = v_j - v_i for the listed pairs on the
device.  Arrays are 2-D ([rows, 2]); pair arrays start at an estimate and
grow on demand instead of the reference's maxn^2/2 - 1 rows.

Deliberate differences from the reference, each a bug or a float32 artefact
there: separations() applies the minimum image before r (the FP64
NeighbourList.separations semantics, neighbour_list.py:63-83, not the
float32 Cython pairsep whose rij ignores the image, pairsep.pyx:66-70);
compress() keeps dv with its pair and sets nip once
(neighbour_list.py:215-222); sort_by_r() permutes rij once (:276-278).
"""
import numpy as np

from . import _bind, _lib

DIM = 2


class NeighbourList:
    """Brute-force list: every pair i<j (neighbour_list.py:46-56)."""

    def __init__(self, particle):
        self.nip = 0
        self.particle = particle
        self.max_interactions = (particle.maxn * particle.maxn) // 2 - 1
        cap = max(16, min(self.max_interactions, 32 * particle.maxn))
        dim = getattr(particle, "dim", DIM)
        self.iap = np.zeros((cap, 2), dtype=np.int64)
        self.rij = np.zeros(cap, dtype=float)
        self.rsq = np.zeros(cap)
        self.drij = np.zeros((cap, dim), dtype=float)
        self.dv = np.zeros((cap, dim), dtype=float)
        self.wij = np.zeros(cap)
        self.wij_lr = np.zeros(cap)
        self.dwij = np.zeros((cap, dim))
        self.dwij_lr = np.zeros((cap, dim))
        self.rebuild_list = False
        self.nforce = 0
        self.forces = []

    # -- device plumbing ------------------------------------------------
    def _device(self):
        p = self.particle
        prm, _ = _bind.params_for(p, self)
        dev = _bind.device_box(p, prm)
        _bind.upload(p, dev)
        return dev

    def _store_pairs(self, iap):
        nip = iap.shape[0]
        _bind.ensure_pair_capacity(self, nip)
        self.iap[:nip] = iap
        self.nip = nip

    # -- reference surface ----------------------------------------------
    def build(self):
        """When in doubt, use brute force: all i<j pairs."""
        self.rebuild_list = False
        n = self.particle.n
        i, j = np.triu_indices(n, 1)
        self._store_pairs(np.stack([i, j], axis=1).astype(np.int64))

    def compress(self):
        """There is no concept of compression for a brute force list."""
        print("Cannot compress a brute list")

    def separations(self):
        """drij (minimum image), rij, rsq, dv for every listed pair."""
        ni = self.nip
        if ni == 0:
            return
        dev = self._device()
        drij, rij, rsq, dv = dev.separations(self.iap[:ni])
        self.drij[:ni, :2] = drij
        self.rij[:ni] = rij
        self.rsq[:ni] = rsq
        self.dv[:ni, :2] = dv

    def find_pair(self, i, j):
        """Index k of pair (i, j), or -1."""
        ni = self.nip
        hit = np.nonzero((self.iap[:ni, 0] == i) & (self.iap[:ni, 1] == j))[0]
        return int(hit[0]) if hit.size else -1

    def apply_minimum_image(self):
        ni = self.nip
        if ni == 0:
            return
        bx = self.particle.box
        dr = np.ascontiguousarray(self.drij[:ni, :2], dtype=np.float64)
        _lib.check(_lib.lib.sph_minimum_image(_lib.context(), ni, _lib.dptr(dr), float(bx.xmax),
                                              float(bx.ymax)))
        self.drij[:ni, :2] = dr

    def minimum_image(self, dr, xmax, ymax, zmax=None):
        """Minimum image of one displacement, in place (first two components)."""
        d = np.ascontiguousarray(np.asarray(dr[:2], dtype=np.float64).reshape(1, 2))
        _lib.check(_lib.lib.sph_minimum_image(_lib.context(), 1, _lib.dptr(d), float(xmax), float(ymax)))
        dr[0] = d[0, 0]
        dr[1] = d[0, 1]


class VerletList(NeighbourList):
    """Pairs within the interaction range plus a tolerance
    (rsq < cutoff^2 + tolerance^2, neighbour_list.py:178)."""

    def __init__(self, particle, cutoff=2.0, tolerance=1.0):
        NeighbourList.__init__(self, particle)
        self.cutoff_radius = cutoff
        self.cutoff_radius_sq = cutoff ** 2
        self.tolerance_sq = tolerance * tolerance
        self.r_old = np.zeros(self.particle.r.shape)

    def build(self):
        """Cell-list build on the GPU; stores iap, drij, dv, rsq and snapshots
        the positions (neighbour_list.py:160-189)."""
        self.rebuild_list = False
        p = self.particle
        self.r_old[:, :] = p.r[:, :]
        if p.n == 0:
            self.nip = 0
            return
        dev = self._device()
        iap = dev.pairs()
        self._store_pairs(iap)
        ni = self.nip
        if ni:
            drij, rij, rsq, dv = dev.separations(iap)
            self.drij[:ni, :2] = drij
            self.dv[:ni, :2] = dv
            self.rsq[:ni] = rsq

    def compress(self):
        """Drop listed pairs now outside cutoff + tolerance, then ponder a
        rebuild (neighbour_list.py:191-223)."""
        ni = self.nip
        if ni:
            dev = self._device()
            drij, rij, rsq, dv = dev.separations(self.iap[:ni])
            keep = rsq < (self.cutoff_radius_sq + self.tolerance_sq)
            q = int(np.count_nonzero(keep))
            self.iap[:q] = self.iap[:ni][keep]
            self.drij[:q, :2] = drij[keep]
            self.dv[:q, :2] = dv[keep]
            self.rsq[:q] = rsq[keep]
            self.nip = q
        self.ponder_rebuild()

    def ponder_rebuild(self):
        """rebuild_list = True if any particle moved more than the tolerance
        since the build (neighbour_list.py:225-234)."""
        n = self.particle.n
        dr = self.r_old[:n] - self.particle.r[:n]
        dsq = float(np.max(np.sum(dr * dr, axis=1))) if n else 0.0
        if dsq > self.tolerance_sq:
            self.rebuild_list = True

    def separations(self):
        NeighbourList.separations(self)


class SmoothVerletList:
    """Placeholder kept for import compatibility (neighbour_list.py:254-260)."""

    def compute(self):
        pass


class SortedVerletList(VerletList):
    """Verlet list sorted by separation, largest first (neighbour_list.py:262-301)."""

    def separations(self):
        VerletList.separations(self)
        self.sort_by_r()

    def sort_by_r(self):
        ni = self.nip
        idx = np.argsort(self.rsq[:ni], kind="stable")[::-1]
        for name in ("iap", "rij", "rsq", "dv", "drij", "wij", "dwij", "wij_lr", "dwij_lr"):
            a = getattr(self, name)
            a[:ni] = a[:ni][idx]

    def sort_by_i(self):
        pass

    def build(self):
        VerletList.build(self)
        self.sort_by_r()

    def compress(self):
        VerletList.compress(self)
        self.sort_by_r()
