"""SPAM trajectory files with the reference's schema (spam_nc.py).

The reference writes a netCDF file roughly after the AMBER trajectory
convention (spam_nc.py:28-118): dimensions ``timestep`` (unlimited),
``particle`` and ``spatial``; coordinate variables of the same names;
per-frame variables ``position`` and ``velocity`` (timestep, particle,
spatial), ``internal_energy`` and ``mass`` (timestep, particle); file
attributes from a dictionary.  ``write_step`` appends one frame
(spam_nc.py:121-145) and ``read_step`` restores a frame, the last by
default, into a particle object (spam_nc.py:148-171).

netCDF4 is not available to this package, so the same schema is stored in
a self-describing binary file:

    8 bytes   magic  b"SPHTRJ1\\n"
    8 bytes   little-endian uint64: header length H
    H bytes   JSON header: dimensions, variables (name, dtype, dims, offset
              within a frame), attributes, frame size in bytes
    frames    fixed-size records, float64 little-endian, in the order
              timestep, position, velocity, internal_energy, mass

Appending a frame never rewrites earlier ones, a frame is found by
arithmetic, and a partly written last frame (a crash while writing) is
ignored by the readers.  ``to_npz`` / ``read_npz`` convert to a NumPy
archive with the netCDF variable names.

``DeviceTrajectory`` writes frames from a device-resident box: after each
chunk of steps it starts an asynchronous snapshot (device copy in stream
order, host copy on a second stream into pinned buffers) that lands while
the next chunk runs.  Its frames carry the step index in ``timestep``
(``write_step`` keeps the reference's frame counter i+1).

Deviation: ``read_step`` also restores ``internal_energy`` (the reference
keeps that line commented out, spam_nc.py:160,168); resuming a run needs it.
"""
import json
import os
import struct

import numpy as np

MAGIC = b"SPHTRJ1\n"
FRAME_VARS = (("position", ("timestep", "particle", "spatial")),
              ("velocity", ("timestep", "particle", "spatial")),
              ("internal_energy", ("timestep", "particle")),
              ("mass", ("timestep", "particle")))


def _header(n, dim, attribs):
    variables = {"timestep": {"dtype": "f8", "dims": ["timestep"], "offset": 0}}
    off = 8
    for name, dims in FRAME_VARS:
        count = n * dim if "spatial" in dims else n
        variables[name] = {"dtype": "f8", "dims": list(dims), "offset": off, "count": count}
        off += 8 * count
    return {"format": "SPAM trajectory (netCDF schema of spam_nc.py in a flat file)",
            "dimensions": {"timestep": None, "particle": int(n), "spatial": int(dim)},
            "coordinates": {"particle": "0..particle-1", "spatial": "0..spatial-1"},
            "variables": variables, "attributes": dict(attribs or {}), "frame_bytes": off}


def create_sph_ncfile(filename, attribs, n, dim):
    """Create an empty trajectory file for n particles in dim dimensions;
    attribs become file attributes (spam_nc.py:28-118)."""
    hdr = dict(_header(n, dim, {"Date": 1, "Creator": "pyticles_amd"}))
    hdr["attributes"].update({str(k): v for k, v in dict(attribs or {}).items()})
    blob = json.dumps(hdr, sort_keys=True).encode()
    with open(filename, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<Q", len(blob)))
        f.write(blob)


def read_header(filename):
    """(header dict, byte offset of the first frame)."""
    with open(filename, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError("%s is not a SPAM trajectory file" % filename)
        (hl,) = struct.unpack("<Q", f.read(8))
        hdr = json.loads(f.read(hl).decode())
    return hdr, 16 + hl


def nframes(filename):
    """Complete frames in the file (a partial trailing frame is ignored)."""
    hdr, start = read_header(filename)
    return (os.path.getsize(filename) - start) // hdr["frame_bytes"]


def append_frame(filename, r, v, u, m, timestep=None):
    """Append one frame (arrays of n particles)."""
    hdr, start = read_header(filename)
    n, dim = hdr["dimensions"]["particle"], hdr["dimensions"]["spatial"]
    i = nframes(filename)
    rec = np.empty(hdr["frame_bytes"] // 8, dtype="<f8")
    rec[0] = float(i + 1 if timestep is None else timestep)  # spam_nc.py:131: timev[i] = i+1
    for name, _ in FRAME_VARS:
        var = hdr["variables"][name]
        a = {"position": r, "velocity": v, "internal_energy": u, "mass": m}[name]
        a = np.asarray(a, dtype=np.float64)
        shape = (n, dim) if name in ("position", "velocity") else (n,)
        if a.shape != shape:
            raise ValueError("%s: expected shape %s, got %s" % (name, shape, a.shape))
        o = var["offset"] // 8
        rec[o:o + var["count"]] = a.ravel()
    with open(filename, "r+b") as f:
        f.seek(start + i * hdr["frame_bytes"])  # overwrite a partial trailing frame
        f.write(rec.tobytes())
        f.truncate()


def read_frame(filename, step="last"):
    """dict(timestep, position, velocity, internal_energy, mass) of a frame."""
    hdr, start = read_header(filename)
    nf = nframes(filename)
    if nf == 0:
        raise ValueError("%s holds no complete frame" % filename)
    i = nf - 1 if step == "last" else int(step)
    if not 0 <= i < nf:
        raise IndexError("frame %d of %d" % (i, nf))
    n, dim = hdr["dimensions"]["particle"], hdr["dimensions"]["spatial"]
    with open(filename, "rb") as f:
        f.seek(start + i * hdr["frame_bytes"])
        rec = np.frombuffer(f.read(hdr["frame_bytes"]), dtype="<f8")
    out = {"timestep": float(rec[0])}
    for name, _ in FRAME_VARS:
        var = hdr["variables"][name]
        o = var["offset"] // 8
        a = rec[o:o + var["count"]].astype(np.float64)
        out[name] = a.reshape(n, dim) if name in ("position", "velocity") else a
    return out


def write_step(filename, p):
    """Append the state of particle object p (spam_nc.py:121-145)."""
    n = p.n
    append_frame(filename, p.r[0:n, :], p.v[0:n, :], p.u[0:n], p.m[0:n])


def read_step(filename, p, step="last"):
    """Make frame `step` (default the last) the state of particle object p
    (spam_nc.py:148-171), internal energy included."""
    fr = read_frame(filename, step)
    n = p.n
    p.r[0:n, :] = fr["position"][0:n, :]
    p.v[0:n, :] = fr["velocity"][0:n, :]
    p.u[0:n] = fr["internal_energy"][0:n]
    p.m[0:n] = fr["mass"][0:n]


def to_npz(filename, out):
    """The whole trajectory as a NumPy archive with the netCDF names."""
    hdr, _ = read_header(filename)
    nf = nframes(filename)
    frames = [read_frame(filename, i) for i in range(nf)]
    n, dim = hdr["dimensions"]["particle"], hdr["dimensions"]["spatial"]
    arrays = {"timestep": np.array([f["timestep"] for f in frames]), "particle": np.arange(n, dtype=np.int32),
              "spatial": np.arange(dim, dtype=np.int32)}
    for name, _ in FRAME_VARS:
        shape = (nf, n, dim) if name in ("position", "velocity") else (nf, n)
        arrays[name] = np.array([f[name] for f in frames]).reshape(shape)
    arrays["attributes"] = np.array(json.dumps(hdr["attributes"]))
    np.savez(out, **arrays)


def read_npz(path):
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    out["attributes"] = json.loads(str(out["attributes"]))
    return out


class DeviceTrajectory:
    """Frames of a device-resident box (device.DeviceBox) every `every`
    steps, written while the next steps run."""

    def __init__(self, box, filename, every=10, attribs=None):
        from .device import PinnedState
        self.box, self.filename, self.every = box, filename, int(every)
        self.bufs = [PinnedState(box.n), PinnedState(box.n)]
        self.frames = 0
        create_sph_ncfile(filename, attribs or {}, box.n, 2)

    def _write(self, buf, t):
        append_frame(self.filename, buf.r, buf.v, buf.u, buf.m, timestep=t)
        self.frames += 1

    def run(self, nsteps, t0=0):
        """nsteps steps in chunks of `every`; a frame after each chunk (and one
        of the starting state).  Returns the number of frames written."""
        cur = 0
        self.box.snapshot(self.bufs[cur])
        self.box.snapshot_wait()
        self._write(self.bufs[cur], t0)
        done = 0
        pending = None
        while done < nsteps:
            k = min(self.every, nsteps - done)
            self.box.step(k)  # returns with this chunk's steps done
            done += k
            if pending is not None:
                self.box.snapshot_wait()
                self._write(self.bufs[pending[0]], pending[1])
            cur ^= 1
            self.box.snapshot(self.bufs[cur])  # overlaps the next chunk
            pending = (cur, t0 + done)
        if pending is not None:
            self.box.snapshot_wait()
            self._write(self.bufs[pending[0]], pending[1])
        return self.frames

    def close(self):
        for b in self.bufs:
            b.close()
