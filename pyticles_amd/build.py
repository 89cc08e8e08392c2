"""Build the gfx950 HIP library libpyticles_hip.so in-tree with hipcc.

    python -m pyticles_amd.build

Sources: pyticles_amd/csrc/*.hip.  Output: pyticles_amd/libpyticles_hip.so.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpyticles_hip.so")
SOURCES = ["sph_sort.hip", "sph_kernels.hip", "sph_capi.hip"]
ARCH = os.environ.get("PYTICLES_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=" + ARCH,
    # packed FP32 instructions stay off (hipcc would otherwise emit
    # v_pk_* forms from ordinary float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds fuse within a source expression only (written fma()
    # calls are kept): with HIP's default (fast) the backend also fuses
    # across statements as each instantiation's schedule allows, so two
    # instantiations of one walk (list entry widths, say) could round the
    # same terms differently
    "-ffp-contract=on",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose=True, out=LIB, defines=()):
    """Compile the library; defines are extra -D flags (tuning variants)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-D" + d for d in defines] + \
        ["-I" + os.path.join(HERE, "..", "include")] + srcs + ["-o", tmp]
    if verbose:
        print("[build] " + " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    log = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + log)
    if log and verbose:
        print(log)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
