"""The closed-form checks of tests/test_oracle_closed_forms.py must reject
a wrong oracle, not only accept the right one.  Each case copies
oracle/oracle.c, plants one plausible mistake in the transport terms
(a factor 2, a flipped sign, a transposed or dropped operand), builds the
copy with the oracle's own flags and runs every closed-form check against
it: at least one must fail.  CPU only (gcc)."""
import os
import shutil
import subprocess

import pytest

from oracle import oracle as O

import test_oracle_closed_forms as CF  # noqa: E402  (tests/ is on sys.path under pytest)

SRC = os.path.join(os.path.dirname(O.__file__), "oracle.c")
CFLAGS = ["-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-shared"]

# name -> [(text in oracle.c, replacement, occurrences)]
MUTATIONS = {
    "grad T doubled": [("double dT = T[j] - T[i];", "double dT = 2. * (T[j] - T[i]);", 1)],
    "heat flux sign": [("qx = -kcond * gradT[2 * i];", "qx = kcond * gradT[2 * i];", 1),
                       ("qy = -kcond * gradT[2 * i + 1];", "qy = kcond * gradT[2 * i + 1];", 1)],
    "heat flux over rho, not rho^2": [("Q[2 * i] = qx / r2;", "Q[2 * i] = qx / rho[i];", 1)],
    "conduction term sign": [("udot[i] += m[j] * (e + c);", "udot[i] += m[j] * (e - c);", 1),
                             ("udot[j] += m[i] * (e - c);", "udot[j] += m[i] * (e + c);", 1)],
    "conduction uses q_i twice": [("(Q[2 * i] + Q[2 * j]) * wx", "(Q[2 * i] + Q[2 * i]) * wx", 1)],
    "zeta sign": [(" - zeta * tr;", " + zeta * tr;", 2)],
    "zeta dropped": [(" - zeta * tr;", ";", 2)],
    "Pi_xy not symmetrised": [("pi_xy = -eta * (gxy + gyx);", "pi_xy = -eta * (gxy + gxy);", 1)],
    "force operand transposed": [("double fx = axx * wx + axy * wy + cs * wlx;",
                                  "double fx = axx * wx + axy * wx + cs * wlx;", 1)],
    "grad v component transposed": [("gradv[4 * i + 1] += vj * dvx * gy;",
                                     "gradv[4 * i + 1] += vj * dvy * gx;", 1)],
    "heating factor 1 instead of 1/2": [("double e = 0.5 * (dvx * fx + dvy * fy);",
                                         "double e = 1.0 * (dvx * fx + dvy * fy);", 1)],
    "heating velocity component": [("double e = 0.5 * (dvx * fx + dvy * fy);",
                                    "double e = 0.5 * (dvx * fx + dvx * fy);", 1)],
}


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    return tmp_path_factory.mktemp("oracle_mutants")


def _build(src_text, path):
    c = str(path) + ".c"
    with open(c, "w") as f:
        f.write(src_text)
    subprocess.run(["gcc"] + CFLAGS + ["-o", str(path), c, "-lm"], check=True, capture_output=True)


def test_unmutated_copy_passes(workdir):
    so = workdir / "liboracle_copy.so"
    _build(open(SRC).read(), so)
    with O.library(str(so)):
        for check in CF.CHECKS:
            check()


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_closed_forms_reject_mutant(workdir, name):
    text = open(SRC).read()
    for old, new, count in MUTATIONS[name]:
        assert text.count(old) == count, "mutation site %r not found %d times" % (old, count)
        text = text.replace(old, new)
    so = workdir / ("liboracle_%s.so" % name.replace(" ", "_").replace("/", "_").replace("^", ""))
    _build(text, so)
    failed = []
    with O.library(str(so)):
        for check in CF.CHECKS:
            try:
                check()
            except AssertionError:
                failed.append(check.__name__)
    assert failed, "mutant %r passed every closed-form check" % name
