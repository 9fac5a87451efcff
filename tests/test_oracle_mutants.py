"""The oracle's pins must fail on plausible transcription slips.

Each case plants one slip in a copy of oracle/oracle.c (an index off by one,
a swapped or transposed metric factor, a wrong sign, a wrong spacing), builds
that copy into a temporary library and runs the momentum pins
(tests/test_oracle_momentum.py) against it in a child process through
ORACLE_LIB: every planted slip must make at least one of them fail.  The
slips sit in the restatement of the momentum right-hand side,
src/mod_calcVel.f90:74-138.  CPU only; the product library is not involved."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# (original text in oracle.c, planted slip)
SLIPS = {
    "w viscous z: dzf_inv(k+1) -> dzf_inv(k)": (
        "dqt = (W(i, j, k + 1) - W(i, j, k)) * dzf_inv[k + 1];",
        "dqt = (W(i, j, k + 1) - W(i, j, k)) * dzf_inv[k];"),
    "w viscous z: dzc_inv -> dzf_inv": (
        "(dqt - dqb) * dzc_inv[k]) * re_inv;", "(dqt - dqb) * dzf_inv[k]) * re_inv;"),
    "w advection z: dzc_inv -> dzf_inv": (
        "(wt * wt - wb * wb) * dzc_inv[k]);", "(wt * wt - wb * wb) * dzf_inv[k]);"),
    "w: z weights of ue swapped": (
        "ue = (U(i, j, k) * dzf[k + 1] + U(i, j, k + 1) * dzf[k]) * dzc_inv[k];\n    uw",
        "ue = (U(i, j, k) * dzf[k] + U(i, j, k + 1) * dzf[k + 1]) * dzc_inv[k];\n    uw"),
    "w: vs on row j instead of j-1": (
        "vs = (V(i, j - 1, k) * dzf[k + 1] + V(i, j - 1, k + 1) * dzf[k]) * dzc_inv[k];",
        "vs = (V(i, j, k) * dzf[k + 1] + V(i, j, k + 1) * dzf[k]) * dzc_inv[k];"),
    "w: wn along x": (
        "wn = (W(i, j, k) + W(i, j + 1, k));\n    ws",
        "wn = (W(i, j, k) + W(i + 1, j, k));\n    ws"),
    "w: frame term with dy_inv": (
        "conv = conv + 0.5 * u_crf * (we - ww) * dx_inv;",
        "conv = conv + 0.5 * u_crf * (we - ww) * dy_inv;"),
    "v viscous x: dx_inv -> dy_inv": (
        "visc = ((dqe - dqw) * dx_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzf_inv[k]) * re_inv;\n    *fv",
        "visc = ((dqe - dqw) * dy_inv + (dqn - dqs) * dy_inv + (dqt - dqb) * dzf_inv[k]) * re_inv;\n    *fv"),
    "v: uw on row j-1 instead of j+1": (
        "uw = (U(i - 1, j, k) + U(i - 1, j + 1, k));",
        "uw = (U(i - 1, j, k) + U(i - 1, j - 1, k));"),
    "v: z weights of vb swapped": (
        "vb = (V(i, j, k) * dzf[k - 1] + V(i, j, k - 1) * dzf[k]) * dzc_inv[k - 1];",
        "vb = (V(i, j, k) * dzf[k] + V(i, j, k - 1) * dzf[k - 1]) * dzc_inv[k - 1];"),
    "v: frame term sign": (
        "conv = conv + 0.5 * u_crf * (ve - vw) * dx_inv;",
        "conv = conv - 0.5 * u_crf * (ve - vw) * dx_inv;"),
    "u: vs shifted to i-1": (
        "vs = (V(i, j - 1, k) + V(i + 1, j - 1, k));",
        "vs = (V(i - 1, j - 1, k) + V(i, j - 1, k));"),
    "u: wb half on level k": (
        "wb = (W(i, j, k - 1) + W(i + 1, j, k - 1));",
        "wb = (W(i, j, k) + W(i + 1, j, k - 1));"),
    "u viscous y: dqs one row up": (
        "dqs = (U(i, j, k) - U(i, j - 1, k)) * dy_inv;",
        "dqs = (U(i, j + 1, k) - U(i, j - 1, k)) * dy_inv;"),
}


def _build(args):
    name, (orig, slip), src, out = args
    assert src.count(orig) >= 1, f"slip site not found for {name!r}"
    cfile = out + ".c"
    with open(cfile, "w") as f:
        f.write(src.replace(orig, slip, 1))
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-fopenmp", "-ffp-contract=off",
                           "-std=gnu11", "-I", os.path.join(ROOT, "oracle"), cfile,
                           "-o", out, "-shared", "-fopenmp", "-lm"])
    return out


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    with open(os.path.join(ROOT, "oracle", "oracle.c")) as f:
        src = f.read()
    jobs = [(n, s, src, str(d / f"m{i}.so")) for i, (n, s) in enumerate(SLIPS.items())]
    with ThreadPoolExecutor(max_workers=4) as ex:
        libs = list(ex.map(_build, jobs))
    return dict(zip(SLIPS, libs))


@pytest.mark.parametrize("name", list(SLIPS))
def test_pins_catch_slip(mutants, name):
    env = dict(os.environ, ORACLE_LIB=mutants[name])
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_momentum.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "failed" in r.stdout, (name, r.stdout[-2000:])
