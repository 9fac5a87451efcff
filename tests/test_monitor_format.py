"""The monitor files' text formats without a GPU (powerllel_channel_amd.monitor).

Fortran's Ew.d output (ISO/IEC 1539 10.7.2.3.2, as gfortran writes a REAL(8)):
mantissa 0.d1...dd, exponent E+xx for |exp| <= 99 and +xxx for 99 < |exp| <= 999,
right-justified in w columns.  Known answers below follow those rules, and a
round trip pins the rounding against Python's correctly rounded conversion.
The Monitor class is driven by a stand-in channel that returns fixed scalars,
to check the files' layout (initMonitor headers, outputMonitor records,
nt_out_moni cadence, append on restart) against src/mod_monitor.f90.
"""
import math
import os

import numpy as np

from powerllel_channel_amd.monitor import Monitor, fortran_a, fortran_e, fortran_i


def _parse_e(s):
    s = s.strip()
    if s in ("NaN",):
        return float("nan")
    if s.endswith("Infinity"):
        return -math.inf if s.startswith("-") else math.inf
    if "E" in s:
        m, e = s.split("E")
    else:  # 0.1234567+101
        cut = max(s.rfind("+"), s.rfind("-"))
        m, e = s[:cut], s[cut:]
    return float(m) * 10.0 ** int(e) if abs(int(e)) < 300 else float(m + "e" + e)


def test_fortran_e_known_answers():
    cases = {
        1.0: "  0.1000000E+01",
        -1.5e-3: " -0.1500000E-02",
        0.0: "  0.0000000E+00",
        0.05: "  0.5000000E-01",
        123456789.0: "  0.1234568E+09",
        0.99999999: "  0.1000000E+01",   # rounds up into the next decade
        0.12345655: "  0.1234565E+00",   # stored just below the decimal tie
        0.12345675: "  0.1234568E+00",   # stored just above it
        1e-100: "  0.1000000E-99",
        1e100: "  0.1000000+101",
        -2.5e-120: " -0.2500000-119",
        float("nan"): "            NaN",
        float("inf"): "       Infinity",
        -float("inf"): "      -Infinity",
    }
    for x, want in cases.items():
        got = fortran_e(x)
        assert got == want, (x, got, want)
        assert len(got) == 15


def test_fortran_e_round_trip():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-30, 30, 2000),
                         [1e-300, -7.25e299, 5e-324]])
    for x in xs:
        s = fortran_e(float(x))
        assert len(s) == 15
        want = float("%.6e" % x)  # 7 significant digits, correctly rounded
        got = _parse_e(s)
        assert got == want or abs(got - want) <= 1e-15 * abs(want), (x, s)


def test_fortran_i_and_a():
    assert fortran_i(100) == "      100"
    assert fortran_i(-5, 4) == "  -5"
    assert fortran_i(123456, 4) == "****"
    assert fortran_a("nt", 9) == "       nt"
    assert fortran_a("u_mean", 15) == "         u_mean"


class _FakeChannel:
    nx, ny, nz = 8, 6, 4
    u_crf = 0.25
    is_forced = (True, False, False)

    def __init__(self):
        self.calls = []

    def mean_h(self, q, stream=None):
        self.calls.append("mean_h")
        return 0.75

    def probe(self, u, v, w, p, ijk, stream=None):
        self.calls.append(("probe", tuple(ijk)))
        return (1.25, -0.5, 0.0, 3.0)

    def skin_friction(self, u, u_ref=1.0, stream=None):
        self.calls.append(("skf", u_ref))
        return (0.0625, 0.0078125), (0.125, 0.03125)


def test_monitor_files_layout(tmp_path):
    ch = _FakeChannel()
    mon = Monitor(ch, out_forcing=True, out_probe_point=True, probe_ijk=(2, 3, 4),
                  out_skf_z=(True, True), nt_out_moni=100, u_ref=2.0, directory=str(tmp_path))
    for nt in (50, 100, 150, 200):
        mon.output(nt, None, None, None, None)
    mon.close()
    read = lambda n: open(os.path.join(tmp_path, n)).read().splitlines()  # noqa: E731
    assert read("monitor_forcing.out") == [
        "       nt         u_mean         v_mean         w_mean",
        "      100  0.1000000E+01  0.0000000E+00  0.0000000E+00",
        "      200  0.1000000E+01  0.0000000E+00  0.0000000E+00"]
    assert read("monitor_probe.out") == [
        "Probe point (   2,   3,   4): ",
        "       nt              u              v              w              p",
        "      100  0.1250000E+01 -0.5000000E+00  0.0000000E+00  0.3000000E+01",
        "      200  0.1250000E+01 -0.5000000E+00  0.0000000E+00  0.3000000E+01"]
    assert read("monitor_skf_z1.out") == [
        "       nt          u_tau             cf",
        "      100  0.6250000E-01  0.7812500E-02",
        "      200  0.6250000E-01  0.7812500E-02"]
    assert read("monitor_skf_z2.out")[1:] == [
        "      100  0.1250000E+00  0.3125000E-01",
        "      200  0.1250000E+00  0.3125000E-01"]
    assert ("skf", 2.0) in ch.calls and ("probe", (2, 3, 4)) in ch.calls
    # restart: append, header written again (mod_monitor.f90:49-56,88-107)
    with Monitor(ch, out_skf_z=(True, False), is_restart=True, directory=str(tmp_path)) as m2:
        m2.output(300, None, None, None, None)
    lines = read("monitor_skf_z1.out")
    assert len(lines) == 5 and lines[3].strip().startswith("nt") and lines[4].startswith("      300")


def test_monitor_probe_outside_domain(tmp_path):
    """A probe point outside the domain gives an empty probe file
    (have_this_point stays false, mod_monitor.f90:58-86)."""
    ch = _FakeChannel()
    with Monitor(ch, out_probe_point=True, probe_ijk=(9, 1, 1), directory=str(tmp_path)) as m:
        m.output(1, None, None, None, None)
    assert open(os.path.join(tmp_path, "monitor_probe.out")).read() == ""
    assert not ch.calls
