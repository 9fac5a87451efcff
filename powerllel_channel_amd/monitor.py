"""Run-time monitors of a channel run, written in the reference's text formats.

Mirrors mod_monitor.f90 as driven by outputData (src/mod_dataIO.f90:59-70):
every `nt_out_moni` steps,

* monitor_forcing.out  (out_forcing, calcAndWriteForcing :262-280):
  nt, Mean_h(u) + u_crf, Mean_h(v), Mean_h(w) for the forced components
  (0 for the others), written only when some component is forced;
* monitor_probe.out    (out_probe_point, extractAndWriteProbePoint :146-167):
  nt, u + u_crf, v, w, p at probe_ijk, written only when the point lies in
  the domain;
* monitor_skf_z1.out / monitor_skf_z2.out (out_skf_z, calcAndWriteSkf
  :225-258 at k = 1 and k = nz): nt, u_tau, cf.

Headers as initMonitor writes them (:44-107), also on a restart, where the
files are opened for append.  Records use the reference's edit descriptors:
'(A9,n(A15))' for headers, '(I9,n(E15.7))' for records.  The reductions run on
the device (Channel.skin_friction, Channel.mean_h, Channel.probe); only the
scalars come to the host.
"""
from __future__ import annotations

import math
import os

__all__ = ["Monitor", "fortran_e", "fortran_i", "fortran_a"]


def fortran_i(n: int, w: int = 9) -> str:
    """Fortran Iw output: right-justified; a number wider than w prints w asterisks."""
    s = "%d" % n
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_a(s: str, w: int) -> str:
    """Fortran Aw output of a shorter string: right-justified in w columns
    (the leftmost w characters of a longer one)."""
    return s.rjust(w) if len(s) <= w else s[:w]


def fortran_e(x: float, w: int = 15, d: int = 7) -> str:
    """Fortran Ew.d output of a REAL(8) (the form gfortran writes): a mantissa
    0.d1..dd with d correctly rounded digits and the exponent as E+xx (three
    digits, +xxx without the E, past 99); zero as 0.000..E+00; NaN and
    Infinity spelled out, right-justified; asterisks when w is too narrow."""
    if math.isnan(x):
        s = "NaN"
    elif math.isinf(x):
        s = "Infinity" if x > 0 else "-Infinity"
        if len(s) > w:
            s = "Inf" if x > 0 else "-Inf"
    else:
        neg = math.copysign(1.0, x) < 0
        a = abs(x)
        if a == 0.0:
            digits, e = "0" * d, 0
        else:
            m, ex = ("%.*e" % (d - 1, a)).split("e")
            digits, e = m.replace(".", ""), int(ex) + 1
        if abs(e) <= 99:
            es = "E%+03d" % e
        elif abs(e) <= 999:
            es = "%+04d" % e
        else:
            return "*" * w
        s = ("-" if neg else "") + "0." + digits + es
        if len(s) > w:  # the optional leading zero goes first
            s = s.replace("0.", ".", 1)
    return s.rjust(w) if len(s) <= w else "*" * w


def _record(nt: int, values) -> str:
    return fortran_i(nt) + "".join(fortran_e(float(v)) for v in values) + "\n"


def _header(names) -> str:
    return fortran_a(names[0], 9) + "".join(fortran_a(n, 15) for n in names[1:]) + "\n"


class Monitor:
    """initMonitor / outputMonitor / freeMonitor for one Channel (one rank
    holds the whole domain, so every file has its writer).

        mon = Monitor(channel, out_skf_z=(True, True), nt_out_moni=100)
        for nt in range(1, nt_end + 1):
            channel.step(...)
            mon.output(nt, u, v, w, p)
        mon.close()
    """

    def __init__(self, channel, *, out_forcing=False, out_probe_point=False,
                 probe_ijk=(1, 1, 1), out_skf_z=(False, False), nt_out_moni=1,
                 u_ref=1.0, is_restart=False, directory=".",
                 fn_forcing="monitor_forcing.out", fn_probe="monitor_probe.out",
                 fn_skf_z=("monitor_skf_z1.out", "monitor_skf_z2.out")):
        if int(nt_out_moni) < 1:
            raise ValueError("nt_out_moni must be >= 1")
        self.ch = channel
        self.nt_out_moni = int(nt_out_moni)
        self.u_ref = float(u_ref)
        self.out_forcing = bool(out_forcing)
        self.out_probe_point = bool(out_probe_point)
        self.out_skf_z = tuple(bool(x) for x in out_skf_z)
        self.probe_ijk = tuple(int(x) for x in probe_ijk)
        mode = "a" if is_restart else "w"
        self._files = {}

        def _open(key, name, header):
            f = open(os.path.join(directory, name), mode)
            self._files[key] = f
            if header:
                f.write(header)
            return f

        if self.out_forcing:                                      # :49-56
            _open("forcing", fn_forcing, _header(["nt", "u_mean", "v_mean", "w_mean"]))
        self.have_probe = False
        if self.out_probe_point:                                  # :58-86
            i, j, k = self.probe_ijk
            self.have_probe = (1 <= i <= channel.nx and 1 <= j <= channel.ny
                               and 1 <= k <= channel.nz)
            head = ""
            if self.have_probe:
                head = ("Probe point (" + fortran_i(i, 4) + "," + fortran_i(j, 4) + ","
                        + fortran_i(k, 4) + "): \n" + _header(["nt", "u", "v", "w", "p"]))
            _open("probe", fn_probe, head)
        for s in range(2):                                        # :88-107
            if self.out_skf_z[s]:
                _open("skf%d" % s, fn_skf_z[s], _header(["nt", "u_tau", "cf"]))

    def output(self, nt, u, v, w, p, stream=None):
        """The monitor branch of outputData (src/mod_dataIO.f90:66-68)."""
        if nt % self.nt_out_moni == 0:
            self.output_now(nt, u, v, w, p, stream)

    def output_now(self, nt, u, v, w, p, stream=None):
        """outputMonitor, src/mod_monitor.f90:109-128."""
        ch = self.ch
        if self.out_forcing and any(ch.is_forced):
            q = (u, v, w)
            mean = [ch.mean_h(q[d], stream) if ch.is_forced[d] else 0.0 for d in range(3)]
            self._files["forcing"].write(_record(nt, (mean[0] + ch.u_crf, mean[1], mean[2])))
        if self.out_probe_point and self.have_probe:
            self._files["probe"].write(_record(nt, ch.probe(u, v, w, p, self.probe_ijk,
                                                            stream)))
        if any(self.out_skf_z):
            walls = ch.skin_friction(u, self.u_ref, stream)
            for s in range(2):
                if self.out_skf_z[s]:
                    self._files["skf%d" % s].write(_record(nt, walls[s]))
        for f in self._files.values():
            f.flush()

    def close(self):
        """freeMonitor, src/mod_monitor.f90:130-140."""
        for f in self._files.values():
            f.close()
        self._files = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
