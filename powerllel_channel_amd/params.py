"""Run parameters of a channel simulation, read from the reference's param.in.

The file is a sequence of Fortran namelists (src/mod_parameters.f90:100-111,
202-203): &PARA, &MESH, &TIME, &RESTART, &PHYSICAL, &FORCE, &STATISTICS,
&OUTPUT and &MONITOR.  This reader takes the subset of namelist syntax those
files use: `name = value` items separated by commas or line breaks, arrays
as consecutive values (`is_forced = T, F, F`) or indexed (`probe_ijk(2) =
5`), repeat counts (`3*T`), logicals T/F/.true./.false., quoted strings,
reals with E or D exponents, `!` comments and a closing `/`.  Defaults are
the reference's initialisers (src/mod_parameters.f90:13-85); an unknown name
is an error, as for Fortran's namelist read.
"""
from __future__ import annotations

import copy
import math
import os
import re

# group -> {name: default}; a list default is an array of that length
_GROUPS = {
    "PARA": {"p_row": 1, "p_col": 1},
    "MESH": {"read_mesh": False, "mesh_type": 1, "stretch_ratio": 0.0, "stretch_func": 0,
             "nx": None, "ny": None, "nz": None, "lx": None, "ly": None, "lz": None},
    "TIME": {"nt_end": None, "dt": None, "nt_check": 100, "cfl_limit": 1.0, "div_limit": 1e-9},
    "RESTART": {"is_restart": False, "fn_prefix_input_inst": "save",
                "fn_prefix_input_stat": "save_stat"},
    "PHYSICAL": {"re": None, "u_ref": 1.0, "l_ref": 1.0, "initial_field": "poi", "u0": 0.0,
                 "init_with_noise": False, "noise_intensity": 0.0, "smooth_wall_visc": False,
                 "u_crf": 0.0},
    "FORCE": {"is_forced": [True, False, False], "vel_force": [1.0, 0.0, 0.0]},
    "STATISTICS": {"nt_init_stat": 10 ** 9, "sample_interval": 0,
                   "stat_which_var": [True] * 11},
    "OUTPUT": {"nt_out_scrn": 100, "nt_out_inst": 10 ** 9, "nt_out_stat": 10 ** 9,
               "nt_out_save": 10 ** 9, "overwrite_save": True, "auto_cleanup": True,
               "num_retained": 0, "nt_out_moni": 100},
    "MONITOR": {"out_forcing": False, "out_probe_point": False, "probe_ijk": [1, 1, 1],
                "out_skf_z": [False, False], "out_region": False, "nt_out_region": 100,
                "region_ijk": [1, 1, 1, 1, 1, 1]},
}
_REQUIRED = ("nx", "ny", "nz", "lx", "ly", "lz", "nt_end", "dt", "re")


class ParamError(ValueError):
    pass


def _strip_comment(line: str) -> str:
    out, q = [], None
    for ch in line:
        if q:
            out.append(ch)
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            out.append(ch)
        elif ch == "!":
            break
        else:
            out.append(ch)
    return "".join(out)


_TOKEN = re.compile(r"""\s*(?:'((?:[^']|'')*)'|"((?:[^"]|"")*)"|([^,\s/=]+\s*=)|(/)|([^,\s/]+))\s*,?""")


def _scalar(tok: str):
    t = tok.strip()
    low = t.lower()
    if low in ("t", ".true.", "true", ".t."):
        return True
    if low in ("f", ".false.", "false", ".f."):
        return False
    try:
        return int(t)
    except ValueError:
        pass
    try:
        return float(low.replace("d", "e"))
    except ValueError:
        raise ParamError(f"cannot read value {tok!r}") from None


def _parse_group(body: str):
    """[(name, index or None, [values])] of one namelist group's body."""
    items, name, idx, vals = [], None, None, []
    pos = 0
    while pos < len(body):
        m = _TOKEN.match(body, pos)
        if not m or m.end() == pos:
            if body[pos:].strip() == "":
                break
            raise ParamError(f"cannot parse namelist text near {body[pos:pos + 30]!r}")
        pos = m.end()
        sq, dq, key, slash, val = m.groups()
        if slash:
            break
        if key:
            if name is not None:
                items.append((name, idx, vals))
            k = key[:-1].strip().lower()
            km = re.fullmatch(r"([a-z_][a-z0-9_]*)(?:\((\d+)\))?", k)
            if not km:
                raise ParamError(f"bad name {key!r}")
            name, idx, vals = km.group(1), int(km.group(2)) if km.group(2) else None, []
            continue
        if name is None:
            raise ParamError("value before any name")
        if sq is not None or dq is not None:
            vals.append((sq if sq is not None else dq).replace("''", "'").replace('""', '"'))
        else:
            rm = re.fullmatch(r"(\d+)\*(.+)", val)
            if rm:
                vals.extend([_scalar(rm.group(2))] * int(rm.group(1)))
            else:
                vals.append(_scalar(val))
    if name is not None:
        items.append((name, idx, vals))
    return items


def parse_namelists(text: str):
    """{GROUP: [(name, index, values)]} of every &GROUP ... / in the text."""
    clean = "\n".join(_strip_comment(l) for l in text.splitlines())
    out = {}
    for m in re.finditer(r"&(\w+)(.*?)(?:/|\Z)", clean, re.S):
        group = m.group(1).upper()
        body = m.group(2)
        out.setdefault(group, []).extend(_parse_group(body + "/"))
    return out


class RunParams:
    """Every namelist value as an attribute (lower case), reference defaults
    for the groups and names the file leaves out."""

    def __init__(self, **overrides):
        for group, names in _GROUPS.items():
            for k, v in names.items():
                setattr(self, k, copy.deepcopy(v))
        for k, v in overrides.items():
            self.set(k, v)

    def set(self, name, value, index=None):
        name = name.lower()
        if not any(name in g for g in _GROUPS.values()):
            raise ParamError(f"unknown parameter {name!r}")
        cur = getattr(self, name)
        if isinstance(cur, list):
            vals = value if isinstance(value, (list, tuple)) else [value]
            start = (index or 1) - 1
            if start < 0 or start + len(vals) > len(cur):
                raise ParamError(f"{name}: {len(vals)} value(s) at {start + 1} overflow {len(cur)}")
            for q, v in enumerate(vals):
                cur[start + q] = v
        else:
            if isinstance(value, (list, tuple)):
                if len(value) != 1:
                    raise ParamError(f"{name} takes one value, got {len(value)}")
                value = value[0]
            setattr(self, name, value)

    @classmethod
    def from_text(cls, text: str) -> "RunParams":
        p = cls()
        for group, items in parse_namelists(text).items():
            if group not in _GROUPS:
                raise ParamError(f"unknown namelist &{group}")
            for name, idx, vals in items:
                if name not in _GROUPS[group]:
                    raise ParamError(f"&{group} has no member {name!r}")
                p.set(name, vals, idx)
        return p

    @classmethod
    def from_file(cls, path: str) -> "RunParams":
        if not os.path.exists(path):
            raise ParamError(f"File {path} doesn't exist!")
        with open(path) as fh:
            return cls.from_text(fh.read())

    @property
    def re_inv(self) -> float:
        """u_ref l_ref / re (src/mod_parameters.f90:187)."""
        return self.u_ref * self.l_ref / self.re

    def check(self) -> None:
        """checkInputParam (src/mod_parameters.f90:242-300) for one GPU."""
        missing = [k for k in _REQUIRED if getattr(self, k) is None]
        if missing:
            raise ParamError("missing parameters: " + ", ".join(missing))
        if self.nx % 2 or self.ny % 2 or self.nz % 2:
            raise ParamError("nx, ny and nz must be even")
        if self.p_row * self.p_col != 1:
            raise ParamError(f"p_row x p_col = {self.p_row} x {self.p_col}: one GPU holds the "
                             "whole domain here, so p_row = p_col = 1")
        if self.initial_field not in ("uni", "poi", "log", "mpi"):
            raise ParamError(f"initial_field {self.initial_field!r} is not uni, poi, log or mpi")
        if self.mesh_type not in (0, 1) or self.stretch_func not in (0, 1):
            raise ParamError("mesh_type must be 0 or 1 and stretch_func 0 or 1 "
                             "(src/mod_mesh.f90:120-162)")
        for k in ("nt_check", "nt_out_scrn", "nt_out_inst", "nt_out_save", "nt_out_moni"):
            if getattr(self, k) < 1:
                raise ParamError(f"{k} must be >= 1")
        if not (self.dt > 0 and math.isfinite(self.dt)):
            raise ParamError("dt must be positive")
