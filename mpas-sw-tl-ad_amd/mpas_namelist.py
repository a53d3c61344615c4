"""Namelist reader with the interface of the reference's ``src/mpas_namelist.py``.

``namelist(nmlname)`` merges a Fortran namelist file over module-level
defaults (``src/mpas_namelist.py:17-47``) and parses the output interval.
f90nml is not available here, so the Fortran-namelist parsing is done by
:func:`read_fortran_namelist` below (groups ``&name ... /``, ``key = value``
with integers, reals, logicals and quoted strings, ``!`` comments; keys are
lower-cased the way f90nml returns them).

As in the reference, ``self.nlatm`` is the module-level defaults dict itself,
so values read by one instance are seen by later instances
(``src/mpas_namelist.py:54-60``).
"""

import json
import re

nl_default = \
    {
        "sw_model": {
            "config_test_case": 5,
            "config_time_integration": "RK4",
            "config_dt": 172.8,
            "config_calendar_type": "gregorian_noleap",
            "config_start_time": "0000-01-01_00:00:00",
            "config_stop_time": "none",
            "config_run_duration": "none",
            "config_stats_interval": 100,
            "config_h_ScaleWithMesh": False,
            "config_h_mom_eddy_visc2": 0.0,
            "config_h_mom_eddy_visc4": 0.0,
            "config_h_tracer_eddy_diff2": 0.0,
            "config_h_tracer_eddy_diff4": 0.0,
            "config_thickness_adv_order": 2,
            "config_tracer_adv_order": 2,
            "config_positive_definite": False,
            "config_monotonic": False,
            "config_wind_stress": False,
            "config_bottom_drag": False,
            "config_apvm_upwinding": 0.5,
            "config_num_halos": 2
        },
        "io": {
            "input": "x1.40962.grid.nc",
            "output": "x1.40962.output.nc",
            "output_interval": "06:00:00"
        }
    }

_TOKEN = re.compile(r"""'((?:[^']|'')*)'|"((?:[^"]|"")*)"|([^,\s]+)""")


def _bare(t):
    low = t.lower()
    if low in (".true.", "true", "t", ".t."):
        return True
    if low in (".false.", "false", "f", ".f."):
        return False
    try:
        return int(t)
    except ValueError:
        pass
    try:
        return float(low.replace("d", "e"))
    except ValueError:
        return t


def _values(rhs):
    out = []
    for m in _TOKEN.finditer(rhs):
        s, d, bare = m.groups()
        if s is not None:
            out.append(s.replace("''", "'"))
        elif d is not None:
            out.append(d.replace('""', '"'))
        else:
            out.append(_bare(bare))
    return out


def read_fortran_namelist(path):
    """Parse a Fortran namelist file into {group: {key: value}}."""
    groups = {}
    cur = None
    with open(path) as f:
        text = f.read()
    for raw in text.splitlines():
        line = raw.split("!", 1)[0].strip()
        if not line:
            continue
        if line.startswith("&"):
            cur = line[1:].strip().lower()
            groups.setdefault(cur, {})
            continue
        if line.startswith("/"):
            cur = None
            continue
        if cur is None or "=" not in line:
            continue
        key, rhs = line.split("=", 1)
        rhs = rhs.strip().rstrip("/").strip()
        vals = _values(rhs)
        groups[cur][key.strip().lower()] = vals[0] if len(vals) == 1 else vals
    return groups


class namelist(object):
    def __init__(self, nmlname='namelist.sw'):
        # defaults, then the file's values (src/mpas_namelist.py:51-61)
        self.nlatm = nl_default
        nl = read_fortran_namelist(nmlname)
        for group in nl:
            for config in nl[group]:
                self.nlatm.setdefault(group, {})[config] = nl[group][config]

        self.file_t0 = self.nlatm['io']['input']
        self.file_output = self.nlatm['io']['output']
        output_interval = self.nlatm['io']['output_interval']

        from pandas import to_timedelta
        self.output_interval = int((to_timedelta(output_interval)).total_seconds())

    def __str__(self):
        return json.dumps(self.nlatm, indent=4)
