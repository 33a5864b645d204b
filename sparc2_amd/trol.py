"""Reader of the reference's control file (``sparc_parameters``).

Follows READ_SPARC_PARAMETERS (initialize.f90:1385-1924) line by line:

* a line is split at its first ``=``; the name is the text before it with
  trailing blanks removed (leading blanks and tabs are kept, so an indented
  name matches nothing), the value is everything after it
  (initialize.f90:1417-1425); lines without ``=`` and unknown names are
  ignored, names are case-sensitive;
* each value is read list-directed (``read(linevalue, *)``): blank- or
  comma-separated items, ``d``/``D`` exponents, ``r*c`` repeats, a ``/``
  ends the list, extra items are ignored;
* a value that fails to read ends the whole parse -- the reference reuses the
  read's iostat as the loop's end-of-file flag (initialize.f90:1414-1416) --
  and the rest of the file is not looked at;
* flags act in file order: ``USE_HPML = 0`` switches USE_PML on
  (initialize.f90:1597-1605), ``PERIODIC`` is refused when USE_HPML is
  already on (:1607-1620), a PML line keeps its width only if the matching
  PML switch is on when the line is read (:1622-1660), ``pulse_*`` values are
  read only when PULSE is already on and take fixed stand-ins otherwise
  (:1689-1751), ``t_xy``/``t_z`` are raised to the timestep read so far
  (:1759-1765);
* after the loop: 2-D adjustments, the derived step counts and cadences
  (:1778-1924).

``read_parameters`` returns every variable the routine sets;
``to_run_config`` maps them onto :class:`RunConfig` for the 3-D pulse-driven
paths this package implements.
"""
from __future__ import annotations

import math
import re
from typing import Dict, List, Optional, Sequence

from .config import RunConfig


class ParameterError(ValueError):
    pass


# --------------------------------------------------------------------------
# list-directed input
# --------------------------------------------------------------------------
_NUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eEdDqQ][+-]?\d+)?$")
_INT = re.compile(r"^[+-]?\d+$")


def _tokens(text: str) -> List[Optional[str]]:
    """Items of a list-directed record: values are separated by blanks or by
    a comma with optional blanks around it; a comma where a value is expected
    gives a null item; a slash ends the record; a quoted string is one item;
    r*c repeats c r times (r* alone repeats nulls)."""
    out: List[Optional[str]] = []
    i, n = 0, len(text)
    expect = True               # at the start and after a comma
    while i < n:
        ch = text[i]
        if ch == " ":
            i += 1
            continue
        if ch == "/":
            break
        if ch == ",":
            if expect:
                out.append(None)
            expect = True
            i += 1
            continue
        if ch in "'\"":
            j, buf = i + 1, []
            while j < n:
                if text[j] == ch:
                    if j + 1 < n and text[j + 1] == ch:   # doubled quote
                        buf.append(ch)
                        j += 2
                        continue
                    break
                buf.append(text[j])
                j += 1
            out.append("".join(buf))
            i = j + 1
        else:
            j = i
            while j < n and text[j] not in " ,/":
                j += 1
            tok = text[i:j]
            m = re.match(r"^(\d+)\*(.*)$", tok)
            if m:
                out.extend([m.group(2) or None] * int(m.group(1)))
            else:
                out.append(tok)
            i = j
        expect = False
    return out


def _conv(tok: str, kind: str):
    if kind == "i":
        if not _INT.match(tok):
            raise ParameterError(f"not an integer: {tok!r}")
        return int(tok)
    if kind == "r":
        if not _NUM.match(tok):
            raise ParameterError(f"not a real: {tok!r}")
        return float(re.sub(r"[dDqQ]", "e", tok))
    return tok  # character


def list_read(text: str, kinds: Sequence[str], current: Sequence = ()) -> list:
    """Read len(kinds) items ('i' integer, 'r' real, 'c' character) from one
    record; a null item keeps the current value.  Raises ParameterError where
    the Fortran read would set a nonzero iostat (bad item, or too few
    items in the record)."""
    toks = _tokens(text)
    vals = []
    for q, kind in enumerate(kinds):
        if q >= len(toks):
            raise ParameterError(f"record ends after {len(toks)} of {len(kinds)} items")
        t = toks[q]
        if t is None:
            vals.append(current[q] if q < len(current) else None)
        else:
            vals.append(_conv(t, kind))
    return vals


# --------------------------------------------------------------------------
# the routine
# --------------------------------------------------------------------------
_SIMPLE = {
    # name: list-directed kind (i integer, r real, c character)
    "dir_out": "c", "dir_bkg": "c", "simname": "c",
    "nx": "i", "ny": "i", "nz": "i",
    "xlength": "r", "ylength": "r",
    "time0": "i", "timestep": "r", "solartime": "r",
    "chunkoutputcad": "r", "fulloutputcad": "r", "minsavtime": "r",
    "xcb": "i", "xct": "i", "ycb": "i", "yct": "i", "zcb": "i", "zct": "i",
    "timestamp_size": "i", "maxva": "r", "TWOP5D": "i",
    "sp_str": "r", "sp_length": "r", "sp_decay": "r",
    "forcingfunc": "c", "cadforcing": "r",
    "randomexcitedepth": "r", "randomexcitesigma": "r", "randomexciteampl": "r",
    "s_xy": "r", "s_z": "r",
}
_SWITCH = {"restart": "RESTART", "STABILISE": "STABILISE", "DAMPING": "DAMPING",
           "MODE": "DISPL", "FLOWS": "FLOWS", "TWOD": "TWOD", "USE_PML": "USE_PML"}
# pulse_* stand-ins when PULSE is off (initialize.f90:1689-1751)
_PULSE_OFF = {"pulse_t0": 1.0, "pulse_dir": 1, "pulse_t1": 1.0, "pulse_st": 1.0,
              "pulse_px": 1.0, "pulse_py": 1.0, "pulse_pz": 1.0, "pulse_sr": 1.0,
              "pulse_sz": 1.0, "pulse_amp": 0.0}
# PML lines: name -> (width variable, coefficient variable, switch)
_PML = {"xpmlt": ("nxpmltop", "USE_HPML"), "xpmlb": ("nxpmlbot", "USE_HPML"),
        "ypmlt": ("nypmltop", "USE_HPML"), "ypmlb": ("nypmlbot", "USE_HPML"),
        "zpmlt": ("nzpmltop", "USE_PML"), "zpmlb": ("nzpmlbot", "USE_PML")}


def _initial() -> Dict:
    """Module variables before the read (all_modules.f90:25-56 declare them
    without initial values; they start as zero / .false. / blank)."""
    p: Dict = {k: (0 if v == "i" else (0.0 if v == "r" else "")) for k, v in _SIMPLE.items()}
    for v in _SWITCH.values():
        p[v] = False
    p.update(PULSE=False, PSRPULSE=False, MAGNETIC=False, QUIET=False, USE_HPML=False,
             PERIODIC=False, wall_time=0.0, g_xy_filt=0, g_z_filt=0, t_xy=0.0, t_z=0.0,
             block_size=[0, 0, 0])
    for k in _PULSE_OFF:
        p[k] = 0 if k == "pulse_dir" else 0.0
    for name, (nvar, _) in _PML.items():
        p[nvar] = 0
        p[name] = [0.0, 0.0, 0.0, 0.0]
    return p


def read_parameters(source: str, is_text: bool = False) -> Dict:
    """Parse a control file (path, or its text with is_text=True).  Returns
    every variable READ_SPARC_PARAMETERS sets plus the derived counts; the key
    ``stopped_at`` holds the 1-based line whose value failed to read (the
    parse ended there), else None."""
    text = source if is_text else open(source, encoding="latin-1").read()
    p = _initial()
    p["stopped_at"] = None
    for lineno, raw in enumerate(text.split("\n"), 1):
        line = raw.rstrip("\r")[:200]              # CHARACTER(LEN=200) buffer
        eq = line.find("=")
        if eq < 0:
            continue
        name = line[:eq].rstrip(" ")[:30]          # trim(); CHARACTER(LEN=30)
        val = line[eq + 1:]
        try:
            _apply(p, name, val)
        except ParameterError as e:
            p["stopped_at"] = lineno
            p["stop_reason"] = f"line {lineno} ({name}): {e}"
            break
    _finish(p)
    return p


def _switch(val: str) -> int:
    return list_read(val, "i")[0]


def _apply(p: Dict, name: str, val: str) -> None:
    if name in _SIMPLE:
        p[name] = list_read(val, _SIMPLE[name], [p[name]])[0]
    elif name in _SWITCH:
        p[_SWITCH[name]] = _switch(val) != 0
    elif name == "wall_time":
        p["wall_time"] = list_read(val, "r", [p["wall_time"] / 3600.0])[0] * 3600.0
    elif name == "PULSE":
        s = _switch(val)
        p["PULSE"] = s != 0
        if s == 2:
            p["PSRPULSE"] = True
    elif name in ("filter_mode_xy", "filter_mode_z"):
        s = _switch(val)
        key = "g_xy_filt" if name.endswith("xy") else "g_z_filt"
        if s == 1:
            p[key] = 5
        elif s == 2:
            p[key] = 3
    elif name == "MAGNETIC":
        s = _switch(val)
        if s == 0:
            p["MAGNETIC"], p["QUIET"] = False, True
        elif s == 1:
            p["MAGNETIC"], p["QUIET"] = False, False
        elif s == 2:
            p["MAGNETIC"], p["QUIET"] = True, False
    elif name == "USE_HPML":
        s = _switch(val)
        if s == 0:
            # the reference switches the vertical PMLs on here (as written)
            p["USE_PML"] = True
            p["USE_HPML"] = False
        else:
            p["USE_HPML"] = True
    elif name == "PERIODIC":
        s = _switch(val)
        p["PERIODIC"] = False if (s == 0 or p["USE_HPML"]) else True
    elif name in _PML:
        nvar, flag = _PML[name]
        v = list_read(val, "irrrr", [p[nvar]] + list(p[name]))
        p[nvar] = v[0]
        p[name] = list(v[1:])
        if not p[flag]:
            p[nvar] = 0
    elif name in _PULSE_OFF:
        if p["PULSE"]:
            p[name] = list_read(val, "i" if name == "pulse_dir" else "r", [p[name]])[0]
        else:
            p[name] = _PULSE_OFF[name]
    elif name in ("t_xy", "t_z"):
        v = list_read(val, "r", [p[name]])[0]
        p[name] = p["timestep"] if v < p["timestep"] else v
    elif name in ("block_x", "block_y", "block_z"):
        p["block_size"]["xyz".index(name[-1])] = list_read(val, "i")[0]


def _finish(p: Dict) -> None:
    """Post-read adjustments and derived counts (initialize.f90:1778-1924)."""
    if p["TWOD"]:
        p.update(ylength=0.0, ny=1, nypmltop=0, nypmlbot=0, pulse_py=0.0)
    # output-name suffixes: dimension, equation form, physics, flows
    name = p["simname"].rstrip(" ") + ("_2D" if p["TWOD"] else "_3D")
    name += "_d" if p["DISPL"] else "_v"
    name += "m" if p["MAGNETIC"] else ("q" if p["QUIET"] else "t")
    if p["FLOWS"]:
        name += "f"
    p["simname"] = name + "_"
    p["nzpml"] = p["nzpmlbot"] + p["nzpmltop"]
    p["nxpml"] = p["nxpmlbot"] + p["nxpmltop"]
    p["nypml"] = p["nypmlbot"] + p["nypmltop"]
    p["nxc"] = p["xct"] - p["xcb"] + 1
    p["nyc"] = p["yct"] - p["ycb"] + 1
    p["nzc"] = p["zct"] - p["zcb"] + 1
    ts = p["timestep"]
    if ts > 0:
        m = int(math.floor(p["solartime"] * 3600.0 / ts))
        p["maxtime"] = 2 * (m // 2)
        p["fulloutputsteps"] = int(math.floor(p["fulloutputcad"] / ts + 1.0e-5))
        p["chunkoutputsteps"] = int(math.floor(p["chunkoutputcad"] / ts + 1.0e-5))
        p["xyfreq_filtering"] = int(math.floor(p["t_xy"] / ts + 1.0e-5))
        p["zfreq_filtering"] = int(math.floor(p["t_z"] / ts + 1.0e-5))
    else:
        p.update(maxtime=0, fulloutputsteps=0, chunkoutputsteps=0, xyfreq_filtering=0,
                 zfreq_filtering=0)
    p["CHUNKOUT"] = p["chunkoutputsteps"] != 0
    p["FULLOUT"] = p["fulloutputsteps"] != 0
    if not p["CHUNKOUT"]:
        p["chunkoutputsteps"] = 1
    if not p["FULLOUT"]:
        p["fulloutputsteps"] = 1


def to_run_config(p: Dict, precision: str = "fp64") -> RunConfig:
    """The RunConfig of a parsed control file, for the 3-D velocity-form runs
    this package implements: the single pulse (PULSE = 1) or the forcing cube
    of a FITS file (PULSE = 0)."""
    if p["TWOD"]:
        raise ParameterError("2-D runs (RHS2D.f90) are not part of this package")
    if p["DISPL"] or p["FLOWS"]:
        raise ParameterError("displacement form and flows are not part of this package")
    if p["PSRPULSE"]:
        raise ParameterError("the PSR pulse driver (PULSE = 2) is not part of this package")
    if p["g_xy_filt"] == 0 or p["g_z_filt"] == 0:
        raise ParameterError("filter_mode_xy and filter_mode_z must be 1 or 2")
    if p["PERIODIC"] and p["USE_HPML"]:
        # PERIODIC = 1 read before USE_HPML = 1: the reference keeps both on
        # (initialize.f90:1600-1624), a combination this package refuses
        raise ParameterError("PERIODIC = 1 precedes USE_HPML = 1, which leaves both on: periodic "
                             "x/y boundaries with horizontal PMLs are not supported")
    magnetic = 2 if p["MAGNETIC"] else (0 if p["QUIET"] else 1)
    npml = (p["nxpmlbot"], p["nxpmltop"], p["nypmlbot"], p["nypmltop"], p["nzpmlbot"], p["nzpmltop"])
    pmlc = tuple(tuple(p[k]) for k in ("xpmlb", "xpmlt", "ypmlb", "ypmlt", "zpmlb", "zpmlt"))
    return RunConfig(
        nx=p["nx"], ny=p["ny"], nz=p["nz"], xlength=p["xlength"], ylength=p["ylength"],
        timestep=p["timestep"], solartime=p["solartime"], magnetic=magnetic,
        periodic=p["PERIODIC"], use_pml=p["USE_PML"], use_hpml=p["USE_HPML"],
        stabilise=p["STABILISE"], damping=p["DAMPING"], npml=npml, pmlc=pmlc,
        filter_mode_xy=1 if p["g_xy_filt"] == 5 else 2,
        filter_mode_z=1 if p["g_z_filt"] == 5 else 2,
        s_xy=p["s_xy"], s_z=p["s_z"], t_xy=p["t_xy"], t_z=p["t_z"],
        pulse=bool(p["PULSE"]), forcingfunc=str(p.get("forcingfunc") or ""),
        cadforcing=p["cadforcing"], randomexcitedepth=p["randomexcitedepth"],
        randomexcitesigma=p["randomexcitesigma"], randomexciteampl=p["randomexciteampl"],
        pulse_dir=p["pulse_dir"], pulse_t0=p["pulse_t0"], pulse_t1=p["pulse_t1"],
        pulse_st=p["pulse_st"], pulse_px=p["pulse_px"], pulse_py=p["pulse_py"],
        pulse_pz=p["pulse_pz"], pulse_sr=p["pulse_sr"], pulse_sz=p["pulse_sz"],
        pulse_amp=p["pulse_amp"], sp_str=p["sp_str"], sp_length=p["sp_length"],
        sp_decay=p["sp_decay"], maxva=p["maxva"], precision=precision)
