"""The reference's HDF5 layouts (bkg.h5, full_NNNN.h5, chunk_NNNN.h5).

The layout is checked with h5dump (the HDF5 tools' own reader, independent
of this code): dataset names, dataspaces in the reversed order the Fortran
API gives a Fortran array (a(nx,ny,nz) shows as (nz, ny, nx)), and the root
attributes xyzdims and time, against what all_modules.f90 writes
(1112-1423, 1433-1550, 1601-1799).  Round trips go through the C library.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from sparc2_amd import h5io
from sparc2_amd.background import make_background
from tests.cases import mhd_cfg, quiet_cfg, random_state

pytestmark = pytest.mark.skipif(not os.path.exists(h5io.LIB), reason="libsparc_h5.so not built (no HDF5)")

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def _dump(path):
    if H5DUMP is None:
        pytest.skip("h5dump not available")
    return subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout


def _spaces(text):
    """{dataset: extents} from an h5dump -H header listing."""
    out = {}
    for name, dims in re.findall(r'DATASET "(\w+)" \{.*?DATASPACE\s+SIMPLE \{ \( ([\d, ]+) \)', text, re.S):
        out[name] = tuple(int(v) for v in dims.split(","))
    return out


class _Stub:
    """The part of Solver the file functions use, holding arrays on the host."""

    def __init__(self, cfg, st, a, psi):
        self.cfg, self.st, self.a, self.psi = cfg, st, a, psi

    def get_state(self, with_psi=False):
        return (self.a.copy(), {f: v.copy() for f, v in self.psi.items()}) if with_psi else self.a.copy()

    def put_state(self, a, psi=None):
        self.a = np.array(a)
        self.psi = {f: np.array(v) for f, v in (psi or {}).items()}

    def get_slab(self, field, k0, k1):
        return self.a[field, k0:k1 + 1].copy()


def _stub(cfg, seed=3):
    from sparc2_amd.initialise import prepare
    bg = make_background(cfg, -5.0, 1.0, R0=6e8) if cfg.mhd else make_background(cfg, -10.0, 0.5)
    a, psi = random_state(cfg, seed)
    return _Stub(cfg, prepare(cfg, bg), a, psi), bg


def test_primitives_layout_and_round_trip(tmp_path):
    p = str(tmp_path / "t.h5")
    h5io.create(p, (7, 5, 3), 1.25)
    x = np.arange(3 * 5 * 7, dtype=np.float64).reshape(3, 5, 7)   # Fortran x(7,5,3)
    h5io.write(p, "f", x)
    assert h5io.attrs(p) == ((7, 5, 3), 1.25)
    assert h5io.has(p, "f") and not h5io.has(p, "g")
    np.testing.assert_array_equal(h5io.read(p, "f", (3, 5, 7)), x)
    with pytest.raises(h5io.H5Error):
        h5io.read(p, "f", (5, 3, 7))          # other extents
    text = _dump(p)
    assert _spaces(text) == {"f": (3, 5, 7)}
    assert re.search(r'ATTRIBUTE "xyzdims".*?\(0\): 7, 5, 3', _dump_full(p), re.S)
    assert re.search(r'ATTRIBUTE "time".*?\(0\): 1.25', _dump_full(p), re.S)


def _dump_full(path):
    return subprocess.run([H5DUMP, "-A", path], capture_output=True, text=True, check=True).stdout


@pytest.mark.parametrize("mhd", [False, True])
def test_bkg_round_trip_and_layout(tmp_path, mhd):
    cfg = mhd_cfg(nx=24, ny=20, nz=18) if mhd else quiet_cfg(nx=24, ny=20, nz=18)
    bg = make_background(cfg, -5.0, 1.0, R0=6e8) if mhd else make_background(cfg, -10.0, 0.5)
    p = str(tmp_path / "bkg.h5")
    h5io.write_bkg(p, cfg, bg)
    back = h5io.read_bkg(p, cfg)
    for k, v in back.items():
        np.testing.assert_array_equal(v, bg[k])
    sp = _spaces(_dump(p))
    want = {"rho0": (18, 20, 24), "pre0": (18, 20, 24), "cs0": (18, 20, 24),
            "zaxis": (18,), "grav": (18,)}
    if mhd:
        want.update({"bx0": (18, 20, 24), "by0": (18, 20, 24), "bz0": (18, 20, 24)})
    assert sp == want
    with pytest.raises(h5io.H5Error):
        h5io.read_bkg(p, cfg.replace(nx=26))


@pytest.mark.parametrize("mhd", [False, True])
def test_full_snapshot_layout_and_restart(tmp_path, mhd):
    cfg = (mhd_cfg(nx=24, ny=20, nz=18, npml=(4, 5, 4, 5, 4, 6)) if mhd
           else quiet_cfg(nx=24, ny=20, nz=18, npml=(0, 0, 0, 0, 4, 6)))
    s, _ = _stub(cfg)
    name = h5io.file_name("run_", "full", 37)
    assert name == "run_full_0037.h5"
    p = str(tmp_path / name)
    h5io.write_full(p, s, time=37 * 12)
    dims, hours = h5io.attrs(p)
    assert dims == (24, 20, 18)
    assert hours == pytest.approx(37 * 12 * cfg.timestep / 3600.0, rel=1e-15)
    sp = _spaces(_dump(p))
    names = ["pre", "rho", "vx", "vy", "vz"] + (["bx", "by", "bz"] if mhd else [])
    for n in names:
        assert sp[n] == (18, 20, 24)
    npsi = 6 if mhd else 2
    # PML_Z*: Fortran (nx, ny, nzpml, pmlsize); PML_X*: (nxpml, ny, nz, pmlsize);
    # PML_Y*: (nx, nypml, nz, pmlsize) (all_modules.f90:1244-1420)
    assert sp["PML_ZBOT"] == (npsi, 4, 20, 24) and sp["PML_ZTOP"] == (npsi, 6, 20, 24)
    if mhd:
        assert sp["PML_XBOT"] == (npsi, 18, 20, 4) and sp["PML_XTOP"] == (npsi, 18, 20, 5)
        assert sp["PML_YBOT"] == (npsi, 18, 4, 24) and sp["PML_YTOP"] == (npsi, 18, 5, 24)
    else:
        assert not any(k.startswith("PML_X") or k.startswith("PML_Y") for k in sp)
    assert len(sp) == len(names) + len(s.psi)
    # physical units in the file: velocity in cm/s, density in g/cm^3
    np.testing.assert_allclose(h5io.read(p, "vx", (18, 20, 24)), s.a[1] * s.st.dimc, rtol=1e-15)
    np.testing.assert_allclose(h5io.read(p, "rho", (18, 20, 24)), s.a[0] * s.st.dimrho, rtol=1e-15)
    # restart: the state comes back to within one rounding of the scaling
    r = _Stub(cfg, s.st, None, None)
    assert h5io.read_full(p, r) == hours
    np.testing.assert_allclose(r.a, s.a, rtol=4e-16, atol=0)
    for f, v in s.psi.items():
        np.testing.assert_array_equal(r.psi[f], v)


def test_chunk_layout(tmp_path):
    cfg = mhd_cfg(nx=24, ny=20, nz=18)
    s, _ = _stub(cfg)
    p = str(tmp_path / h5io.file_name("run_", "chunk", 3))
    h5io.write_chunk(p, s, time=90, xr=(2, 9), yr=(0, 19), zr=(12, 12))
    dims, hours = h5io.attrs(p)
    assert dims == (8, 20, 1)
    sp = _spaces(_dump(p))
    assert sp["xaxis"] == (8,) and sp["yaxis"] == (20,) and sp["zaxis"] == (1,)
    for n in ("vx", "vy", "vz", "rho", "pre", "bx", "by", "bz"):
        assert sp[n] == (1, 20, 8)
    vz = h5io.read(p, "vz", (1, 20, 8))
    np.testing.assert_allclose(vz, s.a[3, 12:13, :, 2:10] * s.st.dimc, rtol=1e-15)
    # axes: x in cm, z in Mm above the surface
    np.testing.assert_allclose(h5io.read(p, "xaxis", (8,)), s.st.xaxis[2:10] * cfg.xlength, rtol=1e-15)
    z = h5io.read(p, "zaxis", (1,))[0]
    assert -5.0 <= z <= 1.0
