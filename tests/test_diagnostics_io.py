"""Diagnostics and restart files: the HDF5 layer and the writers' layouts
(pic2d.f90 fieldsinfo :1626-1800, particlesinfo :1359-1621, energyinfo
:1996-2115), checked with this package's reader and, independently, with the
HDF5 distribution's h5dump where it is installed."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from pbpic_amd import diagnostics as dg
from pbpic_amd import h5

pytestmark = pytest.mark.skipif(not h5.available(), reason="no HDF5 C library")

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def _dump_header(path):
    """{dataset path: [type, shape]} from h5dump -H (groups tracked by their
    braces)."""
    out = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    res, stack, cur = {}, [], None
    for line in out.splitlines():
        s = line.strip()
        m = re.match(r'(HDF5|GROUP|DATASET) "([^"]+)" \{$', s)
        if m:
            stack.append(m.group(2) if m.group(1) != "HDF5" else "")
            if m.group(1) == "DATASET":
                cur = "/".join(x for x in stack if x not in ("", "/"))
            continue
        if s == "}":
            stack.pop()
            continue
        m = re.match(r"DATATYPE\s+(\S+)", s)
        if m and cur:
            res[cur] = [m.group(1), None]
        m = re.match(r"DATASPACE\s+SIMPLE \{ \( ([^)]*) \)", s)
        if m and cur:
            res[cur][1] = tuple(int(x) for x in m.group(1).split(","))
    return res


def test_h5_roundtrip(tmp_path):
    p = str(tmp_path / "a.h5")
    with h5.File(p, "w") as f:
        f.write("x", np.arange(6, dtype=np.float32).reshape(2, 3))
        f.write("g1/g2/y", np.arange(4, dtype=np.int32))
        f.write("z", np.array([1.25, -2.5]))
    with h5.File(p, "a") as f:
        f.write("x", np.ones((2, 3), np.float32))          # overwrite in place
        f.write("g1/w", np.array([7], np.int64))
        assert "g1/g2/y" in f and "g1/w" in f and "nope" not in f and "g1/nope" not in f
    with h5.File(p, "r") as f:
        assert np.array_equal(f.read("x"), np.ones((2, 3), np.float32))
        assert f.read("g1/g2/y").dtype == np.int32
        assert np.array_equal(f.read("z"), [1.25, -2.5])
        assert f.read("g1/w")[0] == 7


def test_ghosted_layout():
    ny, nx = 3, 4
    plane = np.arange(ny * nx, dtype=np.float64).reshape(ny, nx)
    g = dg.ghosted(plane)
    assert g.shape == (ny + 1, nx + 2)
    assert np.array_equal(g[:ny, 1:nx + 1], plane)          # owned region i = 1..mx
    assert np.array_equal(g[:ny, 0], plane[:, nx - 1])       # emf(0,:) = emf(mx,:)
    assert np.array_equal(g[:ny, nx + 1], plane[:, 0])       # emf(mx+1,:) = emf(1,:)
    assert np.array_equal(g[ny], g[0])                       # row my+1 = row 1


def test_divergence_of_a_curl_and_a_gradient():
    """dive (nodes) of E = curl(psi z), psi at the cell centres, is zero; of
    E = grad(phi), phi at the nodes, it is the 5-point Laplacian of phi."""
    rng = np.random.default_rng(2)
    ny, nx, d = 6, 8, 0.1
    psi = rng.standard_normal((ny, nx))             # psi[gy, gx] at (gx+1/2, gy+1/2)
    ex = (psi - np.roll(psi, 1, axis=0)) / d        # Ex at (gx+1/2, gy)
    ey = -(psi - np.roll(psi, 1, axis=1)) / d       # Ey at (gx, gy+1/2)
    dv = dg.divergence(ex, ey, d, d)
    assert np.max(np.abs(dv)) < 1e-12 and not dv[:, 0].any()
    phi = rng.standard_normal((ny, nx))             # phi at the nodes
    ex = (np.roll(phi, -1, axis=1) - phi) / d
    ey = (np.roll(phi, -1, axis=0) - phi) / d
    lap = (np.roll(phi, -1, 1) - 2 * phi + np.roll(phi, 1, 1)) / d**2 + \
          (np.roll(phi, -1, 0) - 2 * phi + np.roll(phi, 1, 0)) / d**2
    assert np.allclose(dg.divergence(ex, ey, d, d)[:ny, 1:nx + 1], lap, rtol=1e-12, atol=1e-9)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_reference_file_layouts(tmp_path):
    """Names, types and shapes as the reference's writers create them (the
    Fortran dims reversed): fields (ny+1, nx+2) float32 + num int32,
    particles 1-D float32/int32 + nptlact, spectra (nbins, 2) int32 per
    %05d group + ebins float32, tote (ntime+1, 4) float32."""
    ny, nx, n = 5, 7, 33
    rng = np.random.default_rng(1)
    planes = {k: rng.standard_normal((ny, nx)) for k in dg.FIELD_NAMES}
    fp = str(tmp_path / "fields00010.h5")
    dg.write_fields(fp, planes, rng.integers(0, 9, (ny, nx)), 0.1, 0.1)
    hd = _dump_header(fp)
    for k in dg.FIELD_NAMES + ("dive",):
        assert hd[k] == ["H5T_IEEE_F32LE", (ny + 1, nx + 2)], k
    assert hd["num"] == ["H5T_STD_I32LE", (ny + 1, nx + 2)]
    back = dg.read_fields(fp)
    for k in dg.FIELD_NAMES:
        assert np.array_equal(back[k], planes[k].astype(np.float32).astype(np.float64))
    pp = str(tmp_path / "particles.h5")
    parts = dict(grid=rng.integers(1, 36, n), dx=rng.random(n) - 0.5, dy=rng.random(n) - 0.5,
                 ux=rng.standard_normal(n), uy=rng.standard_normal(n), uz=rng.standard_normal(n),
                 q=np.where(rng.random(n) < 0.5, 1, -1), pid=np.arange(1, n + 1))
    dg.write_particles(pp, parts)
    hd = _dump_header(pp)
    for k in ("dx", "dy", "ux", "uy", "uz"):
        assert hd[k] == ["H5T_IEEE_F32LE", (n,)]
    for k in ("grid", "q", "pid"):
        assert hd[k] == ["H5T_STD_I32LE", (n,)]
    assert hd["nptlact"] == ["H5T_STD_I32LE", (1,)]
    back = dg.read_particles(pp)
    assert np.array_equal(back["grid"], parts["grid"]) and np.array_equal(back["q"], parts["q"])
    sp = str(tmp_path / "spectra.h5")
    for it in (0, 10):
        dg.write_spectra(sp, it, np.full(dg.NBINS, it, np.int32), np.arange(dg.NBINS, dtype=np.int32))
    hd = _dump_header(sp)
    assert hd["00000/spectra"] == ["H5T_STD_I32LE", (dg.NBINS, 2)]
    assert hd["00010/spectra"] == ["H5T_STD_I32LE", (dg.NBINS, 2)]
    assert hd["ebins"] == ["H5T_IEEE_F32LE", (dg.NBINS, 2)]
    with h5.File(sp, "r") as f:
        eb = f.read("ebins")
        assert np.isclose(eb[0, 0], 0.005) and np.isclose(eb[-1, 0], 5.0) and np.isclose(eb[-1, 1], 25.0)
        assert f.read("00010/spectra")[3, 0] == 10 and f.read("00010/spectra")[3, 1] == 3
    ep = str(tmp_path / "energy_tot.h5")
    tot = rng.random((31, 4))
    dg.write_energy(ep, tot)
    dg.write_energy(ep, 2 * tot)
    assert _dump_header(ep)["tote"] == ["H5T_IEEE_F32LE", (31, 4)]
    with h5.File(ep, "r") as f:
        assert np.array_equal(f.read("tote"), (2 * tot).astype(np.float32))
