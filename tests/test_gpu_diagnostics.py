"""GPU diagnostics and restart: energyspectra (pic2d.f90:1957-1990) and
numdensity (:1343-1354) from the device against numpy on the same particles
(bit-exact: integer counts of FP64 energies computed by the same
expression), the reference-format files written from a run, and exact
checkpoint/restart."""
import os

import numpy as np
import pytest

from cases import gpu_for, make_case, rel_norm
from pbpic_amd import diagnostics as dg
from pbpic_amd import h5

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _spectrum_np(p, ms, nbins, de, emax):
    ux, uy, uz = (np.asarray(p[k], np.float64) for k in ("ux", "uy", "uz"))
    ene = (np.sqrt(1.0 + ux * ux + uy * uy + uz * uz) - 1.0) * ms
    e = ene[ene < emax]
    k = np.clip(np.ceil(e / de) - 1, 0, nbins - 1).astype(np.int64)
    return np.bincount(k, minlength=nbins)


@pytest.mark.parametrize("walls", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_energy_spectrum_and_density(gpu, precision, walls):
    """With conducting walls the guard rows hold no particles: their counts
    are zero."""
    _say(f"spectrum/density precision={precision} walls={walls}")
    case = make_case(nx=40, ny=36, ppc=20, seed=4, vth=(0.1, 0.5), walls=walls)
    ctx = gpu_for(case, precision=precision, compat=0)
    for stage in ("loaded", "pushed"):
        if stage == "pushed":
            ctx.zero_current(); ctx.pusher(); ctx.fieldsolver()
        for s, (q, ms) in enumerate(((1, 25.0), (-1, 1.0))):
            p = ctx.get_particles(s, native=True)
            de, emax = (0.025, 25.0) if q > 0 else (0.005, 5.0)
            got = ctx.energy_spectrum(s)
            assert np.array_equal(got, _spectrum_np(p, ms, 1000, de, emax)), (stage, s)
            # a coarse custom binning too
            assert np.array_equal(ctx.energy_spectrum(s, 37, 0.01, 0.3),
                                  _spectrum_np(p, ms, 37, 0.01, 0.3)), (stage, s)
        nd = np.zeros(case["nx"] * case["ny"], np.int64)
        for s in range(2):
            nd += np.bincount(ctx.get_particles(s)["grid"] - 1, minlength=nd.size)
        assert np.array_equal(ctx.number_density(), nd.reshape(case["nx"], case["ny"]).T), stage
        if walls:
            assert not ctx.number_density()[case["ny"] - 2:].any(), stage


def test_reference_files_from_a_run(gpu, tmp_path):
    """20 steps with totene recorded each step and the three outputs at steps
    10 and 20: the files hold the context's values (float32)."""
    _say("diagnostics files")
    case = make_case(nx=24, ny=20, ppc=8, seed=6)
    ctx = gpu_for(case, precision=0, compat=7, sort_every=10)
    d = dg.Diagnostics(ctx, ntime=20, path=str(tmp_path))
    ctx.countingsort()
    d.record(0)
    for it in range(1, 21):
        ctx.step(1)
        d.record(it)
        if it % 10 == 0:
            d.fieldsinfo(it)
            d.energyinfo(it)
            d.particlesinfo()
    f = dg.read_fields(str(tmp_path / "fields00020.h5"))
    for k in dg.FIELD_NAMES:
        assert np.array_equal(f[k], ctx.get_field(k).astype(np.float32).astype(np.float64)), k
    with h5.File(str(tmp_path / "fields00020.h5"), "r") as fh:
        num = fh.read("num")
        assert np.array_equal(num[:-1, 1:-1], ctx.number_density())
        assert num.sum() == sum(ctx.num_particles(s) for s in range(2))
    with h5.File(str(tmp_path / "energy_tot.h5"), "r") as fh:
        tote = fh.read("tote")
        assert tote.shape == (21, 4)
        assert np.array_equal(tote, d.totene.astype(np.float32))
        e, b = ctx.fieldsenergy()
        assert np.isclose(tote[20, 2], e, rtol=1e-6) and np.isclose(tote[20, 3], b, rtol=1e-6)
    with h5.File(str(tmp_path / "spectra.h5"), "r") as fh:
        assert "00010/spectra" in fh and "00020/spectra" in fh and "ebins" in fh
        sp = fh.read("00020/spectra")
        assert np.array_equal(sp[:, 0], ctx.energy_spectrum(1))
        assert np.array_equal(sp[:, 1], ctx.energy_spectrum(0))
    p = dg.read_particles(str(tmp_path / "particles.h5"))
    g0 = ctx.get_particles(0)
    n0 = len(g0["grid"])
    assert np.array_equal(p["grid"][:n0], g0["grid"]) and np.all(p["q"][:n0] == 1)
    assert np.array_equal(p["ux"][:n0], g0["ux"].astype(np.float32))


@pytest.mark.parametrize("walls", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
def test_checkpoint_restart(gpu, tmp_path, precision, walls):
    """save_checkpoint after 12 steps, load into a fresh context: identical
    state (fields, particles at the device precision, step count, J.E); 8 more
    steps from both agree (FP64 to 1e-12; the split form to FP32 rounding),
    including the sort at step 20 that the restored step count schedules."""
    _say(f"checkpoint precision={precision} walls={walls}")
    case = make_case(nx=24, ny=20, ppc=8, seed=9, walls=walls)
    ctx = gpu_for(case, precision=precision, compat=0, sort_every=10)
    ctx.countingsort()
    ctx.step(12)
    path = str(tmp_path / "ck.h5")
    dg.save_checkpoint(ctx, path)
    ctx2 = gpu_for(make_case(nx=24, ny=20, ppc=8, seed=99, walls=walls), precision=precision,
                   compat=0, sort_every=10)
    dg.load_checkpoint(ctx2, path)
    assert ctx2.counters() == ctx.counters()
    for k in ("ex", "ey", "ez", "bx", "by", "bz"):
        assert np.array_equal(ctx2.get_field(k), ctx.get_field(k)), k
    for s in range(2):
        a, b = ctx.get_particles(s, native=True), ctx2.get_particles(s, native=True)
        for k in ("grid", "dx", "dy", "ux", "uy", "uz", "pid"):
            assert np.array_equal(a[k], b[k]), (s, k)
    ctx.step(8)
    ctx2.step(8)
    d1, d2 = ctx.diag(), ctx2.diag()
    tol = 1e-12 if precision == 0 else 1e-4
    for key in ("ke_e", "ke_i", "w_e", "w_b", "je_int"):
        assert abs(d1[key] - d2[key]) <= tol * abs(d1[key]), (key, d1[key], d2[key])
    assert d1["steps"] == d2["steps"] == 20
    for k in ("ex", "bz"):
        assert rel_norm(ctx2.get_field(k), ctx.get_field(k)) < tol, k


def test_restart_from_reference_files(gpu, tmp_path):
    """fieldsinfo(0) + particlesinfo(0): a new context from the float32 files
    holds the float32-rounded state."""
    _say("restart from reference files")
    case = make_case(nx=24, ny=20, ppc=8, seed=2)
    ctx = gpu_for(case, precision=0, compat=7, sort_every=0)
    ctx.step(3)
    d = dg.Diagnostics(ctx, ntime=3, path=str(tmp_path))
    d.fieldsinfo(3)
    d.particlesinfo()
    ctx2 = gpu_for(make_case(nx=24, ny=20, ppc=8, seed=5), precision=0, compat=7)
    dg.restart_from_reference_files(ctx2, str(tmp_path / "fields00003.h5"),
                                    str(tmp_path / "particles.h5"))
    for k in ("ex", "ey", "ez", "bx", "by", "bz"):
        assert np.array_equal(ctx2.get_field(k), ctx.get_field(k).astype(np.float32).astype(np.float64))
    for s in range(2):
        a, b = ctx.get_particles(s), ctx2.get_particles(s)
        assert np.array_equal(a["grid"], b["grid"]) and np.array_equal(a["pid"], b["pid"])
        assert np.array_equal(b["uz"], a["uz"].astype(np.float32).astype(np.float64))
