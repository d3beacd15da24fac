"""GPU parity with conducting y walls and the Harris-sheet loader, against the
oracle on the same seeded inputs (the walls have no reference counterpart:
pic2d.f90 is periodic in y and comments its conducting boundary out at
:850-859; the loader restates initparticles/initfields, :357-648).

Bars: FP64 one step to 1e-12 (particles element-wise by id, currents and
fields norm-wise over every row, guards included); split precision to the
FP32 bars of the periodic tests; the loader bit-for-bit in cells except where
a position lies within 1e-9 of a cell face (device and host logarithms may
round differently), offsets to 1e-12, momenta to 1e-13.
"""
import numpy as np
import pytest

import oracle
from cases import SPECIES_Q, gpu_for, make_case, oracle_for, oracle_species, rel_norm

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _by_pid(p):
    o = np.argsort(p["pid"])
    return {k: v[o] for k, v in p.items()}


def _wall_case(seed, nx=24, ny=18, ppc=12, vth=(0.1, 0.45)):
    return make_case(nx=nx, ny=ny, ppc=ppc, seed=seed, efield=0.3, bfield=1.2, vth=vth,
                     walls=True)


@pytest.mark.parametrize("precision", [0, 1])
def test_one_step_walls(gpu, precision):
    """push + field solve with reflections at both walls."""
    _say(f"walls one step precision={precision}")
    case = _wall_case(seed=21)
    ctx = gpu_for(case, precision=precision, compat=0)
    o = oracle_for(case, compat=0)
    for f in ("ex", "ey", "ez", "bx", "by", "bz"):             # images applied alike
        assert np.array_equal(ctx.get_field(f), o.get_field(f)), f
    ctx.countingsort(); o.countingsort()
    ctx.zero_current(); ctx.pusher()
    o.zero_current(); o.pusher()
    tol = 1e-12 if precision == 0 else 1e-5
    for f in ("jx", "jy", "jz"):
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < tol, f
    osp = oracle_species(o, case)
    nref = 0
    for k in range(2):
        g, r = _by_pid(ctx.get_particles(k)), _by_pid(osp[k])
        nyp = case["ny"] - 2
        assert np.all((g["grid"] - 1) % case["ny"] < nyp)
        nref += int(np.sum(np.sign(g["uy"]) != np.sign(case["species"][k]["uy"][np.argsort(
            case["species"][k]["pid"])])))
        if precision == 0:
            assert np.array_equal(g["grid"], r["grid"])
            for f in ("dx", "dy", "ux", "uy", "uz"):
                assert np.max(np.abs(g[f] - r[f])) < 1e-12 * max(1.0, np.max(np.abs(r[f]))), f
        else:
            same = g["grid"] == r["grid"]
            near = (np.abs(np.abs(r["dx"]) - 0.5) < 1e-5) | (np.abs(np.abs(r["dy"]) - 0.5) < 1e-5)
            assert np.all(same | near)
            for f in ("dx", "dy"):
                assert np.max(np.abs(g[f] - r[f])[same]) < 2e-6, f
    assert nref > 0                                   # some particles did reflect
    ctx.fieldsolver(); o.fieldsolver()
    for f in ("ex", "ey", "ez", "bx", "by", "bz"):
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < tol, f


def test_multi_step_walls_fp64(gpu):
    """30 steps with a sort every 10: energies, J.E and the Gauss residual
    between the walls track the oracle; the residual's drift stays at
    rounding (charge conservation with reflections)."""
    _say("walls 30 steps fp64")
    case = _wall_case(seed=8, ppc=10)
    ctx = gpu_for(case, precision=0, compat=0, sort_every=10)
    o = oracle_for(case, compat=0)
    ctx.countingsort(); o.countingsort()
    d0 = ctx.diag()
    o.diag()
    ctx.step(30)
    o.step(30, sort_every=10)
    d, od = ctx.diag(), o.diag()
    for key in ("ke_e", "ke_i", "w_e", "w_b", "je_int", "gauss_l2"):
        assert abs(d[key] - od[key]) <= 1e-9 * abs(od[key]), (key, d[key], od[key])
    assert d["gauss_dlinf"] < 1e-10 * d0["gauss_linf"]


@pytest.mark.parametrize("precision", [0, 1])
def test_cavity_mode_walls(gpu, precision):
    """Vacuum TM eigenmode Ez = sin(kx x) sin(ky y), ky = pi/Ly: 60 field
    solves give cos(n w dt) Ez^0 with the discrete Yee frequency."""
    import pbpic_amd
    _say(f"walls cavity precision={precision}")
    nx, nyp, d, cva = 32, 20, 0.1, 15.0
    dt = 0.5 * d / cva
    ctx = pbpic_amd.Context(nx, nyp + 2, d, d, dt, cva, 25.0, 1.0, precision=precision,
                            compat=0, ybound="conducting")
    kx, ky = 2 * np.pi * 2 / (nx * d), np.pi / (nyp * d)
    x = (np.arange(nx) + 0.5) * d
    y = (np.arange(nyp + 2) + 0.5) * d
    ctx.set_field("ez", np.sin(kx * x)[None, :] * np.sin(ky * y)[:, None])
    ez0 = ctx.get_field("ez")
    for _ in range(60):
        ctx.zero_current()
        ctx.fieldsolver()
    s = cva * dt * np.sqrt(np.sin(kx * d / 2) ** 2 + np.sin(ky * d / 2) ** 2) / d
    c = np.cos(60 * 2 * np.arcsin(s))
    err = np.max(np.abs(ctx.get_field("ez")[:nyp] - c * ez0[:nyp]))
    assert err < (1e-11 if precision == 0 else 1e-5), err


HARRIS = {
    "single_y": dict(geometry=1, ppc=16, half_width=0.5, nb=0.2, b0=1.0, psi0=0.1,
                     vthe_cs=0.1443, vthi_cs=0.0645, vthe_bg=0.1443, vthi_bg=0.0645,
                     vde_z=0.0333, vdi_z=-0.1667, thermal_z=1, seed=5, stream=2),
    "double_x": dict(geometry=0, ppc=16, half_width=0.5, nb=0.2, b0=1.0, psi0=0.0,
                     vthe_cs=0.092, vthi_cs=0.0411, vthe_bg=0.0291, vthi_bg=0.0058,
                     vde_z=-0.0667, vdi_z=0.0667, thermal_z=0, seed=9, stream=0),
}


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("geometry", ["single_y", "double_x"])
def test_harris_loader_matches_oracle(gpu, geometry, precision):
    """The device loader against the oracle's; in the split form the
    offsets and momenta are the oracle's rounded to FP32."""
    import pbpic_amd
    _say(f"harris loader {geometry} precision={precision}")
    h = HARRIS[geometry]
    walls = geometry == "single_y"
    nx, ny, d = 64, (34 if walls else 32), 0.1
    ions, ele, fields, ncs = oracle.init_harris(nx, ny, walls, d, d, h)
    n = len(ions["grid"])
    ctx = pbpic_amd.Context(nx, ny, d, d, 0.001, 10.0, 25.0, 0.01, track_pid=True,
                            precision=precision, ybound="conducting" if walls else "periodic")
    si = ctx.add_species(1, n)
    se = ctx.add_species(-1, n)
    ctx.init_harris(si, se, **h)
    for k, ref in ((si, ions), (se, ele)):
        g = ctx.get_particles(k)
        assert np.array_equal(g["pid"], np.arange(1, n + 1))
        x_ref = (ref["grid"] - 1) // ny + 0.5 + ref["dx"]
        y_ref = (ref["grid"] - 1) % ny + 0.5 + ref["dy"]
        x_gpu = (g["grid"] - 1) // ny + 0.5 + g["dx"]
        y_gpu = (g["grid"] - 1) % ny + 0.5 + g["dy"]
        same = g["grid"] == ref["grid"]
        face = (np.abs(np.abs(ref["dx"]) - 0.5) < 1e-9) | (np.abs(np.abs(ref["dy"]) - 0.5) < 1e-9)
        assert np.all(same | face)
        # positions to 1e-12 cells (FP64) or FP32 rounding of the offsets
        tol_x, tol_u = (1e-12, 1e-13) if precision == 0 else (1e-6, 1e-6)
        assert np.max(np.abs(x_gpu - x_ref)) < tol_x and np.max(np.abs(y_gpu - y_ref)) < tol_x
        for f in ("ux", "uy", "uz"):
            assert np.max(np.abs(g[f] - ref[f])) < tol_u * max(1.0, np.max(np.abs(ref[f]))), f
    for f, plane in fields.items():
        assert np.max(np.abs(ctx.get_field(f) - plane)) < 1e-14, f


@pytest.mark.parametrize("precision", [0, 1])
def test_harris_sheet_run(gpu, precision):
    """A C1-shaped sheet (walls, psi perturbation) at 128 x 66 for 200 steps:
    particles stay between the walls, total energy holds within 1% and the
    Gauss residual's drift stays at the precision's rounding."""
    import pbpic_amd
    _say(f"harris run precision={precision}")
    nx, nyp = 128, 64
    d, cva, mime = 0.05, 10.0, 25.0
    dt = 0.6 * d / cva
    h = dict(HARRIS["single_y"], ppc=16)
    ions, _, _, ncs = oracle.init_harris(nx, nyp + 2, True, d, d, h)
    n = len(ions["grid"])
    peak = ncs * d * d / (2 * 0.5 * nx * d)          # sheet particles per cell at n0
    ctx = pbpic_amd.Context(nx, nyp + 2, d, d, dt, cva, mime, 1.0 / peak, precision=precision,
                            compat=0, sort_every=25, ybound="conducting")
    si = ctx.add_species(1, n)
    se = ctx.add_species(-1, n)
    ctx.init_harris(si, se, **h)
    ctx.countingsort()
    tot = lambda dd: dd["ke_e"] + dd["ke_i"] + dd["w_e"] + dd["w_b"]
    d0 = ctx.diag()
    ctx.step(200)
    d1 = ctx.diag()
    assert abs(tot(d1) - tot(d0)) < 1e-2 * tot(d0), (tot(d0), tot(d1))
    assert d1["gauss_dlinf"] < (1e-9 if precision == 0 else 1e-4) * d0["gauss_linf"]
    for k in (si, se):
        g = ctx.get_particles(k)
        assert np.all((g["grid"] - 1) % (nyp + 2) < nyp)
