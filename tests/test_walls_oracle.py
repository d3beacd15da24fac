"""Pins the oracle's conducting y walls and Harris-sheet loader (no reference
counterpart: pic2d.f90 is periodic in y and leaves its conducting boundary
commented out at :850-859; the loader restates initparticles/initfields,
:357-648, with a seeded generator).

  field images (wall_image)       cavity eigenmodes of both polarizations
                                  oscillate as cos(n w dt) with the discrete
                                  Yee dispersion for k_y = m pi / L_y
  particle reflection (push_one)  known answer at both walls
  Jz fold (finish_current)        equals the deposit of the particles plus
                                  their mirror images (charge -q) in a
                                  periodic domain of twice the height
  walls together                  Gauss law conserved while particles reflect
  Harris loader                   sheet fraction, sech^2 profile (KS test),
                                  pairing, drift and thermal moments, div B = 0,
                                  curl B of the tanh profile
"""
import numpy as np
import pytest

import oracle
from cases import make_case, oracle_for

CVA, MIME = 15.0, 25.0


def _vacuum(nx, nyp, d, dt):
    return oracle.Oracle(nx, nyp + 2, d, d, dt, CVA, MIME, 1.0, compat=0, ybound=1)


def _omega_dt(kx, ky, d, dt):
    s = CVA * dt * np.sqrt(np.sin(kx * d / 2) ** 2 + np.sin(ky * d / 2) ** 2) / d
    return 2 * np.arcsin(s)


@pytest.mark.parametrize("m", [1, 2, 5])
def test_cavity_tm_mode(oracle_lib, m):
    """Ez = sin(kx x) sin(ky y) at the Ez points (cell centres), B = 0: with
    the odd images of Ez across the walls this is an eigenvector of the
    discrete curl-curl, so Ez^n = cos(n w dt) Ez^0 to rounding."""
    nx, nyp, d = 16, 12, 0.1
    dt = 0.5 * d / CVA
    o = _vacuum(nx, nyp, d, dt)
    kx, ky = 2 * np.pi * 2 / (nx * d), m * np.pi / (nyp * d)
    x = (np.arange(nx) + 0.5) * d
    y = (np.arange(nyp + 2) + 0.5) * d
    ez0 = np.sin(kx * x)[None, :] * np.sin(ky * y)[:, None]
    ez0[nyp:] = 0.0                                   # guard rows come from the images
    o.set_field("ez", ez0)
    ez0 = o.get_field("ez")
    assert np.allclose(ez0[nyp + 1], -ez0[0]) and np.allclose(ez0[nyp], -ez0[nyp - 1])
    wdt = _omega_dt(kx, ky, d, dt)
    for n in range(1, 61):
        o.zero_current()
        o.fieldsolver()
        if n % 20 == 0:
            ez = o.get_field("ez")
            assert np.max(np.abs(ez[:nyp] - np.cos(n * wdt) * ez0[:nyp])) < 1e-11, n


@pytest.mark.parametrize("m", [1, 3])
def test_cavity_te_mode(oracle_lib, m):
    """E = curl(psi z) with psi = cos(kx x) cos(ky y) at the cell centres (so
    div E = 0 discretely and Ex = 0 on the walls), B = 0: Ex, Ey oscillate as
    cos(n w dt) with the images Ey even, Bz even, Ex zero on the walls."""
    nx, nyp, d = 16, 10, 0.1
    dt = 0.5 * d / CVA
    o = _vacuum(nx, nyp, d, dt)
    kx, ky = 2 * np.pi * 3 / (nx * d), m * np.pi / (nyp * d)
    psi = lambda xx, yy: np.cos(kx * xx) * np.cos(ky * yy)
    gx = np.arange(nx)[None, :]
    gy = np.arange(nyp + 2)[:, None]
    # Ex at (gx+1/2, gy), Ey at (gx, gy+1/2)
    ex0 = (psi((gx + 0.5) * d, (gy + 0.5) * d) - psi((gx + 0.5) * d, (gy - 0.5) * d)) / d
    ey0 = -(psi((gx + 0.5) * d, (gy + 0.5) * d) - psi((gx - 0.5) * d, (gy + 0.5) * d)) / d
    o.set_field("ex", ex0)
    o.set_field("ey", ey0)
    ex0, ey0 = o.get_field("ex"), o.get_field("ey")
    assert np.max(np.abs(ex0[0])) == 0 and np.max(np.abs(ex0[nyp])) == 0
    wdt = _omega_dt(kx, ky, d, dt)
    scale = np.max(np.abs(ey0))
    for n in range(1, 41):
        o.zero_current()
        o.fieldsolver()
    c = np.cos(40 * wdt)
    assert np.max(np.abs(o.get_field("ex")[:nyp + 1] - c * ex0[:nyp + 1])) < 1e-11 * scale
    assert np.max(np.abs(o.get_field("ey")[:nyp] - c * ey0[:nyp])) < 1e-11 * scale


def test_wrong_image_sign_breaks_the_mode(oracle_lib):
    """Sanity of the mode test itself: an Ez pattern that is even across the
    walls (cos instead of sin) is no eigenvector under odd images, so it does
    not oscillate as a single cosine."""
    nx, nyp, d = 16, 12, 0.1
    dt = 0.5 * d / CVA
    o = _vacuum(nx, nyp, d, dt)
    kx, ky = 2 * np.pi * 2 / (nx * d), 1 * np.pi / (nyp * d)
    x = (np.arange(nx) + 0.5) * d
    y = (np.arange(nyp + 2) + 0.5) * d
    ez0 = np.sin(kx * x)[None, :] * np.cos(ky * y)[:, None]
    o.set_field("ez", ez0)
    ez0 = o.get_field("ez")
    for _ in range(20):
        o.zero_current()
        o.fieldsolver()
    c = np.cos(20 * _omega_dt(kx, ky, d, dt))
    assert np.max(np.abs(o.get_field("ez")[:nyp] - c * ez0[:nyp])) > 1e-3


@pytest.mark.parametrize("wall", ["bottom", "top"])
def test_reflection_known_answer(oracle_lib, wall):
    """No fields: a particle 0.2 cells from a wall moving 0.4 cells towards it
    ends 0.2 cells from the wall in the same cell, u_y reversed, u_x, u_z kept."""
    nx, nyp, d = 8, 6, 0.1
    ny = nyp + 2
    dt = 0.005
    o = oracle.Oracle(nx, ny, d, d, dt, CVA, MIME, 0.1, compat=0, ybound=1)
    ux, uz = 0.01, -0.02
    # v_y dt = 0.4 cells: u_y / gamma = a with gamma = sqrt(1 + u^2)
    a = 0.4 * d / (dt * CVA)
    uy = a * np.sqrt((1 + ux * ux + uz * uz) / (1 - a * a))
    cy, dy0, sgn = (0, -0.3, -1.0) if wall == "bottom" else (nyp - 1, 0.3, 1.0)
    grid = 3 * ny + cy + 1
    o.set_particles([grid], [0.1], [dy0], [ux], [sgn * uy], [uz], [-1])
    o.zero_current()
    o.pusher()
    p = o.get_particles()
    assert p["grid"][0] == grid
    assert np.isclose(p["dy"][0], dy0, atol=1e-12)      # |0.5+dy| 0.2 -> -0.2 -> 0.2
    assert np.isclose(p["uy"][0], -sgn * uy, rtol=1e-15)
    assert p["ux"][0] == ux and p["uz"][0] == uz


def _wall_particles(rng, n, nx, nyp, ny):
    """Particles crowded near both walls with thermal momenta."""
    cx = rng.integers(0, nx, n)
    cy = np.where(rng.random(n) < 0.5, rng.integers(0, 2, n), rng.integers(nyp - 2, nyp, n))
    grid = (cx * ny + cy + 1).astype(np.int32)
    u = 0.3 * rng.standard_normal((3, n))
    return grid, rng.random(n) - 0.5, rng.random(n) - 0.5, u[0], u[1], u[2]


def test_jz_fold_equals_image_deposit(oracle_lib):
    """E = B = 0, particles that do not reach a wall this step: the wall
    system's currents equal the deposit of the particles plus their mirror
    images (y -> -y, charge -q, same v_z) in a periodic domain 2 L_y high --
    Jz, Jy on every domain row, Jx on the rows between the walls."""
    rng = np.random.default_rng(12)
    nx, nyp, d, dt = 10, 6, 0.1, 0.002
    ny = nyp + 2
    grid, dx, dy, ux, uy, uz = _wall_particles(rng, 400, nx, nyp, ny)
    q = np.where(rng.random(400) < 0.5, 1, -1).astype(np.int32)
    mk = lambda nyy, yb: oracle.Oracle(nx, nyy, d, d, dt, CVA, MIME, 0.1, compat=0, ybound=yb)
    o = mk(ny, 1)
    o.set_particles(grid, dx, dy, ux, uy, uz, q)
    o.zero_current(); o.pusher()
    keep = o.get_particles()["uy"] == uy            # not reflected
    assert 50 < keep.sum() < 400
    grid, dx, dy, ux, uy, uz, q = (a[keep] for a in (grid, dx, dy, ux, uy, uz, q))
    o = mk(ny, 1)
    o.set_particles(grid, dx, dy, ux, uy, uz, q)
    o.zero_current(); o.pusher()
    p = o.get_particles()
    g0, g1 = grid - 1, p["grid"] - 1
    cx0, cy0, cx1, cy1 = g0 // ny, g0 % ny, g1 // ny, g1 % ny
    dcx = (cx1 - cx0 + 1) % nx - 1
    dcy = cy1 - cy0
    vz = uz * CVA / np.sqrt(1 + ux * ux + uy * uy + uz * uz)
    # doubled periodic domain: the particles and their images
    ny2 = 2 * nyp
    cyi = ny2 - 1 - cy0
    i1 = np.concatenate([cx0, cx0]) + 1
    i2 = np.concatenate([cx0 + dcx, cx0 + dcx]) + 1
    j1 = np.concatenate([cy0, cyi]) + 1
    j2 = np.concatenate([cy0 + dcy, cyi - dcy]) + 1
    dd = oracle.Oracle(nx, ny2, d, d, dt, CVA, MIME, 0.1, compat=0)
    dd.zero_current()
    dd.deposit_batch(i1, i2, j1, j2, np.concatenate([dx, dx]), np.concatenate([p["dx"], p["dx"]]),
                     np.concatenate([dy, -dy]), np.concatenate([p["dy"], -p["dy"]]),
                     np.concatenate([q, -q]), np.concatenate([vz, vz]))
    dd.finish_current()
    for f, rows in (("jz", slice(0, nyp)), ("jy", slice(0, nyp)), ("jx", slice(1, nyp))):
        a, b = o.get_field(f)[rows], dd.get_field(f)[rows]
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(b)), f
    # guard rows of Jz are cleared after the fold
    assert not o.get_field("jz")[nyp:].any()


def test_gauss_law_with_reflections(oracle_lib):
    """A hot plasma crowded at the walls, 30 steps: many reflections, and the
    Gauss residual between the walls stays unchanged (charge-conserving
    deposit of the reflected move)."""
    case = make_case(nx=12, ny=12, ppc=10, seed=5, efield=0.2, bfield=0.8, vth=(0.1, 0.4),
                     walls=True)
    o = oracle_for(case, compat=0)
    d0 = o.diag()
    o.step(30, sort_every=10)
    d1 = o.diag()
    p = o.get_particles()
    assert np.all((p["grid"] - 1) % 12 < 10)
    assert d1["gauss_dlinf"] < 1e-10 * d0["gauss_linf"]


# --------------------------------------------------------------- Harris
C1_LIKE = dict(geometry=1, ppc=16, half_width=0.5, nb=0.2, b0=1.0, psi0=0.1, vthe_cs=0.14,
               vthi_cs=0.064, vthe_bg=0.14, vthi_bg=0.064, vde_z=0.033, vdi_z=-0.167,
               thermal_z=1, seed=11, stream=0)


@pytest.mark.parametrize("geometry", [0, 1])
def test_harris_loader_statistics(oracle_lib, geometry):
    h = dict(C1_LIKE, geometry=geometry)
    if geometry == 0:
        h.update(thermal_z=0, vde_z=-0.0667, vdi_z=0.0667, psi0=0.0)
    nx, d = 64, 0.1
    ny = 34 if geometry == 1 else 32
    walls = geometry == 1
    nyp = ny - 2 if walls else ny
    ions, ele, fields, ncs = oracle.init_harris(nx, ny, walls, d, d, h)
    n = h["ppc"] * nx * nyp
    assert len(ions["grid"]) == n
    ln = (nx if geometry == 0 else nyp) * d
    nsh = 2 if geometry == 0 else 1
    f = 2 * 0.5 * nsh / (2 * 0.5 * nsh + 0.2 * ln)
    assert ncs == 2 * (int(np.ceil(n * f)) // 2)
    # pairs share their cell and the offset across the sheet
    assert np.array_equal(ions["grid"], ele["grid"])
    across = "dx" if geometry == 0 else "dy"
    assert np.array_equal(ions[across], ele[across])
    for p in (ions, ele):
        assert np.all(np.abs(p["dx"]) <= 0.5) and np.all(np.abs(p["dy"]) <= 0.5)
        assert np.all((p["grid"] - 1) % ny < nyp)
    # sheet positions: sech^2 about the centre; KS distance to the CDF
    g = ions["grid"][:ncs] - 1
    pos = (g // ny + 0.5 + ions["dx"][:ncs]) * d if geometry == 0 else \
          (g % ny + 0.5 + ions["dy"][:ncs]) * d
    centres = [0.25 * nx * d] * (ncs // 2) + [0.75 * nx * d] * (ncs // 2) if geometry == 0 \
        else [0.5 * nyp * d] * ncs
    s = np.sort(pos - np.array(centres))
    cdf = 0.5 * (1 + np.tanh(s / 0.5))
    ks = np.max(np.abs(cdf - (np.arange(1, ncs + 1) - 0.5) / ncs))
    assert ks < 1.8 / np.sqrt(ncs)
    # background uniform across the sheets
    gb = ions["grid"][ncs:] - 1
    pb = (gb // ny + 0.5 + ions["dx"][ncs:]) / nx if geometry == 0 else \
         (gb % ny + 0.5 + ions["dy"][ncs:]) / nyp
    sb = np.sort(pb)
    assert np.max(np.abs(sb - (np.arange(1, len(sb) + 1) - 0.5) / len(sb))) < 1.8 / np.sqrt(len(sb))
    # drifts and thermal spreads (u = gamma v / c)
    uc = lambda v: v / np.sqrt(1 - v * v)
    sheet1 = slice(0, ncs // 2) if geometry == 0 else slice(0, ncs)
    for p, vd, vt in ((ions, h["vdi_z"], h["vthi_cs"]), (ele, h["vde_z"], h["vthe_cs"])):
        m = len(p["uz"][sheet1])
        assert abs(np.mean(p["uz"][sheet1]) - uc(vd)) < 5 * uc(vt) / np.sqrt(m) + 1e-15
        assert abs(np.std(p["ux"][:ncs]) / uc(vt) - 1) < 0.05
        if geometry == 0:
            assert np.allclose(p["uz"][ncs // 2:ncs], -uc(vd))     # second sheet reversed
            assert not p["uz"][ncs:].any()                          # cold background in z
    # fields: E = 0; div B = 0 on the nodes; the tanh profile's curl
    assert not any(fields[k].any() for k in ("ex", "ey", "ez", "bz"))
    bx, by = fields["bx"], fields["by"]
    divb = (bx - np.roll(bx, 1, axis=1)) / d + (by - np.roll(by, 1, axis=0)) / d
    rows = slice(1, nyp) if walls else slice(None)
    assert np.max(np.abs(divb[rows])) < 1e-12
    if geometry == 1:
        # Jz = dBy/dx - dBx/dy at the centres: -b0/L sech^2 plus the psi term
        jz = (np.roll(by, -1, axis=1) - by) / d - (np.roll(bx, -1, axis=0) - bx) / d
        yc = (np.arange(ny) + 0.5) * d - 0.5 * nyp * d
        sech2 = 1 / np.cosh(yc / 0.5) ** 2
        assert np.allclose(jz[nyp // 2 - 2:nyp // 2 + 2].mean(axis=1),
                           -2 * sech2[nyp // 2 - 2:nyp // 2 + 2], rtol=0.05)
        assert np.array_equal(by[nyp], -by[nyp - 1]) and np.array_equal(by[nyp + 1], -by[0])
        assert np.array_equal(bx[nyp + 1], bx[1])
    else:
        x = np.arange(nx) * d
        lx = nx * d
        assert np.allclose(by[0], np.tanh((x - lx / 4) / 0.5) - np.tanh((x - 3 * lx / 4) / 0.5) - 1)


# -------------------------------------------------------------- filters
def _vacuum_periodic(nx, ny, d, dt, **kw):
    return oracle.Oracle(nx, ny, d, d, dt, CVA, MIME, 0.1, compat=0, **kw)


def test_marder_damps_a_gauss_error_mode(oracle_lib):
    """divergenceclean (pic2d.f90:867-885): with B = J = rho = 0 and
    Ex = sin(k x) (so dive = F is a single cosine mode), one solve maps F to
    (1 - dt d kappa) F, d = 0.001, kappa = 4 sin^2(k dx/2)/dx^2 -- the
    discrete Laplacian's eigenvalue; Ey stays zero."""
    nx, ny, d = 32, 8, 0.1
    dt = 0.5 * d / CVA
    for m in (1, 5, 12):
        k = 2 * np.pi * m / (nx * d)
        ex = np.sin(k * (np.arange(nx) + 0.5) * d)[None, :] * np.ones((ny, 1))
        divs = []
        for marder in (False, True):
            o = _vacuum_periodic(nx, ny, d, dt, marder=marder)
            o.set_field("ex", ex)
            o.zero_current()
            o.fieldsolver()
            e = o.get_field("ex")
            divs.append((e - np.roll(e, 1, axis=1)) / d)
            if marder:
                assert np.max(np.abs(o.get_field("ey"))) == 0
        kappa = 4 * np.sin(k * d / 2) ** 2 / d ** 2
        f0 = (ex - np.roll(ex, 1, axis=1)) / d
        # without Marder the divergence is unchanged (curl B = J = 0)
        assert np.max(np.abs(divs[0] - f0)) < 1e-12 * np.max(np.abs(f0))
        assert np.max(np.abs(divs[1] - (1 - dt * 0.001 * kappa) * f0)) < 1e-12 * np.max(np.abs(f0))


@pytest.mark.parametrize("comp", ["ez", "ex", "ey"])
def test_cherenkov_damps_transverse_modes(oracle_lib, comp):
    """suppresscerenkov (pic2d.f90:890-907): a transverse E mode (Ez(x),
    Ex(y) or Ey(x) = cos) in vacuum with B = 0 changes in one solve by the
    factor 1 - c^2 kappa (dt^2/2 + dt tau), tau = dt/8: the Yee part
    (-c^2 dt^2/2 curl curl E) plus the filter's -c^2 dt tau curl curl E."""
    nx = ny = 32
    d = 0.1
    dt = 0.4 * d / CVA
    k = 2 * np.pi * 3 / (nx * d)
    g = np.arange(nx)
    if comp == "ez":        # Ez at x = gx + 1/2
        plane = np.cos(k * (g + 0.5) * d)[None, :] * np.ones((ny, 1))
    elif comp == "ex":      # Ex at y = gy
        plane = np.cos(k * g * d)[:, None] * np.ones((1, nx))
    else:                   # Ey at x = gx
        plane = np.cos(k * g * d)[None, :] * np.ones((ny, 1))
    kappa = 4 * np.sin(k * d / 2) ** 2 / d ** 2
    for cher, fac in ((False, 1 - CVA ** 2 * kappa * dt * dt / 2),
                      (True, 1 - CVA ** 2 * kappa * (dt * dt / 2 + dt * dt / 8))):
        o = _vacuum_periodic(nx, ny, d, dt, cherenkov=cher)
        o.set_field(comp, plane)
        o.zero_current()
        o.fieldsolver()
        assert np.max(np.abs(o.get_field(comp) - fac * plane)) < 1e-12, (comp, cher)


@pytest.mark.parametrize("shift", [(0, 5), (7, 0), (3, 11)])
def test_filters_translation_invariant(oracle_lib, shift):
    """A full step with both filters commutes with a periodic shift of fields
    and particles: the ghost cells hold periodic copies wherever the filter
    stencils read them."""
    sx, sy = shift
    case = make_case(nx=16, ny=14, ppc=6, seed=3, efield=0.3, bfield=0.8)
    shifted = dict(case)
    shifted["fields"] = {k: np.roll(np.roll(v, sy, axis=0), sx, axis=1)
                         for k, v in case["fields"].items()}
    sp = []
    for p in case["species"]:
        g = p["grid"] - 1
        cx, cy = g // case["ny"], g % case["ny"]
        q = dict(p)
        q["grid"] = (((cx + sx) % case["nx"]) * case["ny"] + (cy + sy) % case["ny"] + 1).astype(np.int32)
        sp.append(q)
    shifted["species"] = sp
    out = []
    for c in (case, shifted):
        o = oracle_for(c, compat=0)
        o.set_filters(True, True)
        o.step(2, sort_every=0)
        out.append({k: o.get_field(k) for k in ("ex", "ey", "ez", "bx", "by", "bz")})
    for k in out[0]:
        ref = np.roll(np.roll(out[0][k], sy, axis=0), sx, axis=1)
        assert np.max(np.abs(out[1][k] - ref)) < 1e-12 * np.max(np.abs(ref)), k
