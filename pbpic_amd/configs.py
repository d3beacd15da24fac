"""Benchmark configurations mapped onto the reference's normalization.

pic2d.f90 measures lengths in d_i, time in 1/Omega_ci and velocity in v_A
(c = cva), and normalizes the current by in0 = 1/(n0 dx dy) (:15-43).  With
in0 = 1/ppc_e the electron plasma frequency is omega_pe = cva*sqrt(mime)
(linearizing E' = cva^2 (curl B - J) against the Boris push of :964-1000), so
a grid spacing of dx_de electron inertial lengths is dx = dx_de/sqrt(mime) d_i
and a time step of dt_wpe/omega_pe is dt = dt_wpe/(cva*sqrt(mime)).
"""
import math

CVA = 15.0    # the reference's c/v_A (pic2d.f90:20)
MIME = 25.0   # mass ratio of every configuration (pic2d.f90:19)


def electron_units(nx, ny, dx_de, dt_wpe, ppc_e, cva=CVA, mime=MIME):
    s = math.sqrt(mime)
    return dict(nx=nx, ny=ny, dx=dx_de / s, dy=dx_de / s, dt=dt_wpe / (cva * s), cva=cva,
                mime=mime, in0=1.0 / ppc_e)


def thermal_c0(nx=128, ny=128, ppc=16):
    """Plain FP64 PIC, the CPU-runnable case: 128x128, dx=dy=0.5 d_e,
    dt=0.2/wpe, electrons (vth=0.1c) and ions (vth=0.02c), ppc per species."""
    p = electron_units(nx, ny, 0.5, 0.2, ppc)
    p.update(species=[dict(q=1, ppc=ppc, vth=0.02, vdrift=0.0, counter=False),
                      dict(q=-1, ppc=ppc, vth=0.1, vdrift=0.0, counter=False)],
             rho_background=0.0, steps=100)
    return p


def two_stream_c2(nx=2048, ny=2048, ppc=100):
    """Throughput at scale: 2048x2048 periodic, dx=0.5 d_e, two electron beams
    at +-0.2c along x with vth=0.02c, ppc electrons per cell (half per beam),
    immobile neutralizing ions, zero initial fields, 500 steps."""
    p = electron_units(nx, ny, 0.5, 0.2, ppc)
    p.update(species=[dict(q=-1, ppc=ppc, vth=0.02, vdrift=0.2, counter=True)],
             rho_background=float(ppc), steps=500)
    return p


def harris_reference(nx=256, ny=256, ppg=64):
    """The reference's own double Harris sheet (pic2d.f90:12-43, initparticles
    :357-578, initfields :583-648): periodic, sheets across x at Lx/4 and
    3Lx/4, half width dcs = 0.5, nbgcs = 0.2, vthec = 0.092, tiecs = 5,
    tiebg = 1, tbgcs = 0.1, vdrift = 1/cva, cold drift in z."""
    vthe = 0.092
    vthe_bg = vthe * math.sqrt(0.1)
    harris = dict(geometry="double_x", ppc=ppg, half_width=0.5, nb=0.2, b0=1.0, psi0=0.0,
                  vthe_cs=vthe, vthi_cs=vthe * math.sqrt(5.0 / MIME), vthe_bg=vthe_bg,
                  vthi_bg=vthe_bg * math.sqrt(1.0 / MIME), vde_z=-1.0 / CVA, vdi_z=1.0 / CVA,
                  thermal_z=0)
    return dict(nx=nx, ny=ny, dx=0.04, dy=0.04, dt=0.045 / MIME, cva=CVA, mime=MIME,
                in0=1.0 / 291.0, ybound="periodic", harris=harris, steps=30)


def harris_c4(nx=1024, ny=1024, ppc=64):
    """The large differentiable configuration: the Harris sheet of harris_c1
    (same physics and cell size, 0.025 d_i) on nx x ny = 1024 x 1024 domain
    cells, i.e. 25.6 x 25.6 d_i; the adjoint's loss is J.E over the central
    4 x 2 d_i diffusion region (loss_region), 1000 steps, checkpoints every
    50."""
    p = harris_c1(nx, ny, ppc, lx=0.025 * nx, ly=0.025 * ny)
    p.update(steps=1000, ckpt_every=50)
    return p


def loss_region(cfg, wx=4.0, wy=2.0):
    """Weights (ny, nx) of the loss: 1 for the cells whose centre lies in the
    central wx x wy box (d_i), 0 elsewhere (guard rows included)."""
    import numpy as np
    nx, ny = cfg["nx"], cfg["ny"]
    nyp = ny - 2 if cfg.get("ybound") == "conducting" else ny
    x = (np.arange(nx) + 0.5) * cfg["dx"] - 0.5 * nx * cfg["dx"]
    y = (np.arange(ny) + 0.5) * cfg["dy"] - 0.5 * nyp * cfg["dy"]
    m = (np.abs(x)[None, :] < 0.5 * wx) & (np.abs(y)[:, None] < 0.5 * wy)
    m[nyp:] = False
    return m.astype(np.float64)


def harris_c1(nx=1024, ny=512, ppc=64, lx=25.6, ly=12.8):
    """Harris current sheet with conducting y walls: nx x ny domain cells over
    25.6 x 12.8 d_i (the lattice gets ny + 2 rows: two guard rows), half
    thickness 0.5 d_i, mass ratio 25, Ti/Te = 5, background 0.2 n0,
    w_pe/W_ce = 2 (so c/v_A = 2 sqrt(mime) = 10), ppc particles per cell and
    species, 10% flux perturbation psi = 0.1 cos(2 pi x/Lx) cos(pi y/Ly),
    2000 steps.  Pressure balance n0 (Ti + Te) = B0^2/2 gives Te = 1/12 and
    Ti = 5/12 m_i v_A^2; the Harris drifts Vi/Ti = -Ve/Te with Vi - Ve = 1/L
    give Vi = -5/3, Ve = +1/3 v_A (negative ion drift: Bx = tanh(y/L) carries
    Jz < 0).  in0 normalizes the peak sheet density: 1/(sheet particles per
    cell at n0)."""
    mime, ti_te, L, nb = MIME, 5.0, 0.5, 0.2
    cva = 2.0 * math.sqrt(mime)
    dx, dy = lx / nx, ly / ny
    te = 0.5 / (1.0 + ti_te)
    ti = ti_te * te
    vthe = math.sqrt(te * mime) / cva
    vthi = math.sqrt(ti) / cva
    vi, ve = -(1.0 / L) * ti / (ti + te), (1.0 / L) * te / (ti + te)
    n = ppc * nx * ny
    ncs = 2 * (math.ceil(n * 2 * L / (2 * L + nb * ly)) // 2)
    peak = ncs * dx * dy / (2 * L * lx)
    dt = 0.6 * min(dx, dy) / cva
    harris = dict(geometry="single_y", ppc=ppc, half_width=L, nb=nb, b0=1.0, psi0=0.1,
                  vthe_cs=vthe, vthi_cs=vthi, vthe_bg=vthe, vthi_bg=vthi, vde_z=ve / cva,
                  vdi_z=vi / cva, thermal_z=1)
    return dict(nx=nx, ny=ny + 2, dx=dx, dy=dy, dt=dt, cva=cva, mime=mime, in0=1.0 / peak,
                ybound="conducting", harris=harris, steps=2000)
