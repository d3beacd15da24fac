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
