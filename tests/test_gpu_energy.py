"""Multi-step agreement, judged the way diverging trajectories allow: 500 steps
of the CPU-runnable thermal plasma (128 x 128, dx = 0.5 d_e, dt = 0.2/w_pe,
16 particles per cell and species, pic2d.f90:171-188 with a sort every 50
steps) from the same seeded state, on the GPU in both forms and in the FP64
oracle, with the reference's three quirks switched off (compat 0) so that the
deposit conserves charge exactly.  Bars, at every 50th step:
  total energy (kinetic + field, totene):  split within 1% of the oracle's
      total at step 50 from the oracle, FP64 within 0.1%;
  integrated J.E:                           split within 2% of the oracle's
      magnitude, FP64 within 0.5%;
  Gauss-residual drift (charge conservation of the deposit): FP64 below 1e-9
      and split below 1e-4 of the residual itself.
"""
import os

import numpy as np
import pytest

import oracle
from pbpic_amd import configs

pytestmark = pytest.mark.gpu

NSTEPS, EVERY = 500, 50


def _particles(cfg):
    out = []
    for k, sp in enumerate(cfg["species"]):
        p = oracle.init_maxwellian(cfg["nx"], cfg["ny"], sp["ppc"], sp["vth"], sp["vdrift"],
                                   counter_stream=False, seed=20260101, stream=k)
        out.append((sp["q"], p))
    return out


def _oracle_history(cfg, parts):
    threads = min(16, os.cpu_count() or 1)
    o = oracle.Oracle(cfg["nx"], cfg["ny"], cfg["dx"], cfg["dy"], cfg["dt"], cfg["cva"],
                      cfg["mime"], cfg["in0"], prec=8, compat=0, threads=threads)
    cat = lambda k: np.concatenate([p[k] for _, p in parts])
    q = np.concatenate([np.full(len(p["grid"]), qq, np.int32) for qq, p in parts])
    o.set_particles(cat("grid"), cat("dx"), cat("dy"), cat("ux"), cat("uy"), cat("uz"), q)
    o.countingsort()
    hist = [o.diag()]
    for _ in range(NSTEPS // EVERY):
        o.step(EVERY, sort_every=50)
        hist.append(o.diag())
    o.close()
    return hist


def _gpu_history(cfg, parts, precision):
    import pbpic_amd
    ctx = pbpic_amd.Context(cfg["nx"], cfg["ny"], cfg["dx"], cfg["dy"], cfg["dt"], cfg["cva"],
                            cfg["mime"], cfg["in0"], precision=precision, compat=0,
                            sort_every=50)
    for qq, p in parts:
        k = ctx.add_species(qq, len(p["grid"]))
        ctx.set_particles(k, p["grid"], p["dx"], p["dy"], p["ux"], p["uy"], p["uz"])
    ctx.countingsort()
    hist = [ctx.diag()]
    for _ in range(NSTEPS // EVERY):
        ctx.step(EVERY)
        hist.append(ctx.diag())
    ctx.close()
    return hist


def _total(d):
    return d["ke_e"] + d["ke_i"] + d["w_e"] + d["w_b"]


def test_500_steps_energy_je_gauss_against_oracle(gpu):
    cfg = configs.thermal_c0(128, 128, 16)
    parts = _particles(cfg)
    print("oracle FP64 500 steps", flush=True)
    ref = _oracle_history(cfg, parts)
    # the kinetic-energy tally is that of the last push: compare from step 50
    e0 = _total(ref[1])
    je_scale = max(abs(d["je_int"]) for d in ref)
    for precision, tol_e, tol_je, tol_g in ((0, 1e-3, 5e-3, 1e-9), (1, 1e-2, 2e-2, 1e-4)):
        print(f"gpu precision={precision} 500 steps", flush=True)
        got = _gpu_history(cfg, parts, precision)
        for n, (g, r) in enumerate(zip(got, ref)):
            if n > 0:
                assert abs(_total(g) - _total(r)) <= tol_e * e0, (precision, n * EVERY,
                                                                 _total(g), _total(r))
            assert abs(g["je_int"] - r["je_int"]) <= tol_je * je_scale, (precision, n * EVERY,
                                                                        g["je_int"], r["je_int"])
            assert g["gauss_dlinf"] <= tol_g * g["gauss_linf"], (precision, n * EVERY,
                                                                g["gauss_dlinf"], g["gauss_linf"])
        drift = (_total(got[-1]) - _total(got[1])) / _total(got[1])
        drift_ref = (_total(ref[-1]) - e0) / e0
        print(f"precision={precision}: energy drift {drift:.3e} (oracle {drift_ref:.3e})", flush=True)
