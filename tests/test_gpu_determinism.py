"""Run-to-run reproducibility.  The reference's serial deposit loop
(pic2d.f90:934-1063) gives one answer per input.  Every current contribution
is turned into 64-bit fixed point (scale: a power of two per plane): in the
split form rounded once to one integer, in the FP64 form split into two, its
rounded value and the rounded remainder times 2^31 (a quantum 2^-62 of the
plane's bound on one contribution).  All sums, in LDS and in the global
planes, are integer sums, so their order does not matter; the current planes
are one rounding of the exact sums.  The diagnostics are reduced in a fixed
order (per-workgroup partials, then one ordered sum), and the Gauss
residual's charge deposit is fixed point too.  So two runs from the same
state give the same bits, in both forms: particles, J, E, B and every
diagnostic, sorts included, with periodic and with conducting walls."""
import numpy as np
import pytest

from cases import gpu_for, make_case

pytestmark = pytest.mark.gpu

FIELDS = ("jx", "jy", "jz", "rho", "ex", "ey", "ez", "bx", "by", "bz")


def _run(case, precision, steps, sort_every):
    ctx = gpu_for(case, precision=precision, compat=0, sort_every=sort_every)
    for k in range(2):
        ctx.countingsort(k)
    d0 = ctx.diag()
    ctx.step(steps)
    out = dict(diag0=d0, diag=ctx.diag(), fields={f: ctx.get_field(f) for f in FIELDS},
               parts=[ctx.get_particles(k, native=True) for k in range(2)])
    ctx.close()
    return out


@pytest.mark.parametrize("precision,walls", [(1, False), (1, True), (0, False), (0, True)])
def test_runs_are_bit_identical(gpu, precision, walls):
    print(f"determinism precision={precision} walls={walls}", flush=True)
    case = make_case(nx=96, ny=82 if walls else 80, ppc=24, seed=33, vth=(0.05, 0.3),
                     efield=0.5, bfield=1.0, walls=walls)
    a = _run(case, precision, 60, 25)
    b = _run(case, precision, 60, 25)
    for k in a["diag"]:
        assert a["diag"][k] == b["diag"][k], k
        assert a["diag0"][k] == b["diag0"][k], k
    for f in FIELDS:
        assert np.array_equal(a["fields"][f], b["fields"][f]), f
    for pa, pb in zip(a["parts"], b["parts"]):
        for f in ("grid", "dx", "dy", "ux", "uy", "uz", "pid"):
            assert np.array_equal(pa[f], pb[f]), f
    assert a["diag"]["ke_e"] > 0 and np.abs(a["fields"]["jx"]).max() > 0


@pytest.mark.parametrize("precision", [1, 0])
def test_deposit_is_order_independent(gpu, precision):
    """The batched deposit of the same contributions in two different orders
    gives the same current bits (integer sums), in both forms."""
    import pbpic_amd
    rng = np.random.default_rng(4)
    nx, ny, n = 32, 24, 20000
    i1, j1 = rng.integers(1, nx + 1, n), rng.integers(1, ny + 1, n)
    a, b = rng.integers(-1, 2, n), rng.integers(-1, 2, n)
    dx1, dy1 = rng.random(n) - 0.5, rng.random(n) - 0.5
    dx2 = np.where(a == 0, np.clip(dx1 + 0.2 * (rng.random(n) - 0.5), -0.5, 0.5),
                   -a * 0.5 + a * 0.4 * rng.random(n))
    dy2 = np.where(b == 0, np.clip(dy1 + 0.2 * (rng.random(n) - 0.5), -0.5, 0.5),
                   -b * 0.5 + b * 0.4 * rng.random(n))
    q, vz = rng.choice([-1, 1], n), rng.standard_normal(n)
    args = [i1, i1 + a, j1, j1 + b, dx1, dx2, dy1, dy2, q, vz]
    out = []
    for order in (np.arange(n), rng.permutation(n)):
        ctx = pbpic_amd.Context(nx, ny, 0.1, 0.12, 0.003, 15.0, 25.0, 0.05, precision=precision)
        ctx.currentdeposit(*[x[order] for x in args])
        out.append({f: ctx.get_field(f) for f in ("jx", "jy", "jz", "rho")})
        ctx.close()
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f]), f
