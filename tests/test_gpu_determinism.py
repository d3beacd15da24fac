"""Run-to-run reproducibility.  The reference's serial deposit loop
(pic2d.f90:934-1063) gives one answer per input.  In the split form every
current contribution is rounded once to 64-bit fixed point (scale: a power of
two per plane) and all sums, in LDS and in the global planes, are integer
sums, so their order does not matter; the FP32 planes are one rounding of the
exact sums.  The diagnostics are reduced in a fixed order (per-workgroup
partials, then one ordered sum), and the Gauss residual's charge deposit is
fixed point too.  So two runs of the split form from the same state give the
same bits: particles, J, E, B and every diagnostic, sorts included, with
periodic and with conducting walls.  The FP64 form keeps FP64 atomic sums
(exact 1e-12 parity with the oracle's FP64 arithmetic needs them): its runs
agree to rounding, checked here to 1e-12."""
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


@pytest.mark.parametrize("walls", [False, True])
def test_split_runs_are_bit_identical(gpu, walls):
    print(f"determinism split walls={walls}", flush=True)
    case = make_case(nx=96, ny=82 if walls else 80, ppc=24, seed=33, vth=(0.05, 0.3),
                     efield=0.5, bfield=1.0, walls=walls)
    a = _run(case, 1, 60, 25)
    b = _run(case, 1, 60, 25)
    for k in a["diag"]:
        assert a["diag"][k] == b["diag"][k], k
        assert a["diag0"][k] == b["diag0"][k], k
    for f in FIELDS:
        assert np.array_equal(a["fields"][f], b["fields"][f]), f
    for pa, pb in zip(a["parts"], b["parts"]):
        for f in ("grid", "dx", "dy", "ux", "uy", "uz", "pid"):
            assert np.array_equal(pa[f], pb[f]), f
    assert a["diag"]["ke_e"] > 0 and np.abs(a["fields"]["jx"]).max() > 0


def test_split_deposit_is_order_independent(gpu):
    """The batched deposit of the same contributions in two different orders
    gives the same current bits (integer sums)."""
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
        ctx = pbpic_amd.Context(nx, ny, 0.1, 0.12, 0.003, 15.0, 25.0, 0.05, precision=1)
        ctx.currentdeposit(*[x[order] for x in args])
        out.append({f: ctx.get_field(f) for f in ("jx", "jy", "jz", "rho")})
        ctx.close()
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f]), f


def test_fp64_runs_agree_to_rounding(gpu):
    print("determinism fp64", flush=True)
    case = make_case(nx=96, ny=80, ppc=24, seed=33, vth=(0.05, 0.3), efield=0.5, bfield=1.0)
    a = _run(case, 0, 60, 25)
    b = _run(case, 0, 60, 25)
    for k in ("ke_e", "ke_i", "w_e", "w_b", "je_int"):
        assert abs(a["diag"][k] - b["diag"][k]) <= 1e-12 * abs(a["diag"][k]), k
    for f in FIELDS:
        d = np.linalg.norm(a["fields"][f] - b["fields"][f])
        assert d <= 1e-12 * np.linalg.norm(a["fields"][f]), f
