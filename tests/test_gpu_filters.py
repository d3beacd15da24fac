"""GPU parity of the Marder divergence-cleaning and Cherenkov-suppression
filters (pic2d.f90:867-907, commented out at :754-755) against the oracle's
restatement: one full step to 1e-12 (FP64) on the periodic lattice and with
conducting walls, 20 steps of diagnostics to 1e-9, and the split form to
FP32 bars."""
import numpy as np
import pytest

from cases import gpu_for, make_case, oracle_for, rel_norm

pytestmark = pytest.mark.gpu

FIELDS = ("ex", "ey", "ez", "bx", "by", "bz")


def _pair(case, precision, compat, marder, cherenkov, sort_every=0):
    ctx = gpu_for(case, precision=precision, compat=compat, sort_every=sort_every,
                  marder=marder, cherenkov=cherenkov)
    o = oracle_for(case, compat=compat)
    o.set_filters(marder, cherenkov)
    return ctx, o


@pytest.mark.parametrize("walls", [False, True])
@pytest.mark.parametrize("filters", [(True, False), (False, True), (True, True)])
@pytest.mark.parametrize("compat", [7, 0])
def test_one_step_filters_fp64(gpu, walls, filters, compat):
    print(f"filters={filters} walls={walls} compat={compat}", flush=True)
    case = make_case(nx=24, ny=20, ppc=8, seed=17, efield=0.4, bfield=1.0, walls=walls)
    ctx, o = _pair(case, 0, compat, *filters)
    ctx.countingsort(); o.countingsort()
    for _ in range(2):
        ctx.zero_current(); ctx.pusher(); ctx.fieldsolver()
        o.zero_current(); o.pusher(); o.fieldsolver()
    for f in FIELDS:
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < 1e-12, f


def test_filters_multi_step_fp64(gpu):
    print("filters 20 steps", flush=True)
    case = make_case(nx=24, ny=20, ppc=10, seed=3)
    ctx, o = _pair(case, 0, 7, True, True, sort_every=10)
    ctx.countingsort(); o.countingsort()
    ctx.step(20)
    o.step(20, sort_every=10)
    d, od = ctx.diag(), o.diag()
    for key in ("ke_e", "ke_i", "w_e", "w_b", "je_int"):
        assert abs(d[key] - od[key]) <= 1e-9 * abs(od[key]), (key, d[key], od[key])


def test_filters_split(gpu):
    print("filters split", flush=True)
    case = make_case(nx=24, ny=20, ppc=8, seed=5)
    ctx, o = _pair(case, 1, 7, True, True)
    ctx.countingsort(); o.countingsort()
    ctx.zero_current(); ctx.pusher(); ctx.fieldsolver()
    o.zero_current(); o.pusher(); o.fieldsolver()
    for f in FIELDS:
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < 1e-5, f


def test_marder_needs_rho(gpu):
    import pbpic_amd
    with pytest.raises(pbpic_amd.PbpicError):
        pbpic_amd.Context(8, 8, 0.1, 0.1, 0.001, 15.0, 25.0, 0.1, deposit_rho=False, marder=True)
