"""pbpic_deposit_noalloc: the deposit entry point that allocates no device
memory (its inputs go through a fixed per-GPU scratch array) against
pbpic_deposit, through the C ABI, on two calls in a row with the second
smaller, in both precision forms.  Bar: both entry points add every
contribution as a 64-bit fixed-point integer (one word in the split form,
two in the FP64 form), exact in any order of the atomics, so J and rho are
bit-identical."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(rng, nx, ny, n):
    i1 = rng.integers(1, nx + 1, n)
    j1 = rng.integers(1, ny + 1, n)
    a = rng.integers(-1, 2, n)
    b = rng.integers(-1, 2, n)
    dx1 = rng.random(n) - 0.5
    dy1 = rng.random(n) - 0.5
    dx2 = np.where(a == 0, np.clip(dx1 + 0.3 * (rng.random(n) - 0.5), -0.5, 0.5),
                   -a * 0.5 + a * 0.4 * rng.random(n))
    dy2 = np.where(b == 0, np.clip(dy1 + 0.3 * (rng.random(n) - 0.5), -0.5, 0.5),
                   -b * 0.5 + b * 0.4 * rng.random(n))
    q = rng.choice([-1, 1], n)
    vz = rng.standard_normal(n)
    return (i1, i1 + a, j1, j1 + b, dx1, dx2, dy1, dy2, q, vz)


def _rel(a, b):
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


@pytest.mark.parametrize("precision", [0, 1])
def test_deposit_noalloc_matches_deposit(gpu, precision):
    import pbpic_amd
    print(f"deposit_noalloc precision={precision}", flush=True)
    nx, ny = 64, 48
    rng = np.random.default_rng(11 + precision)
    ref = pbpic_amd.Context(nx, ny, 0.1, 0.12, 0.003, 15.0, 25.0, 0.05, precision=precision)
    new = pbpic_amd.Context(nx, ny, 0.1, 0.12, 0.003, 15.0, 25.0, 0.05, precision=precision)
    for n in (pbpic_amd.DEPOSIT_NOALLOC_MAX, 1237):
        args = _batch(rng, nx, ny, n)
        ref.zero_current()
        new.zero_current()
        ref.currentdeposit(*args)
        new.currentdeposit(*args, noalloc=True)
        for f in ("jx", "jy", "jz", "rho"):
            a, b = new.get_field(f), ref.get_field(f)
            assert np.array_equal(a, b), (n, f, _rel(a, b))
            assert np.abs(b).max() > 0, (n, f)
    ref.close()
    new.close()


def test_deposit_noalloc_rejects_oversize(gpu):
    import pbpic_amd
    nx, ny = 16, 16
    ctx = pbpic_amd.Context(nx, ny, 0.1, 0.12, 0.003, 15.0, 25.0, 0.05)
    args = _batch(np.random.default_rng(3), nx, ny, pbpic_amd.DEPOSIT_NOALLOC_MAX + 1)
    with pytest.raises(pbpic_amd.PbpicError) as e:
        ctx.currentdeposit(*args, noalloc=True)
    assert e.value.code == -3 and "exceeds" in str(e.value)
    ctx.close()
