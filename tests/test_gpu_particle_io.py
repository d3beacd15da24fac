"""Host particle I/O through the C ABI (pbpic_set_particles[_f32],
pbpic_get_particles[_f32]): exact round trips in both forms, arrays larger
than one 64 MiB staging chunk, and invalid grid values rejected before any
state changes (the reference would index out of its arrays)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _ctx(precision, nx=64, ny=48, walls=False, track_pid=True):
    import pbpic_amd
    return pbpic_amd.Context(nx, ny, 0.1, 0.1, 0.001, 15.0, 25.0, 0.01, precision=precision,
                             track_pid=track_pid, ybound="conducting" if walls else "periodic")


def _arrays(n, ncell, seed, dtype):
    rng = np.random.default_rng(seed)
    a = dict(grid=rng.integers(1, ncell + 1, n).astype(np.int32),
             dx=rng.uniform(-0.5, 0.5, n), dy=rng.uniform(-0.5, 0.5, n),
             ux=rng.standard_normal(n) * 0.1, uy=rng.standard_normal(n) * 0.1,
             uz=rng.standard_normal(n) * 0.1, pid=np.arange(1, n + 1, dtype=np.int32))
    for k in ("dx", "dy", "ux", "uy", "uz"):
        a[k] = a[k].astype(dtype)
    return a


@pytest.mark.parametrize("precision,host", [(0, np.float64), (1, np.float32), (1, np.float64)])
@pytest.mark.parametrize("n", [1000, 17_000_000])
def test_round_trip(gpu, precision, host, n):
    """Particles in and out unchanged: FP64 exactly, the split form exactly
    for float32 input and rounded to float32 for float64 input.  17M
    particles span two staging chunks per array."""
    _say(f"particle I/O precision={precision} host={np.dtype(host).name} n={n}")
    ctx = _ctx(precision)
    k = ctx.add_species(-1, n)
    a = _arrays(n, 64 * 48, 7, host)
    ctx.set_particles(k, a["grid"], a["dx"], a["dy"], a["ux"], a["uy"], a["uz"], a["pid"])
    assert ctx.num_particles(k) == n
    g = ctx.get_particles(k, native=True)
    assert np.array_equal(g["grid"], a["grid"]) and np.array_equal(g["pid"], a["pid"])
    for f in ("dx", "dy", "ux", "uy", "uz"):
        want = a[f] if precision == 0 else a[f].astype(np.float32)
        assert np.array_equal(g[f], want), f
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_invalid_grid_leaves_the_species_unchanged(gpu, precision):
    """grid 0, grid beyond nx*ny and (with conducting walls) a grid in a
    guard row are rejected with PBPIC_ERR_CELL; the species keeps the
    particles it had."""
    import pbpic_amd
    _say(f"invalid grid precision={precision}")
    for walls in (False, True):
        ctx = _ctx(precision, walls=walls)
        ny = 48
        k = ctx.add_species(1, 5000)
        good = _arrays(5000, 64 * ny, 3, np.float64)
        if walls:
            good["grid"] = ((good["grid"] - 1) // ny * ny + (good["grid"] - 1) % (ny - 2) + 1).astype(np.int32)
        ctx.set_particles(k, good["grid"], good["dx"], good["dy"], good["ux"], good["uy"],
                          good["uz"], good["pid"])
        bad_values = [0, 64 * ny + 1] + ([ny - 1] if walls else [])   # ny - 1 = guard row nyp
        for v in bad_values:
            b = {key: val.copy() for key, val in good.items()}
            b["grid"][4321] = v
            with pytest.raises(pbpic_amd.PbpicError) as e:
                ctx.set_particles(k, b["grid"], b["dx"], b["dy"], b["ux"], b["uy"], b["uz"], b["pid"])
            assert "grid[4321]" in str(e.value), (v, str(e.value))
            g = ctx.get_particles(k)
            assert ctx.num_particles(k) == 5000
            assert np.array_equal(g["grid"], good["grid"]), (walls, v)
        ctx.close()
