"""GPU parity: the HIP path through the C ABI against the CPU restatement of
pic2d.f90 (the oracle) on the same seeded inputs.

Bars: bit-exact for the sort (counts and permutation); FP64 single step to
1e-12 relative (particles element-wise, currents and fields norm-wise, since
the current sums are accumulated in a different order); the split-precision
form against the FP64 oracle within FP32 tolerances stated per test.
"""
import numpy as np
import pytest

import oracle
from cases import (SPECIES_Q, gpu_for, make_case, oracle_for, oracle_species, rel_norm,
                   supertile_key)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _say(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("kind", ["random", "nearly_sorted", "one_cell", "reverse", "ragged"])
@pytest.mark.parametrize("nx,ny", [(16, 12), (300, 257)])
def test_countingsort_bitexact(gpu, kind, nx, ny):
    """countingsort (pic2d.f90:1303-1329): counts == bincount, permutation ==
    stable argsort, and both == the oracle's restatement."""
    _say(f"sort kind={kind} nx={nx} ny={ny}")
    rng = np.random.default_rng(11)
    ncell = nx * ny
    n = {"random": 20000, "nearly_sorted": 30000, "one_cell": 5000, "reverse": 9000,
         "ragged": 4097}[kind]
    if kind == "random":
        grid = rng.integers(1, ncell + 1, n)
    elif kind == "nearly_sorted":
        grid = np.sort(rng.integers(1, ncell + 1, n))
        grid = np.clip(grid + rng.integers(-3, 4, n), 1, ncell)
    elif kind == "one_cell":
        grid = np.full(n, ncell // 2 + 1)
    elif kind == "reverse":
        grid = np.sort(rng.integers(1, ncell + 1, n))[::-1].copy()
    else:
        grid = rng.integers(1, min(ncell, 40) + 1, n)
    case = make_case(nx=nx, ny=ny, ppc=1, seed=3)
    case["species"] = [dict(grid=grid.astype(np.int32), dx=rng.random(n) - 0.5,
                            dy=rng.random(n) - 0.5, ux=rng.standard_normal(n),
                            uy=rng.standard_normal(n), uz=rng.standard_normal(n),
                            pid=np.arange(1, n + 1, dtype=np.int32))]
    import pbpic_amd
    ctx = pbpic_amd.Context(nx, ny, 0.1, 0.1, 0.01, 15.0, 25.0, 0.1, track_pid=True,
                            push_order=False)
    k = ctx.add_species(-1, n)
    s = case["species"][0]
    ctx.set_particles(k, s["grid"], s["dx"], s["dy"], s["ux"], s["uy"], s["uz"], s["pid"])
    counts, perm = ctx.countingsort(k, want=True)
    assert np.array_equal(counts, np.bincount(grid - 1, minlength=ncell))
    assert np.array_equal(perm, np.argsort(grid, kind="stable"))
    o = oracle.Oracle(nx, ny, 0.1, 0.1, 0.01, 15.0, 25.0, 0.1)
    o.set_particles(s["grid"], s["dx"], s["dy"], s["ux"], s["uy"], s["uz"],
                    np.full(n, -1, np.int32), s["pid"])
    oc, op = o.countingsort(want=True)
    assert np.array_equal(counts, oc) and np.array_equal(perm, op)
    g = ctx.get_particles(k)
    assert np.array_equal(g["grid"], grid[perm])
    assert np.array_equal(g["pid"], s["pid"][perm])
    assert np.array_equal(g["ux"], s["ux"][perm])


def _push_order(ref_perm, V):
    """The push order of a cell-order permutation: inside each full tile of
    4096 slots, slot q = c*64V + l*V + j (round c, lane l, element j) holds
    cell-order rank l*64 + c*V + j; a ragged last tile keeps cell order."""
    out = ref_perm.copy()
    q = np.arange(4096)
    rank = ((q % (64 * V)) // V) * 64 + (q // (64 * V)) * V + q % V
    assert np.array_equal(np.sort(rank), q)                 # a bijection
    for t0 in range(0, len(ref_perm) - 4095, 4096):
        out[t0:t0 + 4096] = ref_perm[t0 + rank]
    return out


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [4095, 4096, 20000, 3 * 4096 + 17])
@pytest.mark.parametrize("nx,ny", [(40, 37), (40, 130), (200, 20)])
def test_countingsort_push_order(gpu, precision, n, nx, ny):
    """With the push order (the default) the particles are sorted stably by
    supertile key (8 x 8 cell blocks), every full 4096-particle tile is laid
    out so that the 64 lanes of a wave-wide load hold particles 64 ranks
    apart, and the per-cell counts are the reference's."""
    _say(f"push-order sort precision={precision} n={n} nx={nx} ny={ny}")
    rng = np.random.default_rng(5 + n)
    grid = rng.integers(1, nx * ny + 1, n).astype(np.int32)
    ux = rng.standard_normal(n)
    pid = np.arange(1, n + 1, dtype=np.int32)
    import pbpic_amd
    ctx = pbpic_amd.Context(nx, ny, 0.1, 0.1, 0.01, 15.0, 25.0, 0.1, precision=precision,
                            track_pid=True)
    k = ctx.add_species(-1, n)
    z = np.zeros(n)
    ctx.set_particles(k, grid, z, z, ux, z, z, pid)
    counts, perm = ctx.countingsort(k, want=True)
    ref = np.argsort(supertile_key(grid - 1, nx, ny), kind="stable")
    assert np.array_equal(counts, np.bincount(grid - 1, minlength=nx * ny))
    for t0 in range(0, n, 4096):
        assert np.array_equal(np.sort(perm[t0:t0 + 4096]), np.sort(ref[t0:t0 + 4096]))
    assert np.array_equal(perm, _push_order(ref, 2 if precision == 0 else 4))
    g = ctx.get_particles(k)
    assert np.array_equal(g["grid"], grid[perm]) and np.array_equal(g["pid"], pid[perm])
    want_ux = ux[perm] if precision == 0 else ux.astype(np.float32)[perm]
    assert np.array_equal(g["ux"], want_ux)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("push_order", [False, True])
@pytest.mark.parametrize("nx,ny,ppc,boxed", [(128, 96, 30, True), (300, 257, 7, False),
                                             (40, 40, 200, True)])
def test_countingsort_after_push_bitexact(gpu, precision, push_order, nx, ny, ppc, boxed):
    """A sort right after a push in push order takes the tile-box path
    (records per tile box cell, radix-sorted by key, scanned, then one stable
    scatter).  Its
    permutation is the stable argsort of the keys (cell ids in reference
    order, supertile keys in push order, laid out in the push kernel's lane
    order inside full tiles), bit for bit, and the particles move with it."""
    _say(f"sort after push precision={precision} push_order={push_order} {nx}x{ny}x{ppc}")
    case = make_case(nx=nx, ny=ny, ppc=ppc, seed=nx + ppc, vth=(0.1, 0.45), efield=0.2,
                     bfield=1.0)
    ctx = gpu_for(case, precision=precision, push_order=push_order)
    for k in range(2):
        ctx.countingsort(k)
    for _ in range(3):
        ctx.zero_current()
        ctx.pusher()
        ctx.fieldsolver()
    V = 2 if precision == 0 else 4
    radix0 = ctx.diag()["sorts_radix"]
    for k in range(2):
        before = ctx.get_particles(k, native=True)
        perm = ctx.countingsort(k, want="perm")
        after = ctx.get_particles(k, native=True)
        c = before["grid"] - 1
        key = supertile_key(c, nx, ny) if push_order else c
        ref = np.argsort(key, kind="stable")
        if push_order:
            ref = _push_order(ref, V)
        assert np.array_equal(perm, ref)
        for f in ("grid", "dx", "dy", "ux", "uy", "uz", "pid"):
            assert np.array_equal(after[f], before[f][perm]), f
    if push_order and boxed:
        # both took the tile-box path.  The push leaves a tile's box unknown
        # when it spans half the domain or more: in reference order a tile is
        # a strip of whole columns; at 300 x 257 x 7 a push-order tile spans
        # 9 supertiles and some of them the height's half.  Those sorts take
        # the radix path.
        assert ctx.diag()["sorts_radix"] == radix0
    ctx.close()


def test_countingsort_empty(gpu):
    import pbpic_amd
    ctx = pbpic_amd.Context(8, 8, 0.1, 0.1, 0.01, 15.0, 25.0, 0.1)
    k = ctx.add_species(1, 10)
    counts, perm = ctx.countingsort(k, want=True)
    assert counts.sum() == 0 and len(perm) == 0


# ------------------------------------------------------------- deposit
@pytest.mark.parametrize("compat", [7, 0])
def test_currentdeposit_batch(gpu, compat):
    """currentdeposit (pic2d.f90:1226-1296), every crossing pattern, incl.
    boundary cells where the reference drops jy at column mx+2."""
    _say(f"deposit compat={compat}")
    nx, ny = 10, 9
    rng = np.random.default_rng(5)
    n = 4000
    i1 = rng.integers(1, nx + 1, n)
    j1 = rng.integers(1, ny + 1, n)
    a = rng.integers(-1, 2, n)
    b = rng.integers(-1, 2, n)
    i1[:50] = nx
    a[:50] = 1                      # right-boundary crossers (jy column mx+2)
    dx1 = rng.random(n) - 0.5
    dy1 = rng.random(n) - 0.5
    dx2 = np.where(a == 0, np.clip(dx1 + 0.3 * (rng.random(n) - 0.5), -0.5, 0.5),
                   -a * 0.5 + a * 0.4 * rng.random(n))
    dy2 = np.where(b == 0, np.clip(dy1 + 0.3 * (rng.random(n) - 0.5), -0.5, 0.5),
                   -b * 0.5 + b * 0.4 * rng.random(n))
    q = rng.choice([-1, 1], n)
    vz = rng.standard_normal(n)
    args = (i1, i1 + a, j1, j1 + b, dx1, dx2, dy1, dy2, q, vz)
    import pbpic_amd
    ctx = pbpic_amd.Context(nx, ny, 0.1, 0.12, 0.003, 15.0, 25.0, 0.05, compat=compat)
    ctx.currentdeposit(*args)
    o = oracle.Oracle(nx, ny, 0.1, 0.12, 0.003, 15.0, 25.0, 0.05, compat=compat)
    o.deposit_batch(*args)
    o.finish_current()
    for f in ("jx", "jy", "jz", "rho"):
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < 1e-13, f


@pytest.mark.parametrize("precision", [0, 1])
def test_currentdeposit_hot_cell(gpu, precision):
    """300,000 same-sign contributions into one cell and its neighbours: the
    fixed-point sums (one word split, two words FP64) stay exact far beyond
    the few hundred contributions a node gets in a step, and agree with the
    oracle's sequential FP64 sums (FP64 1e-12, split 1e-5)."""
    _say(f"deposit hot cell precision={precision}")
    nx, ny, n = 10, 9, 300000
    rng = np.random.default_rng(17)
    i1 = np.full(n, 5)
    j1 = np.full(n, 4)
    a = rng.integers(-1, 2, n)
    b = rng.integers(-1, 2, n)
    dx1 = rng.random(n) - 0.5
    dy1 = rng.random(n) - 0.5
    dx2 = np.where(a == 0, np.clip(dx1 + 0.3 * (rng.random(n) - 0.5), -0.5, 0.5),
                   -a * 0.5 + a * 0.4 * rng.random(n))
    dy2 = np.where(b == 0, np.clip(dy1 + 0.3 * (rng.random(n) - 0.5), -0.5, 0.5),
                   -b * 0.5 + b * 0.4 * rng.random(n))
    q = np.ones(n, dtype=np.int64)
    vz = np.abs(rng.standard_normal(n)) + 1.0
    args = (i1, i1 + a, j1, j1 + b, dx1, dx2, dy1, dy2, q, vz)
    import pbpic_amd
    ctx = pbpic_amd.Context(nx, ny, 0.1, 0.12, 0.003, 15.0, 25.0, 0.05, precision=precision,
                            compat=0)
    ctx.currentdeposit(*args)
    o = oracle.Oracle(nx, ny, 0.1, 0.12, 0.003, 15.0, 25.0, 0.05, compat=0)
    o.deposit_batch(*args)
    o.finish_current()
    tol = 1e-12 if precision == 0 else 1e-5
    for f in ("jx", "jy", "jz", "rho"):
        g, r = ctx.get_field(f), o.get_field(f)
        assert np.abs(r).max() > 1e3, f
        assert rel_norm(g, r) < tol, f


# ------------------------------------------------------------- fields
def test_fieldsolver_fp64(gpu):
    """fieldsolver (pic2d.f90:678-860) with random E, B, J: agrees to rounding."""
    case = make_case(nx=70, ny=37, ppc=1, seed=9)
    case["species"] = []
    rng = np.random.default_rng(2)
    jx, jy, jz = (rng.standard_normal((37, 70)) for _ in range(3))
    ctx = gpu_for(case)
    o = oracle_for(case)
    for name, plane in (("jx", jx), ("jy", jy), ("jz", jz)):
        ctx.set_field(name, plane)
        o.set_field(name, plane)
    ctx.fieldsolver()
    o.fieldsolver()
    for f in ("ex", "ey", "ez", "bx", "by", "bz"):
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < 1e-14, f
    ge, gb = ctx.fieldsenergy()
    oe, ob = o.fieldsenergy()
    assert abs(ge - oe) <= 1e-12 * abs(oe) and abs(gb - ob) <= 1e-12 * abs(ob)
    d, od = ctx.diag(), o.diag()
    assert abs(d["je_step"] - od["je_step"]) <= 1e-12 * abs(od["je_step"])


@pytest.mark.parametrize("precision", [0, 1])
def test_fieldsolver_many_tiles(gpu, precision):
    """fieldsolver (pic2d.f90:678-860) at 2048 x 2048 cells, where the FDTD
    launch has 8192 workgroups, far more than are resident at once: tiles
    that run late must still read their neighbours' E^n and B^n as halo, not
    fields a finished neighbour has already advanced.  Three solves with
    random E, B and J against the oracle: FP64 to 1e-14, the split form
    (FP32 increments on FP64 fields) to 1e-6."""
    nx = ny = 2048
    _say(f"fieldsolver {nx}x{ny} precision={precision}")
    rng = np.random.default_rng(31)
    case = dict(nx=nx, ny=ny, dx=0.1, dy=0.1, dt=0.2 / (15.0 * 5.0), cva=15.0, mime=25.0,
                in0=1.0 / 16, species=[], sort=False, ybound=0,
                fields={f: rng.standard_normal((ny, nx)) for f in ("ex", "ey", "ez", "bx", "by", "bz")})
    ctx = gpu_for(case, precision=precision)
    o = oracle_for(case)
    tol = 1e-14 if precision == 0 else 1e-6
    for it in range(3):
        for name in ("jx", "jy", "jz"):
            plane = rng.standard_normal((ny, nx))
            ctx.set_field(name, plane)
            o.set_field(name, plane)
        ctx.fieldsolver()
        o.fieldsolver()
    for f in ("ex", "ey", "ez", "bx", "by", "bz"):
        err = rel_norm(ctx.get_field(f), o.get_field(f))
        assert err < tol, (f, err)
    ctx.close()


def test_emfgrid_matches(gpu):
    case = make_case(nx=20, ny=16, ppc=1, seed=4)
    case["species"] = []
    ctx = gpu_for(case)
    o = oracle_for(case)
    for f in ("gex", "gey", "gez", "gbx", "gby", "gbz"):
        assert np.array_equal(ctx.get_field(f), o.get_field(f)), f


# ------------------------------------------------------------- one step
@pytest.mark.parametrize("compat", [7, 0])
@pytest.mark.parametrize("shuffle", [False, True])
def test_one_step_fp64(gpu, compat, shuffle):
    """zero J, pusher, fieldsolver (pic2d.f90:173-178) in FP64: particles to
    1e-12 element-wise (cells exact), J/rho/E/B to 1e-12 norm-wise.  shuffle
    feeds unsorted particles (the global-atomic fall-back path)."""
    _say(f"one step fp64 compat={compat} shuffle={shuffle}")
    case = make_case(nx=40, ny=33, ppc=12, seed=21, shuffle=shuffle)
    ctx = gpu_for(case, compat=compat)
    o = oracle_for(case, compat=compat)
    ctx.zero_current(); ctx.pusher()
    o.zero_current(); o.pusher()
    for f in ("jx", "jy", "jz", "rho"):
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < FP64_TOL, f
    ctx.fieldsolver(); o.fieldsolver()
    for f in ("ex", "ey", "ez", "bx", "by", "bz"):
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < FP64_TOL, f
    osp = oracle_species(o, case)
    for k in range(2):
        g = ctx.get_particles(k)
        assert np.array_equal(g["pid"], osp[k]["pid"])
        assert np.array_equal(g["grid"], osp[k]["grid"])
        for f in ("dx", "dy", "ux", "uy", "uz"):
            err = np.max(np.abs(g[f] - osp[k][f]) / np.maximum(np.abs(osp[k][f]), 1e-300))
            assert err < FP64_TOL, (k, f, err)
    d, od = ctx.diag(), o.diag()
    for key in ("ke_e", "ke_i", "w_e", "w_b", "je_step"):
        assert abs(d[key] - od[key]) <= 1e-12 * abs(od[key]) + 1e-300, key


def _by_pid(sp):
    order = np.argsort(sp["pid"])
    return {k: v[order] for k, v in sp.items()}


@pytest.mark.parametrize("precision", [0, 1])
def test_one_step_crowded_cell(gpu, precision):
    """Every particle of each species in one cell (15,360 per species, four
    push tiles each): each tile's LDS window takes up to 8,192 contributions
    on a node of a plane, the most a tile can put there, so the window sums
    reach their largest values (below 2^62 in the FP64 form's fixed point).
    One push against the oracle."""
    _say(f"one step crowded cell precision={precision}")
    case = make_case(nx=32, ny=30, ppc=16, seed=19, efield=0.05, bfield=0.5)
    for k, sp in enumerate(case["species"]):
        sp["grid"][:] = 30 * (11 + k) + 10          # cell (12 + k, 10), 1-based
    ctx = gpu_for(case, precision=precision, compat=0)
    o = oracle_for(case, compat=0)
    ctx.zero_current(); ctx.pusher()
    o.zero_current(); o.pusher()
    tol = FP64_TOL if precision == 0 else 1e-5
    for f in ("jx", "jy", "jz", "rho"):
        r = o.get_field(f)
        assert np.abs(r).max() > 0, f
        assert rel_norm(ctx.get_field(f), r) < tol, f
    assert ctx.diag()["tiles_spilled"] == 0
    if precision == 0:
        osp = oracle_species(o, case)
        for k in range(2):
            g, r = _by_pid(ctx.get_particles(k)), _by_pid(osp[k])
            assert np.array_equal(g["grid"], r["grid"])
            for f in ("dx", "dy", "ux", "uy", "uz"):
                err = np.max(np.abs(g[f] - r[f]) / np.maximum(np.abs(r[f]), 1e-300))
                assert err < FP64_TOL, (k, f, err)


@pytest.mark.parametrize("precision", [0, 1])
def test_push_adds_to_uploaded_current(gpu, precision):
    """A push adds its deposit to the current as it stands: current uploaded
    with set_field, then one push without zero_current, gives the upload plus
    the deposit the oracle's push makes from zero.  The upload goes into the
    fixed-point sums the push adds to (two words a node in the FP64 form).
    (The library holds the current unnormalized and applies in0 at its
    boundary, so a push adds in0 times its raw deposit.  The reference's
    pusher multiplies the whole planes by in0 after its loop, :1065-1069,
    and its time loop always starts it from zero planes, :173-176.)"""
    _say(f"push onto uploaded current precision={precision}")
    case = make_case(nx=40, ny=36, ppc=16, seed=41, efield=0.05, bfield=0.5)
    ctx = gpu_for(case, precision=precision, compat=0)
    o = oracle_for(case, compat=0)
    o.zero_current()
    o.pusher()
    rng = np.random.default_rng(8)
    up = {}
    for f in ("jx", "jy", "jz", "rho"):
        up[f] = np.abs(o.get_field(f)).max() * rng.standard_normal((36, 40))
        ctx.set_field(f, up[f])
    ctx.pusher()
    tol = FP64_TOL if precision == 0 else 1e-5
    for f in ("jx", "jy", "jz", "rho"):
        g, dep = ctx.get_field(f), o.get_field(f)
        assert rel_norm(g, up[f] + dep) < tol, f
        assert rel_norm(g - up[f], dep) < 10 * tol, f


@pytest.mark.parametrize("nx,ny", [(32, 64), (45, 37)])
@pytest.mark.parametrize("precision,compat", [(0, 7), (0, 0), (1, 7)])
def test_one_step_full_tiles(gpu, precision, compat, nx, ny):
    """Sorted species of full 4096-particle tiles, 64 particles per cell: at
    32 x 64 cells 32 tiles of one 8 x 8 supertile each, 12 of them interior;
    at 45 x 37 (odd sizes: partial supertiles at the far edges) 26 full tiles
    and a ragged one.  The push of whole tiles in the push order against the
    oracle's push of the same particles in cell order, matched by particle
    id.  FP64 to 1e-12; split precision to the bars of test_one_step_split32."""
    _say(f"one step full tiles precision={precision} compat={compat} {nx}x{ny}")
    case = make_case(nx=nx, ny=ny, ppc=64, seed=31, efield=0.05, bfield=0.5)
    ctx = gpu_for(case, precision=precision, compat=compat)
    o = oracle_for(case, compat=compat)
    ctx.countingsort(); o.countingsort()
    ctx.zero_current(); ctx.pusher()
    o.zero_current(); o.pusher()
    tol_j = FP64_TOL if precision == 0 else 1e-5
    for f in ("jx", "jy", "jz", "rho"):
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < tol_j, f
    osp = oracle_species(o, case)
    for k in range(2):
        g, r = _by_pid(ctx.get_particles(k)), _by_pid(osp[k])
        assert np.array_equal(g["pid"], r["pid"])
        if precision == 0:
            assert np.array_equal(g["grid"], r["grid"])
            for f in ("dx", "dy", "ux", "uy", "uz"):
                err = np.max(np.abs(g[f] - r[f]) / np.maximum(np.abs(r[f]), 1e-300))
                assert err < FP64_TOL, (k, f, err)
        else:
            same = g["grid"] == r["grid"]
            near_face = (np.abs(np.abs(r["dx"]) - 0.5) < 1e-5) | (np.abs(np.abs(r["dy"]) - 0.5) < 1e-5)
            assert np.all(same | near_face)
            for f in ("dx", "dy"):
                assert np.max(np.abs(g[f] - r[f])[same]) < 2e-6, f
            for f in ("ux", "uy", "uz"):
                assert np.max(np.abs(g[f] - r[f])) < 1e-5 * np.max(np.abs(r[f])), f
    d, od = ctx.diag(), o.diag()
    tol_d = 1e-12 if precision == 0 else 1e-5
    for key in ("ke_e", "ke_i", "je_step"):
        assert abs(d[key] - od[key]) <= tol_d * abs(od[key]), key


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nx,ny,path", [(30, 26, "middle pass"), (40, 36, "last pass"),
                                        (72, 70, "global atomics")])
def test_one_step_oversized_windows(gpu, precision, nx, ny, path):
    """Unsorted particles: every 4096-particle tile spans the whole domain, a
    window of (nx+3) x (ny+3) nodes.  At 30 x 26 (957 nodes) it exceeds the
    split form's first window (896) and fits its middle one (1243 with the
    rho plane); the FP64 form's first launch, at two workgroups per CU, holds
    it (998 with the rho plane).  At 40 x 36 (1677) only the last launch's
    window holds it (2377, 1920); at 72 x 70 none, and the tiles deposit with
    global atomics.  All against the oracle."""
    _say(f"oversized windows precision={precision} {nx}x{ny} ({path})")
    case = make_case(nx=nx, ny=ny, ppc=12, seed=13, shuffle=True)
    ctx = gpu_for(case, precision=precision, compat=7)
    o = oracle_for(case, compat=7)
    ctx.zero_current(); ctx.pusher()
    o.zero_current(); o.pusher()
    tol = FP64_TOL if precision == 0 else 1e-5
    for f in ("jx", "jy", "jz", "rho"):
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < tol, f
    osp = oracle_species(o, case)
    for k in range(2):
        g = ctx.get_particles(k)
        if precision == 0:
            assert np.array_equal(g["grid"], osp[k]["grid"])
            assert np.max(np.abs(g["ux"] - osp[k]["ux"])) < 1e-12 * np.max(np.abs(osp[k]["ux"]))
    spilled = ctx.diag()["tiles_spilled"]
    assert (spilled > 0) if path == "global atomics" else (spilled == 0), spilled


@pytest.mark.parametrize("precision", [0, 1])
def test_gauss_residual_tiled_charge(gpu, precision):
    """The Gauss residual's charge deposit, per push tile: after a sort (no
    recorded tile boxes: per-particle atomics) and after a push (tile boxes
    recorded: LDS node histogram per tile), L2 and Linf norms of
    divE - c^2 in0 rho against the oracle's.  FP64 to 1e-9 (atomic sum order),
    split precision to 1e-5."""
    _say(f"gauss tiled charge precision={precision}")
    case = make_case(nx=32, ny=64, ppc=64, seed=41, efield=0.05, bfield=0.5)
    ctx = gpu_for(case, precision=precision, compat=0)
    o = oracle_for(case, compat=0)
    ctx.countingsort(); o.countingsort()
    tol = 1e-9 if precision == 0 else 1e-5
    for stage in ("sorted", "pushed"):
        if stage == "pushed":
            ctx.zero_current(); ctx.pusher(); ctx.fieldsolver()
            o.zero_current(); o.pusher(); o.fieldsolver()
        d, od = ctx.diag(), o.diag()
        for key in ("gauss_l2", "gauss_linf"):
            assert abs(d[key] - od[key]) <= tol * od[key], (stage, key, d[key], od[key])


def test_multi_step_fp64_diagnostics(gpu):
    """30 steps with a sort every 10: diagnostics track the oracle (trajectories
    stay within rounding for a few dozen steps of a weakly driven plasma)."""
    case = make_case(nx=32, ny=32, ppc=16, seed=5, efield=0.0, bfield=0.0)
    ctx = gpu_for(case, sort_every=10)
    o = oracle_for(case)
    ctx.countingsort(); o.countingsort()
    ctx.step(30)
    o.step(30, sort_every=10)
    d, od = ctx.diag(), o.diag()
    for key in ("ke_e", "ke_i", "w_e", "w_b", "je_int"):
        assert abs(d[key] - od[key]) <= 1e-9 * abs(od[key]), (key, d[key], od[key])
    assert abs(d["gauss_l2"] - od["gauss_l2"]) <= 1e-9 * od["gauss_l2"]


# ------------------------------------------------------- split precision
def test_one_step_split32(gpu):
    """Split precision (int cell + FP32 offset, FP32 push/deposit) against the
    FP64 oracle: offsets within 2e-6 cells, momenta within 1e-5 relative,
    cells exact except where an FP64 offset lies within 1e-5 of a cell face."""
    case = make_case(nx=40, ny=33, ppc=12, seed=22)
    ctx = gpu_for(case, precision=1, compat=7)
    o = oracle_for(case, compat=7)
    ctx.zero_current(); ctx.pusher(); ctx.fieldsolver()
    o.zero_current(); o.pusher(); o.fieldsolver()
    osp = oracle_species(o, case)
    for k in range(2):
        g = ctx.get_particles(k)
        same = g["grid"] == osp[k]["grid"]
        near_face = (np.abs(np.abs(osp[k]["dx"]) - 0.5) < 1e-5) | (np.abs(np.abs(osp[k]["dy"]) - 0.5) < 1e-5)
        assert np.all(same | near_face)
        for f in ("dx", "dy"):
            assert np.max(np.abs(g[f] - osp[k][f])[same]) < 2e-6, f
        for f in ("ux", "uy", "uz"):
            assert np.max(np.abs(g[f] - osp[k][f])) < 1e-5 * np.max(np.abs(osp[k][f])), f
    for f in ("jx", "jy", "jz"):
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < 1e-5, f
    for f in ("ex", "ey", "ez", "bx", "by", "bz"):
        assert rel_norm(ctx.get_field(f), o.get_field(f)) < 1e-6, f


# ------------------------------------------------------------- loader
def test_device_loader_matches_oracle(gpu):
    """The device Philox loader produces the oracle loader's particles: cells
    and offsets bit-identical, Box-Muller momenta within 1e-13."""
    import pbpic_amd
    nx, ny, ppc = 17, 9, 5
    ctx = pbpic_amd.Context(nx, ny, 0.1, 0.1, 0.01, 15.0, 25.0, 0.2)
    k = ctx.add_species(-1, nx * ny * ppc)
    ctx.init_maxwellian(k, ppc, 0.1, 0.2, counter_stream=True, seed=99, stream=3)
    g = ctx.get_particles(k)
    r = oracle.init_maxwellian(nx, ny, ppc, 0.1, 0.2, counter_stream=True, seed=99, stream=3)
    assert np.array_equal(g["grid"], r["grid"])
    assert np.array_equal(g["dx"], r["dx"]) and np.array_equal(g["dy"], r["dy"])
    for f in ("ux", "uy", "uz"):
        assert np.max(np.abs(g[f] - r[f])) < 1e-13


# ---------------------------------------------------------- error paths
def test_cfl_violation_reported(gpu):
    """A particle moving more than one cell (the reference mishandles it
    silently, pic2d.f90:1008-1023) is reported as PBPIC_ERR_CFL."""
    import pbpic_amd
    ctx = pbpic_amd.Context(8, 8, 0.01, 0.01, 0.01, 15.0, 25.0, 0.1)
    k = ctx.add_species(-1, 4)
    ctx.set_particles(k, [1, 2, 3, 4], [0.0] * 4, [0.0] * 4, [50.0, 0, 0, 0], [0.0] * 4, [0.0] * 4)
    ctx.zero_current(); ctx.pusher()
    with pytest.raises(pbpic_amd.PbpicError) as e:
        ctx.sync()
    assert e.value.code == -4
    ctx.sync()   # flags are cleared once reported


def test_nonfinite_reported(gpu):
    import pbpic_amd
    ctx = pbpic_amd.Context(8, 8, 0.1, 0.1, 0.001, 15.0, 25.0, 0.1)
    k = ctx.add_species(1, 2)
    ctx.set_particles(k, [1, 2], [0.0, 0.0], [0.0, 0.0], [np.nan, 0.0], [0.0, 0.0], [0.0, 0.0])
    ctx.zero_current(); ctx.pusher()
    with pytest.raises(pbpic_amd.PbpicError) as e:
        ctx.sync()
    assert e.value.code == -5


@pytest.mark.parametrize("precision", [0, 1])
def test_tile_box_of_half_the_domain(gpu, precision):
    """A tile whose cells span exactly half the domain after a push (rows 0..8
    of 16): such a box is ambiguous under the periodic unwrap and must not be
    recorded for the next push (which would report PBPIC_ERR_STATE).  One full
    tile per species in column 3, rows 0..7 in cell order; the particles of
    row 7 move up, no others leave.  Three pushes against the oracle."""
    _say(f"tile box of half the domain precision={precision}")
    rng = np.random.default_rng(5)
    case = make_case(nx=8, ny=16, ppc=1, seed=5, efield=0.0, bfield=0.0)
    sp = []
    for k in range(2):
        n = 4096
        row = np.repeat(np.arange(8), n // 8)
        p = dict(grid=(3 * 16 + row + 1).astype(np.int32),
                 dx=rng.uniform(-0.4, 0.4, n), dy=rng.uniform(-0.5, 0.49, n),
                 ux=np.zeros(n), uy=np.where(row == 7, 0.3, 0.0), uz=np.zeros(n),
                 pid=(np.arange(1, n + 1) + k * 10_000_000).astype(np.int32))
        sp.append(p)
    case["species"] = sp
    ctx = gpu_for(case, precision=precision, compat=7)
    o = oracle_for(case, compat=7)
    for step in range(3):
        ctx.zero_current(); ctx.pusher()
        o.zero_current(); o.pusher()
        ctx.sync()
    osp = oracle_species(o, case)
    for k in range(2):
        g = ctx.get_particles(k)
        crossed = np.sum((g["grid"] - 1) % 16 >= 8)
        assert crossed > 0
        if precision == 0:
            assert np.array_equal(g["grid"], osp[k]["grid"])
            assert np.max(np.abs(g["dy"] - osp[k]["dy"])) < 1e-12
    # jx is zero up to rounding (no particle moves in x): currents are held
    # to the tolerance relative to the largest current plane
    tol = FP64_TOL if precision == 0 else 1e-5
    scale = max(np.linalg.norm(o.get_field(f)) for f in ("jx", "jy", "jz"))
    for f in ("jx", "jy", "jz"):
        assert np.linalg.norm(ctx.get_field(f) - o.get_field(f)) < tol * scale, f


def test_record_writing_push_matches_packing_sort(gpu):
    """Split form: on a step whose sort follows, the push writes the sort's
    24-byte particle records and the sort skips its packing pass.  One step
    with a sort every step (its push writes the records) against one step
    without an in-step sort and then the sort on its own (which packs the
    records).  The pushed particles depend only on the initial fields, so
    they, their order and the cell counts agree bit for bit.  The fields
    agree to FP32 rounding: tiles add their current windows into J with
    FP32 atomics in no fixed order.  45 x 37 cells at 64 particles per cell:
    full tiles on the fast path, edge tiles on the generic one and a ragged
    last tile."""
    _say("record-writing push vs packing sort")
    case = make_case(nx=45, ny=37, ppc=64, seed=5, efield=0.05, bfield=0.5)
    a = gpu_for(case, precision=1, compat=7, sort_every=1)
    b = gpu_for(case, precision=1, compat=7, sort_every=0)
    a.countingsort(); b.countingsort()
    a.step(1)
    b.step(1)
    b.countingsort()
    for k in range(2):
        pa, pb = a.get_particles(k, native=True), b.get_particles(k, native=True)
        for f in ("grid", "dx", "dy", "ux", "uy", "uz", "pid"):
            assert np.array_equal(pa[f], pb[f]), (k, f)
        ca, _ = a.countingsort(k, want=True)
        cb, _ = b.countingsort(k, want=True)
        assert np.array_equal(ca, cb), k
    for f in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"):
        assert rel_norm(a.get_field(f), b.get_field(f)) < 1e-6, f
    assert a.counters()[0] == b.counters()[0] == 1


@pytest.mark.parametrize("precision", [0, 1])
def test_steps_with_an_empty_species(gpu, precision):
    """A species with no particles beside one with a ragged last tile (3840
    electrons): 12 steps with a sort every 5 against the oracle, and the
    empty species' particle I/O and counts."""
    _say(f"empty species precision={precision}")
    case = make_case(nx=24, ny=20, ppc=8, seed=3)
    e = case["species"][0]
    case["species"][0] = {k: v[:0] for k, v in e.items()}
    ctx = gpu_for(case, precision=precision, compat=7, sort_every=5)
    o = oracle_for(case, compat=7)
    assert ctx.num_particles(0) == 0
    ctx.countingsort(); o.countingsort()
    ctx.step(12)
    o.step(12, sort_every=5)
    d, od = ctx.diag(), o.diag()
    tol = 1e-10 if precision == 0 else 1e-4
    for key in ("ke_e", "w_e", "w_b", "je_int"):
        assert abs(d[key] - od[key]) <= tol * abs(od[key]), (key, d[key], od[key])
    assert d["ke_i"] == 0.0
    g = ctx.get_particles(0)
    assert len(g["grid"]) == 0
    assert ctx.energy_spectrum(0).sum() == 0
    counts, perm = ctx.countingsort(0, want=True)
    assert counts.sum() == 0 and len(perm) == 0
