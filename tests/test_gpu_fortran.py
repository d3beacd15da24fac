"""The Fortran drop-in: pbpic_amd/fortran/pic2d_gpu runs the reference's time
loop (pic2d.f90:171-188) through module pbpic_gpu (pusher, fieldsolver,
countingsort, fieldsenergy) on the thermal plasma of the CPU-runnable case
(128 x 128, 16 particles per cell and species).  Its energies match the same
run through the Python binding of the same C ABI: to the printed digits after
10 steps, and within 1% after 100 steps (the runs' trajectories diverge once
the order of the atomic current sums differs; the plasma heats by a few
percent over the run, dx being 5 Debye lengths)."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "pbpic_amd", "fortran", "pic2d_gpu")

pytestmark = pytest.mark.gpu


def _run_driver(ntime, precision, mode=0):
    if not os.path.exists(DRIVER):
        pytest.skip("pic2d_gpu is not built (python -m pbpic_amd.build --fortran)")
    print(f"fortran driver ntime={ntime} precision={precision} mode={mode}", flush=True)
    out = subprocess.run([DRIVER, str(ntime), str(precision), str(mode)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    rows = []
    for line in out.stdout.splitlines():
        f = line.split()
        if f and f[0] == "itime":
            rows.append((int(f[1]), [float(x) for x in f[2:6]], float(f[7])))
    return rows


def _python_run(nsteps, precision):
    """The driver's run through pbpic_amd.Context (same seeds and order)."""
    import pbpic_amd
    nx = ny = 128
    ppg, mime, cva = 16, 25.0, 15.0
    dx = 0.5 / math.sqrt(mime)
    dt = 0.2 / (cva * math.sqrt(mime))
    ctx = pbpic_amd.Context(nx, ny, dx, dx, dt, cva, mime, 1.0 / ppg, precision=precision,
                            sort_every=50)
    ion = ctx.add_species(1, nx * ny * ppg)
    ele = ctx.add_species(-1, nx * ny * ppg)
    ctx.init_maxwellian(ion, ppg, 0.02, 0.0, counter_stream=False, seed=20260101, stream=1)
    ctx.init_maxwellian(ele, ppg, 0.1, 0.0, counter_stream=False, seed=20260101, stream=0)
    ctx.countingsort()
    ctx.step(nsteps)
    d = ctx.diag()
    ctx.close()
    return [d["ke_e"], d["ke_i"], d["w_e"], d["w_b"]]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("precision", [0, 1])
def test_fortran_driver_matches_python_binding(gpu, precision, mode):
    """Mode 0 runs the reference's loop body (zerocurrent, pusher,
    fieldsolver, countingsort every 50); mode 1 one pbpic_step per step."""
    rows = _run_driver(100, precision, mode)
    assert [r[0] for r in rows] == list(range(10, 101, 10))
    tot = np.array([sum(r[1]) for r in rows])
    assert np.all(np.isfinite(tot))
    # the first diagnostics line equals the Python binding's run of 10 steps
    ref = _python_run(10, precision)
    tol = 2e-8 if precision == 0 else 1e-4          # the driver prints 9 digits
    for k in range(4):
        assert abs(rows[0][1][k] - ref[k]) <= tol * abs(ref[k]) + 1e-12, (k, rows[0][1], ref)
    ref100 = _python_run(100, precision)
    assert abs(tot[-1] - sum(ref100)) <= 1e-2 * sum(ref100), (tot[-1], ref100)


def _write_input(path, case, dep):
    sp = case["species"]
    cat = lambda k: np.concatenate([s[k] for s in sp])
    q = np.concatenate([np.full(len(s["grid"]), qq, np.int32) for s, qq in zip(sp, (1, -1))])
    n = len(q)
    with open(path, "wb") as f:
        np.array([case["nx"], case["ny"], n, len(dep[0])], np.int32).tofile(f)
        np.array([case[k] for k in ("dx", "dy", "dt", "cva", "mime", "in0")], np.float64).tofile(f)
        for a in (cat("grid"), q, cat("pid")):
            np.asarray(a, np.int32).tofile(f)
        for k in ("dx", "dy", "ux", "uy", "uz"):
            np.asarray(cat(k), np.float64).tofile(f)
        for k in ("ex", "ey", "ez", "bx", "by", "bz"):
            np.asarray(case["fields"][k], np.float64).tofile(f)      # (ny, nx): x fastest
        for a in dep[:5]:
            np.asarray(a, np.int32).tofile(f)
        for a in dep[5:]:
            np.asarray(a, np.float64).tofile(f)


def _read_output(path, nx, ny):
    with open(path, "rb") as f:
        n = int(np.fromfile(f, np.int32, 1)[0])
        ints = np.fromfile(f, np.int32, 3 * n).reshape(3, n)
        reals = np.fromfile(f, np.float64, 5 * n).reshape(5, n)
        planes = np.fromfile(f, np.float64, 12 * nx * ny).reshape(12, ny, nx)
    p = dict(grid=ints[0], q=ints[1], pid=ints[2], dx=reals[0], dy=reals[1], ux=reals[2],
             uy=reals[3], uz=reals[4])
    names = ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "djx", "djy", "djz")
    return p, dict(zip(names, planes))


@pytest.mark.parametrize("kind", ["r8", "r4"])
def test_reference_state_through_the_module(gpu, tmp_path, kind):
    """tests/fortran/ref_state_step.f90 declares the reference's `type
    particle`, `type fields` and `type density` (pic2d.f90:47-67) in default
    real, fills ptl(:) and emf(0:mx+1,0:my+1), and moves them through module
    pbpic_gpu (pbpic_put_ptl / pbpic_put_field, pbpic_take_ptl /
    pbpic_take_field), runs one time-loop body (zerocurrent, pusher,
    fieldsolver) and then a batch of currentdeposit calls with default-real
    actuals.  Built with -fdefault-real-8 (kind r8, FP64 context): particles
    to 1e-12 of the oracle by pid with cells exact, fields and currents to
    1e-12 norm-wise, the deposit to 1e-13.  Built with the reference's 4-byte
    default real (kind r4, split context, the state rounded to FP32 on the
    way in): positions to 1e-5 cells, momenta to 1e-5, fields to 1e-6 and
    currents to 1e-4 relative."""
    import oracle
    from cases import make_case, oracle_for, rel_norm
    exe = os.path.join(ROOT, "tests", "fortran", "ref_state_step_" + kind)
    if not os.path.exists(exe):
        pytest.skip("ref_state_step is not built (python -m pbpic_amd.build --fortran)")
    print(f"fortran reference state kind={kind}", flush=True)
    case = make_case(nx=40, ny=33, ppc=12, seed=21)
    nx, ny = case["nx"], case["ny"]
    rng = np.random.default_rng(5)
    m = 3000
    i1, j1 = rng.integers(1, nx + 1, m), rng.integers(1, ny + 1, m)
    a, b = rng.integers(-1, 2, m), rng.integers(-1, 2, m)
    dx1, dy1 = rng.random(m) - 0.5, rng.random(m) - 0.5
    dx2 = np.where(a == 0, np.clip(dx1 + 0.3 * (rng.random(m) - 0.5), -0.5, 0.5),
                   -a * 0.5 + a * 0.4 * rng.random(m))
    dy2 = np.where(b == 0, np.clip(dy1 + 0.3 * (rng.random(m) - 0.5), -0.5, 0.5),
                   -b * 0.5 + b * 0.4 * rng.random(m))
    qd = rng.choice([-1, 1], m)
    vz = rng.standard_normal(m)
    dep = (i1, i1 + a, j1, j1 + b, qd, dx1, dx2, dy1, dy2, vz)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, case, dep)
    out = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    g, gf = _read_output(fout, nx, ny)
    o = oracle_for(case)
    o.zero_current()
    o.pusher()
    o.fieldsolver()
    op = o.get_particles()
    inv = np.empty(op["pid"].max() + 1, np.int64)
    inv[op["pid"]] = np.arange(len(op["pid"]))
    sel = inv[g["pid"]]
    assert np.array_equal(g["q"], op["q"][sel])
    r8 = kind == "r8"
    if r8:
        assert np.array_equal(g["grid"], op["grid"][sel])
        for k in ("dx", "dy"):
            assert np.abs(g[k] - op[k][sel]).max() <= 1e-12, k
    else:
        for k, comp, mm in (("dx", lambda c: c // ny, nx), ("dy", lambda c: c % ny, ny)):
            d = comp(g["grid"] - 1) + g[k] - (comp(op["grid"][sel] - 1) + op[k][sel])
            assert np.abs((d + 0.5 * mm) % mm - 0.5 * mm).max() <= 1e-5, k
    for k in ("ux", "uy", "uz"):
        ref = op[k][sel]
        assert np.abs(g[k] - ref).max() <= (1e-12 if r8 else 1e-5) * np.abs(ref).max(), k
    for f in ("ex", "ey", "ez", "bx", "by", "bz"):
        assert rel_norm(gf[f], o.get_field(f)) <= (1e-12 if r8 else 1e-6), f
    for f in ("jx", "jy", "jz"):
        assert rel_norm(gf[f], o.get_field(f)) <= (1e-12 if r8 else 1e-4), f
    od = oracle.Oracle(nx, ny, case["dx"], case["dy"], case["dt"], case["cva"], case["mime"],
                       case["in0"])
    od.deposit_batch(*dep[:4], dx1, dx2, dy1, dy2, qd, vz)
    od.finish_current()
    for f in ("jx", "jy", "jz"):
        assert rel_norm(gf["d" + f], od.get_field(f)) <= (1e-13 if r8 else 1e-5), f
