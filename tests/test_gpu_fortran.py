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
