"""ctypes binding of the CPU restatement (TEST INFRASTRUCTURE ONLY).

This package is the oracle that the HIP path is checked against.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import it;
the product package pbpic_amd never does.  The C source restates
src/programs/pic2d.f90 of xiaocanli/pbpic line by line (see pic_oracle_impl.h).

Parity status of each oracle function (what pins it) is listed in
tests/test_oracle.py's docstrings.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpbpic_oracle.so")
FIELDS = {"ex": 0, "ey": 1, "ez": 2, "bx": 3, "by": 4, "bz": 5, "jx": 6, "jy": 7,
          "jz": 8, "rho": 9, "gex": 10, "gey": 11, "gez": 12, "gbx": 13, "gby": 14,
          "gbz": 15}
COMPAT_REFERENCE = 7
DIAG_NAMES = ["ke_e", "ke_i", "w_e", "w_b", "je_step", "je_int", "gauss_l2",
              "gauss_linf", "gauss_dl2", "gauss_dlinf", "steps"]

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        pi32, pdbl = P(ctypes.c_int32), P(ctypes.c_double)
        for pre in ("orc8", "orc4"):
            sig = {
                "create": (vp, [i32, i32, dbl, dbl, dbl, dbl, dbl, dbl, i32]),
                "destroy": (None, [vp]),
                "set_particles": (i32, [vp, i64, pi32, pdbl, pdbl, pdbl, pdbl, pdbl, pi32, pi32]),
                "num_particles": (i64, [vp]),
                "get_particles": (None, [vp, pi32, pdbl, pdbl, pdbl, pdbl, pdbl, pi32, pi32]),
                "set_field": (None, [vp, i32, pdbl]),
                "get_field": (None, [vp, i32, pdbl]),
                "zero_current": (None, [vp]),
                "pusher": (None, [vp]),
                "deposit_batch": (None, [vp, i64, pi32, pi32, pi32, pi32, pdbl, pdbl, pdbl, pdbl, pi32, pdbl]),
                "finish_current": (None, [vp]),
                "fieldsolver": (None, [vp]),
                "fieldsenergy": (None, [vp, pdbl]),
                "countingsort": (None, [vp, pi32, pi32]),
                "step": (None, [vp, i32, i32]),
                "diag": (None, [vp, dbl, pdbl]),
                "set_threads": (None, [vp, i32]),
                "set_ybound": (i32, [vp, i32]),
                "set_filters": (i32, [vp, i32, i32]),
            }
            for name, (res, args) in sig.items():
                f = getattr(L, f"{pre}_{name}")
                f.restype = res
                f.argtypes = args
        L.orc_init_maxwellian.restype = None
        L.orc_init_maxwellian.argtypes = [i32, i32, i32, dbl, dbl, i32, ctypes.c_uint64,
                                          ctypes.c_uint32, pi32, pdbl, pdbl, pdbl, pdbl, pdbl]
        L.orc_harris_count.restype = i64
        L.orc_harris_count.argtypes = [i32, i32, i32, dbl, dbl, i32, i32, pdbl, P(ctypes.c_int64)]
        L.orc_init_harris.restype = None
        L.orc_init_harris.argtypes = [i32, i32, i32, dbl, dbl, i32, i32, pdbl, i32, ctypes.c_uint64,
                                      ctypes.c_uint32, pi32, pdbl, pi32, pdbl, pdbl]
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct)) if a is not None else None


class Oracle:
    """The reference step on the CPU.  prec=8: real(8); prec=4: default real.

    Particles are the reference's single interleaved array with a per-particle
    charge q (`type particle`, pic2d.f90:47-50)."""

    def __init__(self, nx, ny, dx, dy, dt, cva, mime, in0, prec=8, compat=COMPAT_REFERENCE,
                 threads=1, ybound=0, marder=False, cherenkov=False):
        self.L = lib()
        self.pre = "orc8" if prec == 8 else "orc4"
        self.nx, self.ny = nx, ny
        self.h = self._f("create")(nx, ny, dx, dy, dt, cva, mime, in0, compat)
        if threads > 1:
            self._f("set_threads")(self.h, threads)
        if ybound:
            if self._f("set_ybound")(self.h, int(ybound)):
                raise ValueError("conducting y walls need ny >= 4")
        if marder or cherenkov:
            self._f("set_filters")(self.h, int(marder), int(cherenkov))

    def _f(self, name):
        return getattr(self.L, f"{self.pre}_{name}")

    def close(self):
        if self.h:
            self._f("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_filters(self, marder, cherenkov):
        self._f("set_filters")(self.h, int(marder), int(cherenkov))

    def set_threads(self, n):
        self._f("set_threads")(self.h, n)

    def set_particles(self, grid, dx, dy, ux, uy, uz, q, pid=None):
        g = np.ascontiguousarray(grid, np.int32)
        arrs = [np.ascontiguousarray(a, np.float64) for a in (dx, dy, ux, uy, uz)]
        qq = np.ascontiguousarray(q, np.int32)
        pd = np.ascontiguousarray(pid, np.int32) if pid is not None else None
        rc = self._f("set_particles")(self.h, len(g), _p(g, ctypes.c_int32),
                                      *[_p(a, ctypes.c_double) for a in arrs],
                                      _p(qq, ctypes.c_int32), _p(pd, ctypes.c_int32))
        if rc:
            raise MemoryError("oracle set_particles failed")

    def num_particles(self):
        return self._f("num_particles")(self.h)

    def get_particles(self):
        n = self.num_particles()
        g = np.empty(n, np.int32)
        a = [np.empty(n, np.float64) for _ in range(5)]
        q = np.empty(n, np.int32)
        pid = np.empty(n, np.int32)
        self._f("get_particles")(self.h, _p(g, ctypes.c_int32), *[_p(x, ctypes.c_double) for x in a],
                                 _p(q, ctypes.c_int32), _p(pid, ctypes.c_int32))
        return dict(grid=g, dx=a[0], dy=a[1], ux=a[2], uy=a[3], uz=a[4], q=q, pid=pid)

    def set_field(self, name, plane):
        a = np.ascontiguousarray(plane, np.float64)
        assert a.shape == (self.ny, self.nx)
        self._f("set_field")(self.h, FIELDS[name], _p(a, ctypes.c_double))

    def get_field(self, name):
        a = np.empty((self.ny, self.nx), np.float64)
        self._f("get_field")(self.h, FIELDS[name], _p(a, ctypes.c_double))
        return a

    def zero_current(self):
        self._f("zero_current")(self.h)

    def pusher(self):
        self._f("pusher")(self.h)

    def deposit_batch(self, i1, i2, j1, j2, dx1, dx2, dy1, dy2, q, vz):
        ints = [np.ascontiguousarray(np.atleast_1d(a), np.int32) for a in (i1, i2, j1, j2)]
        dbl = [np.ascontiguousarray(np.atleast_1d(a), np.float64) for a in (dx1, dx2, dy1, dy2)]
        qq = np.ascontiguousarray(np.atleast_1d(q), np.int32)
        v = np.ascontiguousarray(np.atleast_1d(vz), np.float64)
        self._f("deposit_batch")(self.h, len(qq), *[_p(a, ctypes.c_int32) for a in ints],
                                 *[_p(a, ctypes.c_double) for a in dbl], _p(qq, ctypes.c_int32),
                                 _p(v, ctypes.c_double))

    def finish_current(self):
        self._f("finish_current")(self.h)

    def fieldsolver(self):
        self._f("fieldsolver")(self.h)

    def fieldsenergy(self):
        out = (ctypes.c_double * 2)()
        self._f("fieldsenergy")(self.h, out)
        return out[0], out[1]

    def countingsort(self, want=False):
        n = self.num_particles()
        if not want:
            self._f("countingsort")(self.h, None, None)
            return None
        counts = np.empty(self.nx * self.ny, np.int32)
        perm = np.empty(max(n, 1), np.int32)
        self._f("countingsort")(self.h, _p(counts, ctypes.c_int32), _p(perm, ctypes.c_int32))
        return counts, perm[:n]

    def step(self, nsteps, sort_every=50):
        self._f("step")(self.h, nsteps, sort_every)

    def diag(self, rho_bg=0.0):
        out = (ctypes.c_double * 16)()
        self._f("diag")(self.h, rho_bg, out)
        return {k: out[i] for i, k in enumerate(DIAG_NAMES)}


def init_maxwellian(nx, ny, ppc, vth, vdrift_x=0.0, counter_stream=False, seed=20260101, stream=0):
    """Seeded loader (Philox4x32-10 + Box-Muller), cell-ordered, grid 1-based."""
    n = nx * ny * ppc
    g = np.empty(n, np.int32)
    a = [np.empty(n, np.float64) for _ in range(5)]
    lib().orc_init_maxwellian(nx, ny, ppc, vth, vdrift_x, int(bool(counter_stream)), seed, stream,
                              _p(g, ctypes.c_int32), *[_p(x, ctypes.c_double) for x in a])
    return dict(grid=g, dx=a[0], dy=a[1], ux=a[2], uy=a[3], uz=a[4])


def wall_rows(grid, nyp, ny):
    """Re-index 1-based cell ids of an nx x nyp lattice onto the rows 0..nyp-1
    of an nx x ny lattice (the domain of conducting y walls, ny = nyp + 2)."""
    g = np.asarray(grid, np.int64) - 1
    return ((g // nyp) * ny + g % nyp + 1).astype(np.int32)


HARRIS_KEYS = ("half_width", "nb", "b0", "psi0", "vthe_cs", "vthi_cs", "vthe_bg", "vthi_bg",
               "vde_z", "vdi_z")


def init_harris(nx, ny, ywall, dx, dy, h):
    """Seeded Harris-sheet loader (pic2d.f90:357-648): returns (ions,
    electrons, fields) with particle dicts (grid 1-based) and the six field
    planes (ny, nx) keyed ex..bz.  h: dict with geometry, ppc, HARRIS_KEYS,
    thermal_z, seed, stream."""
    L = lib()
    prm = np.array([float(h[k]) for k in HARRIS_KEYS], np.float64)
    ncs = ctypes.c_int64()
    n = L.orc_harris_count(nx, ny, int(ywall), dx, dy, int(h["geometry"]), int(h["ppc"]),
                           _p(prm, ctypes.c_double), ctypes.byref(ncs))
    gi, ge = np.empty(n, np.int32), np.empty(n, np.int32)
    pi, pe = np.empty(5 * n, np.float64), np.empty(5 * n, np.float64)
    f = np.empty(6 * nx * ny, np.float64)
    L.orc_init_harris(nx, ny, int(ywall), dx, dy, int(h["geometry"]), int(h["ppc"]),
                      _p(prm, ctypes.c_double), int(h.get("thermal_z", 0)), int(h.get("seed", 1)),
                      int(h.get("stream", 0)), _p(gi, ctypes.c_int32), _p(pi, ctypes.c_double),
                      _p(ge, ctypes.c_int32), _p(pe, ctypes.c_double), _p(f, ctypes.c_double))
    def unpack(g, a):
        return dict(grid=g, dx=a[:n], dy=a[n:2 * n], ux=a[2 * n:3 * n], uy=a[3 * n:4 * n],
                    uz=a[4 * n:])
    fields = {k: f[i * nx * ny:(i + 1) * nx * ny].reshape(ny, nx)
              for i, k in enumerate(("ex", "ey", "ez", "bx", "by", "bz"))}
    return unpack(gi, pi), unpack(ge, pe), fields, ncs.value


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)
