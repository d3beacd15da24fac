"""pbpic_amd -- MI355X-native 2D electromagnetic PIC step of xiaocanli/pbpic.

Host-side mirror of the reference's per-step subroutines
(src/programs/pic2d.f90): :class:`Context` exposes ``pusher``,
``currentdeposit``, ``fieldsolver``, ``fieldsenergy`` and ``countingsort``
with the reference's argument meaning, on top of the C ABI of
``include/pbpic.h`` implemented by ``libpbpic.so`` (hand-written HIP kernels
for gfx950).  Where the reference aborts (MPI's default error handler), these
methods raise :class:`PbpicError`.  There is no CPU fallback: without the
built library every call fails.
"""
import ctypes
import os

import numpy as np

__all__ = ["Context", "PbpicError", "Params", "load_library", "FIELDS",
           "PREC_FP64", "PREC_SPLIT32", "COMPAT_REFERENCE", "DIAG_NAMES"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpbpic.so")

PREC_FP64, PREC_SPLIT32 = 0, 1
COMPAT_BORIS_OLD_U, COMPAT_RHO_MIRROR, COMPAT_JY_DROP = 1, 2, 4
COMPAT_REFERENCE = 7
FLAG_DEPOSIT_RHO, FLAG_TRACK_PID, FLAG_PUSH_ORDER = 1, 2, 4
FLAG_MARDER, FLAG_CHERENKOV = 8, 16
DEPOSIT_NOALLOC_MAX = 65536     # PBPIC_DEPOSIT_NOALLOC_MAX
YBOUND_PERIODIC, YBOUND_CONDUCTING = 0, 1
HARRIS_DOUBLE_X, HARRIS_SINGLE_Y = 0, 1
FIELDS = {"ex": 0, "ey": 1, "ez": 2, "bx": 3, "by": 4, "bz": 5, "jx": 6,
          "jy": 7, "jz": 8, "rho": 9, "gex": 10, "gey": 11, "gez": 12,
          "gbx": 13, "gby": 14, "gbz": 15}
DIAG_NAMES = ["ke_e", "ke_i", "w_e", "w_b", "je_step", "je_int", "gauss_l2",
              "gauss_linf", "gauss_dl2", "gauss_dlinf", "steps", "tiles_spilled",
              "sorts_radix"]
_ERR = {-1: "invalid argument", -2: "HIP runtime error", -3: "capacity",
        -4: "CFL violation", -5: "non-finite value", -6: "cell out of range",
        -7: "bad state", -8: "out of device memory", -9: "sort spin limit"}


class PbpicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pbpic error {code} ({_ERR.get(code, '?')}): {msg}")
        self.code = code


class Params(ctypes.Structure):
    """pbpic_params (include/pbpic.h); the run-time form of pic2d.f90's
    `module parameters` constants (:4-142)."""
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("compat", ctypes.c_int32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("dt", ctypes.c_double), ("cva", ctypes.c_double),
                ("mime", ctypes.c_double), ("in0", ctypes.c_double),
                ("rho_background", ctypes.c_double),
                ("sort_every", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("ybound", ctypes.c_int32), ("reserved", ctypes.c_int32 * 7)]


class Harris(ctypes.Structure):
    """pbpic_harris (include/pbpic.h): the Harris-sheet loader's parameters,
    the run-time form of the constants initparticles uses (pic2d.f90:22-30)."""
    _fields_ = [("geometry", ctypes.c_int32), ("ppc", ctypes.c_int32),
                ("half_width", ctypes.c_double), ("nb", ctypes.c_double),
                ("b0", ctypes.c_double), ("psi0", ctypes.c_double),
                ("vthe_cs", ctypes.c_double), ("vthi_cs", ctypes.c_double),
                ("vthe_bg", ctypes.c_double), ("vthi_bg", ctypes.c_double),
                ("vde_z", ctypes.c_double), ("vdi_z", ctypes.c_double),
                ("thermal_z", ctypes.c_int32), ("reserved0", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("stream", ctypes.c_uint32),
                ("reserved", ctypes.c_int32 * 7)]


_lib = None


def load_library(path=None):
    """Loads libpbpic.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise PbpicError(-7, f"{p} is missing; build it with `python -m pbpic_amd.build`")
    # One HIP runtime per process.  PyTorch wheels bundle their own
    # libamdhip64.so.7 (same soname as ROCm's): loaded first, it is the one the
    # library binds to.  Loaded after the library, it would be a second
    # runtime beside /opt/rocm's, and device pointers, streams and contexts
    # would not be shared between the two.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(p)
    P = ctypes.POINTER
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    pi32, pdbl, pflt, pi64 = P(ctypes.c_int32), P(ctypes.c_double), P(ctypes.c_float), P(ctypes.c_int64)
    sig = {
        "pbpic_params_default": (None, [P(Params)]),
        "pbpic_ctx_create": (i32, [P(Params), P(vp)]),
        "pbpic_ctx_destroy": (i32, [vp]),
        "pbpic_last_error": (ctypes.c_char_p, [vp]),
        "pbpic_sync": (i32, [vp]),
        "pbpic_add_species": (i32, [vp, i32, i64, pi32]),
        "pbpic_num_particles": (i32, [vp, i32, pi64]),
        "pbpic_set_particles": (i32, [vp, i32, i64, pi32, pdbl, pdbl, pdbl, pdbl, pdbl, pi32]),
        "pbpic_get_particles": (i32, [vp, i32, i64, pi64, pi32, pdbl, pdbl, pdbl, pdbl, pdbl, pi32]),
        "pbpic_set_particles_f32": (i32, [vp, i32, i64, pi32, pflt, pflt, pflt, pflt, pflt, pi32]),
        "pbpic_get_particles_f32": (i32, [vp, i32, i64, pi64, pi32, pflt, pflt, pflt, pflt, pflt, pi32]),
        "pbpic_init_maxwellian": (i32, [vp, i32, i32, dbl, dbl, i32, ctypes.c_uint64, ctypes.c_uint32]),
        "pbpic_init_harris": (i32, [vp, i32, i32, P(Harris)]),
        "pbpic_energy_spectrum": (i32, [vp, i32, i32, dbl, dbl, pi32]),
        "pbpic_number_density": (i32, [vp, pi32]),
        "pbpic_get_counters": (i32, [vp, pi64, pdbl]),
        "pbpic_set_counters": (i32, [vp, i64, pdbl]),
        "pbpic_set_field": (i32, [vp, i32, pdbl]),
        "pbpic_get_field": (i32, [vp, i32, pdbl]),
        "pbpic_zero_current": (i32, [vp]),
        "pbpic_push": (i32, [vp]),
        "pbpic_deposit": (i32, [vp, i64, pi32, pi32, pi32, pi32, pdbl, pdbl, pdbl, pdbl, pi32, pdbl]),
        "pbpic_deposit_noalloc": (i32, [vp, i64, pi32, pi32, pi32, pi32, pdbl, pdbl, pdbl, pdbl,
                                        pi32, pdbl]),
        "pbpic_fieldsolve": (i32, [vp]),
        "pbpic_fieldsenergy": (i32, [vp, pdbl]),
        "pbpic_kineticenergy": (i32, [vp, pdbl]),
        "pbpic_countingsort": (i32, [vp]),
        "pbpic_countingsort_species": (i32, [vp, i32, pi32, pi32]),
        "pbpic_step": (i32, [vp, i32]),
        "pbpic_diag": (i32, [vp, pdbl]),
        "pbpic_species_device_ptrs": (i32, [vp, i32, P(vp)]),
        "pbpic_field_device_ptrs": (i32, [vp, P(vp)]),
        "pbpic_stream": (vp, [vp]),
        "pbpic_profile": (i32, [vp, i32]),
        "pbpic_profile_read": (i32, [vp, pdbl, pi64]),
        "pbpic_set_particles_device": (i32, [vp, i32, i64, vp, vp, vp, vp, vp, vp]),
        "pbpic_set_field_device": (i32, [vp, i32, vp]),
        "pbpic_adjoint_forward": (i32, [vp, i32, i32, pdbl]),
        "pbpic_set_loss_region": (i32, [vp, pdbl]),
        "pbpic_adjoint_backward": (i32, [vp, P(vp), P(vp)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct)) if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Context:
    """One simulation on one GPU: fields, currents and particle species.

    Mirrors the state of pic2d.f90's `module parameters` (:4-142) and its
    step subroutines.  Field planes are (ny, nx) numpy arrays (x fastest,
    the owned region emf(1:mx,1:my)); particle cell ids use the reference's
    1-based `grid = my*(ix-1)+iy` (:422).
    """

    def __init__(self, nx, ny, dx, dy, dt, cva, mime, in0, *, precision=PREC_FP64,
                 compat=COMPAT_REFERENCE, sort_every=50, deposit_rho=True,
                 track_pid=False, push_order=True, rho_background=0.0, ybound=YBOUND_PERIODIC,
                 marder=False, cherenkov=False):
        self._lib = load_library()
        p = Params()
        self._lib.pbpic_params_default(ctypes.byref(p))
        p.nx, p.ny, p.dx, p.dy, p.dt = nx, ny, dx, dy, dt
        p.cva, p.mime, p.in0 = cva, mime, in0
        p.precision, p.compat, p.sort_every = precision, compat, sort_every
        p.flags = ((FLAG_DEPOSIT_RHO if deposit_rho else 0) | (FLAG_TRACK_PID if track_pid else 0) |
                   (FLAG_PUSH_ORDER if push_order else 0) | (FLAG_MARDER if marder else 0) |
                   (FLAG_CHERENKOV if cherenkov else 0))
        p.rho_background = rho_background
        p.ybound = {"periodic": YBOUND_PERIODIC, "conducting": YBOUND_CONDUCTING}.get(ybound, ybound)
        self.params = p
        self.nx, self.ny = nx, ny
        self.precision = precision
        h = ctypes.c_void_p()
        rc = self._lib.pbpic_ctx_create(ctypes.byref(p), ctypes.byref(h))
        if rc != 0:
            raise PbpicError(rc, "pbpic_ctx_create failed")
        self._h = h
        self.species_q = []

    # ------------------------------------------------------------ plumbing
    def _chk(self, rc):
        if rc != 0:
            raise PbpicError(rc, self._lib.pbpic_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pbpic_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def sync(self):
        self._chk(self._lib.pbpic_sync(self._h))

    # ------------------------------------------------------------ species
    def add_species(self, q, capacity):
        s = ctypes.c_int32()
        self._chk(self._lib.pbpic_add_species(self._h, int(q), int(capacity), ctypes.byref(s)))
        self.species_q.append(int(q))
        return s.value

    def num_particles(self, species):
        n = ctypes.c_int64()
        self._chk(self._lib.pbpic_num_particles(self._h, species, ctypes.byref(n)))
        return n.value

    @property
    def tracks_pid(self):
        return bool(self.params.flags & FLAG_TRACK_PID)

    def set_particles(self, species, grid, dx, dy, ux, uy, uz, pid=None):
        """Host particle state (grid 1-based).  float32 arrays go through the
        FP32 entry point, which is exact for the split-precision form."""
        n = len(grid)
        g = _c(grid, np.int32)
        pd = _c(pid, np.int32) if pid is not None else None
        f32 = all(np.asarray(a).dtype == np.float32 for a in (dx, dy, ux, uy, uz))
        if f32:
            arrs = [_c(a, np.float32) for a in (dx, dy, ux, uy, uz)]
            self._chk(self._lib.pbpic_set_particles_f32(
                self._h, species, n, _ptr(g, ctypes.c_int32),
                *[_ptr(a, ctypes.c_float) for a in arrs], _ptr(pd, ctypes.c_int32)))
            return
        arrs = [_c(a, np.float64) for a in (dx, dy, ux, uy, uz)]
        self._chk(self._lib.pbpic_set_particles(
            self._h, species, n, _ptr(g, ctypes.c_int32), *[_ptr(a, ctypes.c_double) for a in arrs],
            _ptr(pd, ctypes.c_int32)))

    def get_particles(self, species, native=False):
        """Host copy of a species (grid 1-based; pid zero unless tracked).
        native=True returns the offsets and momenta at the device precision
        (float32 for the split form) instead of float64."""
        n = self.num_particles(species)
        g = np.empty(n, np.int32)
        pid = np.empty(n, np.int32)
        nn = ctypes.c_int64()
        if native and self.precision == PREC_SPLIT32:
            a = [np.empty(n, np.float32) for _ in range(5)]
            self._chk(self._lib.pbpic_get_particles_f32(
                self._h, species, n, ctypes.byref(nn), _ptr(g, ctypes.c_int32),
                *[_ptr(x, ctypes.c_float) for x in a], _ptr(pid, ctypes.c_int32)))
        else:
            a = [np.empty(n, np.float64) for _ in range(5)]
            self._chk(self._lib.pbpic_get_particles(
                self._h, species, n, ctypes.byref(nn), _ptr(g, ctypes.c_int32),
                *[_ptr(x, ctypes.c_double) for x in a], _ptr(pid, ctypes.c_int32)))
        return dict(grid=g, dx=a[0], dy=a[1], ux=a[2], uy=a[3], uz=a[4], pid=pid)

    def init_maxwellian(self, species, ppc, vth, vdrift_x=0.0, counter_stream=False,
                        seed=20260101, stream=0):
        self._chk(self._lib.pbpic_init_maxwellian(self._h, species, int(ppc), float(vth),
                                                  float(vdrift_x), int(bool(counter_stream)),
                                                  int(seed), int(stream)))

    def init_harris(self, ion, electron, **kw):
        """Seeded Harris current-sheet loader (initparticles / initfields,
        pic2d.f90:357-648; pbpic_init_harris).  Keywords are the fields of
        :class:`Harris`; ``geometry`` may be "double_x" or "single_y"."""
        h = Harris()
        geo = kw.pop("geometry", HARRIS_DOUBLE_X)
        h.geometry = {"double_x": HARRIS_DOUBLE_X, "single_y": HARRIS_SINGLE_Y}.get(geo, geo)
        h.b0 = 1.0
        for k, v in kw.items():
            if k not in dict(Harris._fields_) or k.startswith("reserved"):
                raise TypeError(f"unknown Harris parameter {k}")
            setattr(h, k, v)
        self._chk(self._lib.pbpic_init_harris(self._h, ion, electron, ctypes.byref(h)))

    # ------------------------------------------------------------ fields
    def set_field(self, name, plane):
        a = _c(plane, np.float64)
        if a.shape != (self.ny, self.nx):
            raise ValueError(f"field plane must have shape (ny, nx)={(self.ny, self.nx)}")
        self._chk(self._lib.pbpic_set_field(self._h, FIELDS[name], _ptr(a, ctypes.c_double)))

    def get_field(self, name):
        a = np.empty((self.ny, self.nx), np.float64)
        self._chk(self._lib.pbpic_get_field(self._h, FIELDS[name], _ptr(a, ctypes.c_double)))
        return a

    # ------------------------------------------------ reference subroutines
    def zero_current(self):
        """`denjr = 0` at the top of the time loop (pic2d.f90:173-176)."""
        self._chk(self._lib.pbpic_zero_current(self._h))

    def pusher(self):
        """`pusher` (pic2d.f90:912-1217)."""
        self._chk(self._lib.pbpic_push(self._h))

    def currentdeposit(self, i1, i2, j1, j2, dx1, dx2, dy1, dy2, q0, uz, noalloc=False):
        """`currentdeposit(i1,i2,j1,j2,dx1,dx2,dy1,dy2,q0,uz)` (pic2d.f90:1226),
        batched over equal-length arrays; `uz` is the velocity vz passed at :1029.
        noalloc=True calls pbpic_deposit_noalloc (no device allocation, at most
        DEPOSIT_NOALLOC_MAX entries)."""
        ints = [_c(np.atleast_1d(a), np.int32) for a in (i1, i2, j1, j2)]
        dbl = [_c(np.atleast_1d(a), np.float64) for a in (dx1, dx2, dy1, dy2)]
        q = _c(np.atleast_1d(q0), np.int32)
        vz = _c(np.atleast_1d(uz), np.float64)
        n = len(q)
        f = self._lib.pbpic_deposit_noalloc if noalloc else self._lib.pbpic_deposit
        self._chk(f(self._h, n, *[_ptr(a, ctypes.c_int32) for a in ints],
                    *[_ptr(a, ctypes.c_double) for a in dbl], _ptr(q, ctypes.c_int32),
                    _ptr(vz, ctypes.c_double)))

    def fieldsolver(self):
        """`fieldsolver` (pic2d.f90:678-860)."""
        self._chk(self._lib.pbpic_fieldsolve(self._h))

    def fieldsenergy(self):
        """`fieldsenergy` (pic2d.f90:653-673): (electric, magnetic) energy."""
        out = (ctypes.c_double * 2)()
        self._chk(self._lib.pbpic_fieldsenergy(self._h, out))
        return out[0], out[1]

    def kineticenergy(self):
        """totene(1:2) of `pusher` (pic2d.f90:1061-1062): (electron, ion)
        kinetic energy of the current momenta."""
        out = (ctypes.c_double * 2)()
        self._chk(self._lib.pbpic_kineticenergy(self._h, out))
        return out[0], out[1]

    def countingsort(self, species=None, want=False):
        """`countingsort` (pic2d.f90:1303-1329).  With a species and want=True
        returns (counts, perm): per-cell counts indexed by grid-1 and the
        0-based pre-sort index of each particle after the sort; want="perm"
        returns the permutation alone (a sort right after a push then takes
        the tile-box path, which computes no per-cell counts)."""
        if species is None:
            self._chk(self._lib.pbpic_countingsort(self._h))
            return None
        if not want:
            self._chk(self._lib.pbpic_countingsort_species(self._h, species, None, None))
            return None
        if want == "perm":
            perm = np.empty(max(self.num_particles(species), 1), np.int32)
            self._chk(self._lib.pbpic_countingsort_species(self._h, species, None,
                                                           _ptr(perm, ctypes.c_int32)))
            return perm[: self.num_particles(species)]
        counts = np.empty(self.nx * self.ny, np.int32)
        perm = np.empty(max(self.num_particles(species), 1), np.int32)
        self._chk(self._lib.pbpic_countingsort_species(
            self._h, species, _ptr(counts, ctypes.c_int32), _ptr(perm, ctypes.c_int32)))
        return counts, perm[: self.num_particles(species)]

    def step(self, nsteps=1):
        """The time-loop body (pic2d.f90:171-188), nsteps times."""
        self._chk(self._lib.pbpic_step(self._h, int(nsteps)))

    def energy_spectrum(self, species, nbins=1000, de=None, emax=None):
        """`energyspectra` (pic2d.f90:1957-1990) of one species: counts per
        energy bin k of ceil(ene/de) = k+1, ene = (gamma-1) ms < emax.  The
        reference's bins (:136-140) are the defaults: de = 0.005, emax = 5 for
        electrons, de = 0.025, emax = 25 for ions."""
        ion = self.species_q[species] > 0
        de = (0.025 if ion else 0.005) if de is None else de
        emax = (25.0 if ion else 5.0) if emax is None else emax
        out = np.empty(nbins, np.int32)
        self._chk(self._lib.pbpic_energy_spectrum(self._h, species, nbins, float(de), float(emax),
                                                  _ptr(out, ctypes.c_int32)))
        return out

    def number_density(self):
        """`numdensity` (pic2d.f90:1343-1354): particles per cell, all
        species, as a (ny, nx) array indexed [iy-1, ix-1]."""
        out = np.empty(self.nx * self.ny, np.int32)
        self._chk(self._lib.pbpic_number_density(self._h, _ptr(out, ctypes.c_int32)))
        return out.reshape(self.nx, self.ny).T.copy()

    def counters(self):
        """(steps, (je_step_raw, je_int)) for checkpoint/restart."""
        steps = ctypes.c_int64()
        je = (ctypes.c_double * 2)()
        self._chk(self._lib.pbpic_get_counters(self._h, ctypes.byref(steps), je))
        return steps.value, (je[0], je[1])

    def set_counters(self, steps, je):
        arr = (ctypes.c_double * 2)(*je)
        self._chk(self._lib.pbpic_set_counters(self._h, int(steps), arr))

    def diag(self):
        out = (ctypes.c_double * 16)()
        self._chk(self._lib.pbpic_diag(self._h, out))
        return {k: out[i] for i, k in enumerate(DIAG_NAMES)}

    # ------------------------------------------------------- device views
    def species_device_ptrs(self, species):
        arr = (ctypes.c_void_p * 7)()
        self._chk(self._lib.pbpic_species_device_ptrs(self._h, species, arr))
        return [arr[i] for i in range(7)]

    def field_device_ptrs(self):
        arr = (ctypes.c_void_p * 10)()
        self._chk(self._lib.pbpic_field_device_ptrs(self._h, arr))
        return [arr[i] for i in range(10)]

    def stream(self):
        return self._lib.pbpic_stream(self._h)

    # ------------------------------------------- device-side state, adjoint
    def set_field_device(self, which, ptr):
        """Copies one field plane from device memory (nx*ny doubles)."""
        w = FIELDS[which] if isinstance(which, str) else int(which)
        self._chk(self._lib.pbpic_set_field_device(self._h, w, ptr))

    def set_particles_device(self, species, n, cell, dx, dy, ux, uy, uz):
        """Copies a species' state from device memory (cell = grid - 1)."""
        self._chk(self._lib.pbpic_set_particles_device(self._h, species, int(n), cell, dx, dy,
                                                       ux, uy, uz))

    def adjoint_forward(self, nsteps, ckpt_every):
        """Runs nsteps steps keeping checkpoints every ckpt_every steps and
        returns the time-integrated J.E (the differentiable loss)."""
        out = ctypes.c_double()
        self._chk(self._lib.pbpic_adjoint_forward(self._h, int(nsteps), int(ckpt_every),
                                                  ctypes.byref(out)))
        return out.value

    def set_loss_region(self, weights=None):
        """Weights (ny, nx) of each cell's J.E term in the adjoint's loss
        (None: the whole domain)."""
        if weights is None:
            self._chk(self._lib.pbpic_set_loss_region(self._h, None))
            return
        a = _c(weights, np.float64)
        if a.shape != (self.ny, self.nx):
            raise ValueError(f"weights must have shape (ny, nx)={(self.ny, self.nx)}")
        self._chk(self._lib.pbpic_set_loss_region(self._h, _ptr(a, ctypes.c_double)))

    def adjoint_backward(self, field_grad_ptrs, particle_grad_ptrs):
        """Writes d(loss)/d(initial state) into caller device buffers: 6 field
        planes and 5 arrays per species (dx, dy, ux, uy, uz)."""
        fg = (ctypes.c_void_p * 6)(*field_grad_ptrs)
        pg = (ctypes.c_void_p * len(particle_grad_ptrs))(*particle_grad_ptrs)
        self._chk(self._lib.pbpic_adjoint_backward(self._h, fg, pg))

    # ------------------------------------------------------------ timing
    PROF_SLOTS = ("push", "field", "sort", "zero", "adjoint")

    def profile(self, enable=True):
        """Starts (clearing earlier samples) or stops HIP-event timing of each
        push / field / sort / zero launch on the context's stream."""
        self._chk(self._lib.pbpic_profile(self._h, int(bool(enable))))

    def profile_read(self):
        ms = (ctypes.c_double * 5)()
        cnt = (ctypes.c_int64 * 5)()
        self._chk(self._lib.pbpic_profile_read(self._h, ms, cnt))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.PROF_SLOTS)}
