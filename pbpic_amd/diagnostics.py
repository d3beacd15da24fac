"""Diagnostics and restart files in the layouts of pic2d.f90's writers.

  fieldsinfo(1)    :1626-1800  fields%05d.h5  ex ey ez bx by bz jx jy jz rho dive
                                (float32) and num (int32), Fortran (nx+2, ny+1),
                                i.e. C shape (ny+1, nx+2): columns 0..nx+1 are
                                the reference's i = 0..mx+1 with its periodic
                                ghosts, rows j = 1..my+1
  fieldsinfo(0)    :1828-1928  restart: E and B from such a file
  particlesinfo    :1359-1621  particles.h5   dx dy ux uy uz (float32), grid q
                                pid (int32), nptlact (int32, one entry per rank)
  energyinfo       :1996-2115  spectra.h5     /%05d/spectra int32 (nbins, 2)
                                [bin, electron|ion], ebins float32 (nbins, 2);
                                energy_tot.h5  tote float32 (ntime+1, 4) rows
                                itime: KE_e, KE_i, E-energy, B-energy

Where the reference's own output is an artefact this writes the intended
value, and says so: energy_tot.h5 keeps the last time row the reference drops
(its dims are (4, ntime) for totene2(4, 0:ntime), :2018, :2042); spectra are
indexed [bin, species] (the reference allocates espectra(2, nbins) and indexes
it (ibin, j), :1987, :2040); the ghost row j = my+1 of J and rho holds the
periodic copy of row 1 (the reference leaves the pre-exchange partial sums
there); dive uses the y range 0:my (:795 writes 0:mx).  num keeps zero ghosts
and dive a zero column 0, as the reference's arrays do.

save_checkpoint / load_checkpoint write and restore the complete state at its
own precision (FP64 fields, particles in the context's precision, step count
and J.E accumulators), for an exact restart; the float32 reference files are
for interchange.
"""
import os

import numpy as np

from . import h5

FIELD_NAMES = ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz", "rho")
# energyspectra bins (pic2d.f90:133-140): minene = 0, maxenee = 5,
# maxenei = 5 tiecs, dee = 0.005, dei = 0.005 tiecs, nbins = 1000
NBINS, DEE, DEI = 1000, 0.005, 0.025


# ------------------------------------------------------------ pure writers
def ghosted(plane):
    """(ny, nx) compact plane -> (ny+1, nx+2): column c holds x index c-1
    (periodic), row r holds y index r (row ny = row 0)."""
    ny, nx = plane.shape
    cols = (np.arange(nx + 2) - 1) % nx
    rows = np.arange(ny + 1) % ny
    return plane[np.ix_(rows, cols)]


def divergence(ex, ey, dx, dy):
    """dive (pic2d.f90:794-795) at the nodes, ghosted: column 0 stays 0."""
    d = (ex - np.roll(ex, 1, axis=1)) / dx + (ey - np.roll(ey, 1, axis=0)) / dy
    g = ghosted(d)
    g[:, 0] = 0.0
    return g


def write_fields(path, planes, num, dx, dy):
    """fields%05d.h5 from compact (ny, nx) planes (FIELD_NAMES) and per-cell
    counts num (ny, nx)."""
    ny, nx = planes["ex"].shape
    with h5.File(path, "w") as f:
        for k in FIELD_NAMES:
            f.write(k, ghosted(planes[k]).astype(np.float32))
        g = np.zeros((ny + 1, nx + 2), np.int32)
        g[:ny, 1:nx + 1] = num
        f.write("num", g)
        f.write("dive", divergence(planes["ex"], planes["ey"], dx, dy).astype(np.float32))


def read_fields(path):
    """The compact (ny, nx) planes of a fields%05d.h5 file (float64)."""
    out = {}
    with h5.File(path, "r") as f:
        for k in FIELD_NAMES:
            a = f.read(k)
            out[k] = a[:-1, 1:-1].astype(np.float64)
    return out


def write_particles(path, parts):
    """particles.h5 from a dict of 1-D arrays grid, dx, dy, ux, uy, uz, q, pid."""
    n = len(parts["grid"])
    with h5.File(path, "w") as f:
        for k in ("dx", "dy", "ux", "uy", "uz"):
            f.write(k, np.asarray(parts[k], np.float32))
        for k in ("grid", "q", "pid"):
            f.write(k, np.asarray(parts[k], np.int32))
        f.write("nptlact", np.array([n], np.int32))


def read_particles(path):
    with h5.File(path, "r") as f:
        n = int(f.read("nptlact").sum())
        return {k: f.read(k)[:n] for k in ("dx", "dy", "ux", "uy", "uz", "grid", "q", "pid")}


def ebins(nbins=NBINS, dee=DEE, dei=DEI):
    """enebins (pic2d.f90:2033-2038): upper bin edges (nbins, 2)."""
    i = np.arange(1, nbins + 1, dtype=np.float64)
    return np.stack([i * dee, i * dei], axis=1).astype(np.float32)


def write_spectra(path, itime, spec_e, spec_i, nbins=NBINS):
    """Appends group %05d with the (nbins, 2) spectra; adds ebins on creation."""
    new = not os.path.exists(path)
    with h5.File(path, "a") as f:
        f.write(f"{itime:05d}/spectra", np.stack([spec_e, spec_i], axis=1).astype(np.int32))
        if new:
            f.write("ebins", ebins(nbins))


def write_energy(path, totene):
    """energy_tot.h5 with tote (ntime+1, 4) float32."""
    with h5.File(path, "a") as f:
        f.write("tote", np.asarray(totene, np.float32))


# ------------------------------------------------------ context-level I/O
class Diagnostics:
    """The reference's diagnostic cadence for one Context (pic2d.f90:158-188):
    call record(itime) every step (totene), fieldsinfo / energyinfo /
    particlesinfo at the output steps.  Files go to `path`."""

    def __init__(self, ctx, ntime, path=".", species_e=None, species_i=None):
        self.ctx = ctx
        self.path = path
        os.makedirs(path, exist_ok=True)
        self.totene = np.zeros((ntime + 1, 4), np.float64)
        qs = ctx.species_q
        self.se = species_e if species_e is not None else next((k for k, q in enumerate(qs) if q < 0), None)
        self.si = species_i if species_i is not None else next((k for k, q in enumerate(qs) if q > 0), None)

    def record(self, itime):
        """totene(1:4, itime): kinetic energies of the current momenta
        (:1061-1062) and field energies (:653-673)."""
        ke_e, ke_i = self.ctx.kineticenergy()
        we, wb = self.ctx.fieldsenergy()
        self.totene[itime] = (ke_e, ke_i, we, wb)

    def fieldsinfo(self, itime):
        c = self.ctx
        planes = {k: c.get_field(k) for k in FIELD_NAMES}
        p = c.params
        write_fields(os.path.join(self.path, f"fields{itime:05d}.h5"), planes,
                     c.number_density(), p.dx, p.dy)

    def energyinfo(self, itime):
        c = self.ctx
        z = np.zeros(NBINS, np.int32)
        se = c.energy_spectrum(self.se, NBINS, DEE, NBINS * DEE) if self.se is not None else z
        si = c.energy_spectrum(self.si, NBINS, DEI, NBINS * DEI) if self.si is not None else z
        write_spectra(os.path.join(self.path, "spectra.h5"), itime, se, si)
        write_energy(os.path.join(self.path, "energy_tot.h5"), self.totene)

    def particlesinfo(self):
        write_particles(os.path.join(self.path, "particles.h5"), gather_particles(self.ctx))


def gather_particles(ctx):
    """All species in species order with their charge; pid from the tracked
    ids, else 1..n per species."""
    out = {k: [] for k in ("grid", "dx", "dy", "ux", "uy", "uz", "q", "pid")}
    for s, q in enumerate(ctx.species_q):
        p = ctx.get_particles(s)
        n = len(p["grid"])
        for k in ("grid", "dx", "dy", "ux", "uy", "uz"):
            out[k].append(p[k])
        out["q"].append(np.full(n, q, np.int32))
        out["pid"].append(p["pid"] if ctx.tracks_pid else np.arange(1, n + 1, dtype=np.int32))
    return {k: np.concatenate(v) if v else np.zeros(0) for k, v in out.items()}


def restart_from_reference_files(ctx, fields_path, particles_path):
    """fieldsinfo(0) + particlesinfo(0) (:1828-1928, :1572-1597): E and B from
    a fields%05d.h5, particles from particles.h5 split by charge into the
    context's species (one per charge)."""
    planes = read_fields(fields_path)
    for k in ("ex", "ey", "ez", "bx", "by", "bz"):
        ctx.set_field(k, planes[k])
    p = read_particles(particles_path)
    for s, q in enumerate(ctx.species_q):
        m = p["q"] == q
        ctx.set_particles(s, p["grid"][m], p["dx"][m], p["dy"][m], p["ux"][m], p["uy"][m],
                          p["uz"][m], p["pid"][m])


def save_checkpoint(ctx, path):
    """Exact state: FP64 fields, particles at the context's precision, the
    step count and the J.E accumulators."""
    steps, je = ctx.counters()
    p = ctx.params
    with h5.File(path, "w") as f:
        f.write("meta/geometry", np.array([p.nx, p.ny, p.precision, p.ybound], np.int64))
        f.write("meta/steps", np.array([steps], np.int64))
        f.write("meta/je", np.array(je, np.float64))
        f.write("meta/q", np.array(ctx.species_q, np.int32))
        for k in ("ex", "ey", "ez", "bx", "by", "bz"):
            f.write(f"fields/{k}", ctx.get_field(k))
        for s in range(len(ctx.species_q)):
            g = ctx.get_particles(s, native=True)
            for k in ("grid", "dx", "dy", "ux", "uy", "uz"):
                f.write(f"species{s}/{k}", g[k])
            if ctx.tracks_pid:
                f.write(f"species{s}/pid", g["pid"])


def load_checkpoint(ctx, path):
    """Restores a save_checkpoint file into a context of the same geometry,
    precision and species charges."""
    p = ctx.params
    with h5.File(path, "r") as f:
        geo = f.read("meta/geometry")
        if tuple(int(x) for x in geo) != (p.nx, p.ny, p.precision, p.ybound):
            raise ValueError(f"checkpoint geometry {tuple(geo)} does not match the context")
        q = [int(x) for x in f.read("meta/q")]
        if q != list(ctx.species_q):
            raise ValueError(f"checkpoint species {q} do not match {ctx.species_q}")
        for k in ("ex", "ey", "ez", "bx", "by", "bz"):
            ctx.set_field(k, f.read(f"fields/{k}"))
        for s in range(len(q)):
            arr = {k: f.read(f"species{s}/{k}") for k in ("grid", "dx", "dy", "ux", "uy", "uz")}
            pid = f.read(f"species{s}/pid") if f"species{s}/pid" in f else None
            ctx.set_particles(s, arr["grid"], arr["dx"], arr["dy"], arr["ux"], arr["uy"],
                              arr["uz"], pid)
        ctx.set_counters(int(f.read("meta/steps")[0]), tuple(f.read("meta/je")))
