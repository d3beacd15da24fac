"""Seeded test cases shared by the oracle and GPU parity tests.

A case is a small periodic plasma in the reference's normalization
(pic2d.f90:15-43): two species (ions q=+1 first, electrons q=-1), optional
random fields.  The oracle holds both species in one interleaved array like
the reference's `ptl`; the GPU holds one array per species.
"""
import numpy as np

import oracle

SPECIES_Q = (1, -1)


def make_case(nx=24, ny=20, ppc=8, seed=7, vth=(0.05, 0.3), efield=0.3, bfield=1.5,
              shuffle=False, sort=True, dx=0.1, dy=0.1, cva=15.0, mime=25.0, walls=False):
    """walls: conducting y walls; ny counts the two guard rows, the plasma
    fills rows 0..ny-3."""
    rng = np.random.default_rng(seed)
    dt = 0.2 / (cva * np.sqrt(mime))
    in0 = 1.0 / ppc
    sp = []
    nyp = ny - 2 if walls else ny
    for k, q in enumerate(SPECIES_Q):
        p = oracle.init_maxwellian(nx, nyp, ppc, vth[k], seed=seed, stream=k)
        if walls:
            p["grid"] = oracle.wall_rows(p["grid"], nyp, ny)
        if shuffle:
            order = rng.permutation(len(p["grid"]))
            p = {key: v[order] for key, v in p.items()}
        p["pid"] = np.arange(1, len(p["grid"]) + 1, dtype=np.int32) + k * 10_000_000
        sp.append(p)
    fields = {}
    for name, amp in (("ex", efield), ("ey", efield), ("ez", efield),
                      ("bx", bfield), ("by", bfield), ("bz", bfield)):
        fields[name] = amp * rng.standard_normal((ny, nx)) if amp else np.zeros((ny, nx))
    return dict(nx=nx, ny=ny, dx=dx, dy=dy, dt=dt, cva=cva, mime=mime, in0=in0,
                species=sp, fields=fields, sort=sort, ybound=1 if walls else 0)


def oracle_for(case, prec=8, compat=oracle.COMPAT_REFERENCE):
    o = oracle.Oracle(case["nx"], case["ny"], case["dx"], case["dy"], case["dt"],
                      case["cva"], case["mime"], case["in0"], prec=prec, compat=compat,
                      ybound=case.get("ybound", 0))
    sp = case["species"]
    if sp:
        cat = lambda k: np.concatenate([s[k] for s in sp])
        q = np.concatenate([np.full(len(s["grid"]), qq, np.int32) for s, qq in zip(sp, SPECIES_Q)])
        o.set_particles(cat("grid"), cat("dx"), cat("dy"), cat("ux"), cat("uy"), cat("uz"), q,
                        cat("pid"))
    else:
        o.set_particles([], [], [], [], [], [], [])
    for name, plane in case["fields"].items():
        o.set_field(name, plane)
    return o


def oracle_species(o, case):
    """Splits the oracle's interleaved particle array back into species."""
    p = o.get_particles()
    out = []
    for q in SPECIES_Q:
        m = p["q"] == q
        out.append({k: v[m] for k, v in p.items()})
    return out


def gpu_for(case, precision=0, compat=oracle.COMPAT_REFERENCE, sort_every=0, deposit_rho=True,
            **kw):
    import pbpic_amd
    ctx = pbpic_amd.Context(case["nx"], case["ny"], case["dx"], case["dy"], case["dt"],
                            case["cva"], case["mime"], case["in0"], precision=precision,
                            compat=compat, sort_every=sort_every, deposit_rho=deposit_rho,
                            track_pid=True, ybound=case.get("ybound", 0), **kw)
    for q, s in zip(SPECIES_Q, case["species"]):
        k = ctx.add_species(q, len(s["grid"]))
        ctx.set_particles(k, s["grid"], s["dx"], s["dy"], s["ux"], s["uy"], s["uz"], s["pid"])
    for name, plane in case["fields"].items():
        ctx.set_field(name, plane)
    return ctx


def rel_norm(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def gilbert_order(w, h):
    """Cells (x, y) of a w x h grid in the order of the generalized Hilbert
    curve (J. Cerveny's "gilbert" construction): a rectangle more than 1.5
    times as long as it is wide is cut in two along its length, otherwise
    walked up / across / down, with cuts moved by one where a half would be
    odd."""
    out = []

    def sgn(v):
        return (v > 0) - (v < 0)

    def walk(x, y, ax, ay, bx, by):
        w_, h_ = abs(ax + ay), abs(bx + by)
        dax, day, dbx, dby = sgn(ax), sgn(ay), sgn(bx), sgn(by)
        if h_ == 1:
            for _ in range(w_):
                out.append((x, y))
                x, y = x + dax, y + day
            return
        if w_ == 1:
            for _ in range(h_):
                out.append((x, y))
                x, y = x + dbx, y + dby
            return
        ax2, ay2, bx2, by2 = ax // 2, ay // 2, bx // 2, by // 2
        w2, h2 = abs(ax2 + ay2), abs(bx2 + by2)
        if 2 * w_ > 3 * h_:
            if w2 % 2 and w_ > 2:
                ax2, ay2 = ax2 + dax, ay2 + day
            walk(x, y, ax2, ay2, bx, by)
            walk(x + ax2, y + ay2, ax - ax2, ay - ay2, bx, by)
        else:
            if h2 % 2 and h_ > 2:
                bx2, by2 = bx2 + dbx, by2 + dby
            walk(x, y, bx2, by2, ax2, ay2)
            walk(x + bx2, y + by2, ax, ay, bx - bx2, by - by2)
            walk(x + (ax - dax) + (bx2 - dbx), y + (ay - day) + (by2 - dby),
                 -bx2, -by2, -(ax - ax2), -(ay - ay2))

    if w >= h:
        walk(0, 0, w, 0, 0, h)
    else:
        walk(0, 0, 0, h, w, 0)
    return out


def supertile_key(cell, nx, ny):
    """Push-order key of a 0-based cell id: 8 x 8 supertiles ranked along the
    generalized Hilbert curve over the supertile grid, cells x-major inside
    a supertile."""
    cell = np.asarray(cell, dtype=np.int64)
    cx, cy = cell // ny, cell % ny
    nbx, nby = (nx + 7) // 8, (ny + 7) // 8
    rank = np.zeros(nbx * nby, dtype=np.int64)
    for r, (x, y) in enumerate(gilbert_order(nbx, nby)):
        rank[x * nby + y] = r
    return rank[(cx >> 3) * nby + (cy >> 3)] * 64 + (cx & 7) * 8 + (cy & 7)
