"""Differentiable restatement of the PIC step in PyTorch (TEST INFRASTRUCTURE).

A second CPU restatement of pic2d.f90's step, written with torch tensor ops
in float64 so that torch's reverse mode yields a reference gradient of the
time-integrated J.E; the hand-written GPU adjoint is checked against it.  Its
forward is pinned to the C oracle (tests/test_torch_ref.py), its gradient to
central finite differences.  Only tests use it.

Layout: compact periodic planes (ny, nx), x fastest; particles per species
with 0-based cell = cx*ny + cy and offsets in (-1/2, 1/2].
  gather + Boris + move      pic2d.f90:938-1028
  deposit (relative form)    :1237-1282 (algebraically equal to the
                             absolute-index form of the source)
  x fold of jy column mx+2   :1156 (compat bit 4)
  fieldsolver                :686-803
With P["nyp"] set (conducting y walls, no counterpart in the reference): the
domain is rows 0..nyp-1 of the ny = nyp + 2 rows, particles are reflected at
the walls before their deposit, Jz beyond a wall folds back with sign -1,
Ex vanishes on the walls and the two guard rows hold the fields' mirror
images (the same construction as the C oracle's wall_image).
"""
import math

import torch

BORIS_OLD_U, RHO_MIRROR, JY_DROP = 1, 2, 4


def _rx(a, s):   # value at x - s (periodic), planes are (ny, nx)
    return torch.roll(a, s, dims=1)


def _ry(a, s):
    return torch.roll(a, s, dims=0)


def node_fields(E, B):
    """emfgrid (pic2d.f90:837-845) as 6 node planes."""
    ex, ey, ez = E
    bx, by, bz = B
    return torch.stack([(ex + _rx(ex, 1)) * 0.5, (ey + _ry(ey, 1)) * 0.5,
                        (ez + _rx(ez, 1) + _ry(ez, 1) + _rx(_ry(ez, 1), 1)) * 0.25,
                        (bx + _rx(bx, 1)) * 0.5, (by + _ry(by, 1)) * 0.5,
                        (bz + _rx(bz, 1) + _ry(bz, 1) + _rx(_ry(bz, 1), 1)) * 0.25])


def _cross(a, b):
    return torch.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                        a[0] * b[1] - a[1] * b[0]])


def push_deposit(P, sp, N, J, compat):
    """One species: gather, Boris, move, deposit into J (3 flat planes)."""
    nx, ny = P["nx"], P["ny"]
    cell = sp["cell"]
    cx, cy = cell // ny, cell % ny
    cx1, cy1 = (cx + 1) % nx, (cy + 1) % ny
    ox, oy = sp["dx"], sp["dy"]
    u = torch.stack([sp["ux"], sp["uy"], sp["uz"]])
    s00 = (0.5 - ox) * (0.5 - oy)
    s10 = (0.5 + ox) * (0.5 - oy)
    s01 = (0.5 - ox) * (0.5 + oy)
    s11 = (0.5 + ox) * (0.5 + oy)
    F = N[:, cy, cx] * s00 + N[:, cy, cx1] * s10 + N[:, cy1, cx] * s01 + N[:, cy1, cx1] * s11
    E, B = F[:3], F[3:]
    q = float(sp["q"])
    ms = P["mime"] if q > 0 else 1.0
    c0 = q * P["dt"] * P["mime"] / ms / 2.0
    c1 = c0 / P["cva"]
    um = u + c1 * E
    g1 = torch.sqrt(1.0 + (um * um).sum(0))
    t = (c0 / g1) * B
    s = 2.0 / (1.0 + (t * t).sum(0)) * t
    u0 = um + _cross(u if compat & BORIS_OLD_U else um, t)
    un = um + _cross(u0, s) + c1 * E
    gn = torch.sqrt(1.0 + (un * un).sum(0))
    c3 = P["dt"] * P["cva"] / gn
    dx1 = ox + c3 * un[0] / P["dx"]
    dy1 = oy + c3 * un[1] / P["dy"]
    a = torch.where(dx1.abs() > 0.5, torch.sign(dx1), torch.zeros_like(dx1)).detach()
    b = torch.where(dy1.abs() > 0.5, torch.sign(dy1), torch.zeros_like(dy1)).detach()
    ox2, oy2 = dx1 - a, dy1 - b
    nyp = P.get("nyp")
    if nyp:
        # specular reflection at a wall: the offset mirrored into the cell
        # the particle left, u_y reversed
        refl = ((cy == 0) & (b < 0)) | ((cy == nyp - 1) & (b > 0))
        b = torch.where(refl, torch.zeros_like(b), b)
        oy2 = torch.where(refl, -oy2, oy2)
        un = torch.stack([un[0], torch.where(refl, -un[1], un[1]), un[2]])
    vz = un[2] * P["cva"] / gn
    qx, qy = q * P["dx"] / P["dt"], q * P["dy"] / P["dt"]

    def axis(a_, o1, o2, qf):
        imr = (a_ + o1 + o2 + 1.0) * 0.5
        lo = torch.clamp(a_, min=0.0)
        hi = torch.clamp(a_, max=0.0) + 1.0
        irr = torch.minimum(hi, torch.maximum(lo, imr))
        return (qf * (irr - o1 - 0.5), qf * (a_ + o2 + 0.5 - irr), (irr + o1 + 0.5) * 0.5,
                (irr - a_ + o2 + 0.5) * 0.5, imr - 0.5)

    fx1, fx2, wx1, wx2, dim = axis(a, ox, ox2, qx)
    fy1, fy2, wy1, wy2, djm = axis(b, oy, oy2, qy)
    ai, bi = a.long(), b.long()

    def at(ddx, ddy):
        return ((cy + ddy) % ny) * nx + (cx + ddx) % nx

    jx, jy, jz = J
    jx.index_add_(0, at(0, 0), fx1 * (1 - wy1))
    jx.index_add_(0, at(0, 1), fx1 * wy1)
    jy.index_add_(0, at(0, 0), fy1 * (1 - wx1))
    jy.index_add_(0, at(1, 0), fy1 * wx1)
    jx.index_add_(0, at(ai, bi), fx2 * (1 - wy2))
    jx.index_add_(0, at(ai, bi + 1), fx2 * wy2)
    jy.index_add_(0, at(ai, bi), fy2 * (1 - wx2))
    keep = torch.ones_like(fy2)
    if compat & JY_DROP:
        keep = torch.where((cx == nx - 1) & (ai == 1), 0.0, 1.0).to(fy2.dtype)
    jy.index_add_(0, at(ai + 1, bi), fy2 * wx2 * keep)
    di = torch.where(torch.signbit(dim), -1, 1)
    dj = torch.where(torch.signbit(djm), -1, 1)
    A, Bq = dim.abs(), djm.abs()
    quz = q * vz
    jz.index_add_(0, at(0, 0), quz * (1 - A) * (1 - Bq))
    jz.index_add_(0, at(di, 0), quz * A * (1 - Bq))
    jz.index_add_(0, at(0, dj), quz * (1 - A) * Bq)
    jz.index_add_(0, at(di, dj), quz * A * Bq)
    new = dict(sp)
    new.update(cell=((cx + ai) % nx) * ny + (cy + bi) % ny, dx=ox2, dy=oy2,
               ux=un[0], uy=un[1], uz=un[2])
    return new


def _image_rows(a, nyp, sign):
    """Guard rows nyp, nyp+1 of a half-row plane <- sign x rows nyp-1, 0."""
    return torch.cat([a[:nyp], sign * a[nyp - 1:nyp], sign * a[0:1]], dim=0)


def wall_images(E, B, nyp):
    ex, ey, ez = E
    bx, by, bz = B
    z = torch.zeros_like(ex[0:1])
    ex = torch.cat([z, ex[1:nyp], z, -ex[1:2]], dim=0)
    bx = torch.cat([bx[:nyp + 1], bx[1:2]], dim=0)
    return (torch.stack([ex, _image_rows(ey, nyp, 1.0), _image_rows(ez, nyp, -1.0)]),
            torch.stack([bx, _image_rows(by, nyp, -1.0), _image_rows(bz, nyp, 1.0)]))


def fold_jz(jz, nyp):
    """Jz of the guard rows folded back with sign -1 (jz: (ny, nx) plane)."""
    top, bot = jz[nyp], jz[nyp + 1]
    rows = [jz[0] - bot] + [jz[r] for r in range(1, nyp - 1)] + [jz[nyp - 1] - top]
    if nyp == 1:
        rows = [jz[0] - bot - top]
    z = torch.zeros_like(jz[0])
    return torch.stack(rows + [z, z])


def fieldsolve(P, E, B, J):
    """fieldsolver (pic2d.f90:686-803) on compact periodic planes; J raw."""
    ex, ey, ez = E
    bx, by, bz = B
    jx, jy, jz = (j * P["in0"] for j in J)
    idx, idy = 1.0 / P["dx"], 1.0 / P["dy"]
    dth = P["dt"] / 2.0
    dtc2 = P["dt"] * (P["cva"] * P["cva"])
    bx = bx - dth * ((ez - _ry(ez, 1)) * idy)
    by = by + dth * ((ez - _rx(ez, 1)) * idx)
    bz = bz + dth * ((_ry(ex, -1) - ex) * idy - (_rx(ey, -1) - ey) * idx)
    exn = ex + dtc2 * ((bz - _ry(bz, 1)) * idy - jx)
    eyn = ey - dtc2 * ((bz - _rx(bz, 1)) * idx + jy)
    ezn = ez + dtc2 * ((_rx(by, -1) - by) * idx - (_ry(bx, -1) - bx) * idy - jz)
    nyp = P.get("nyp")
    if nyp:
        z = torch.zeros_like(exn[0:1])
        exn = torch.cat([z, exn[1:nyp], z, exn[nyp + 1:]], dim=0)   # tangential E on the walls
        ezn = _image_rows(ezn, nyp, -1.0)
    bx = bx - dth * ((ezn - _ry(ezn, 1)) * idy)
    by = by + dth * ((ezn - _rx(ezn, 1)) * idx)
    bz = bz + dth * ((_ry(exn, -1) - exn) * idy - (_rx(eyn, -1) - eyn) * idx)
    E1, B1 = torch.stack([exn, eyn, ezn]), torch.stack([bx, by, bz])
    if nyp:
        E1, B1 = wall_images(E1, B1, nyp)
    return E1, B1


def run(P, E, B, species, nsteps, compat=7, sort_every=0, step0=0, loss_mask=None):
    """nsteps steps; returns (loss = time-integrated J.E, final state).
    loss_mask: (ny, nx) weights of each cell's J.E term (None: all 1)."""
    nx, ny = P["nx"], P["ny"]
    w = P["dt"] * P["mime"] / (P["cva"] * P["cva"])
    loss = torch.zeros((), dtype=torch.float64, device=E.device)
    for n in range(nsteps):
        N = node_fields(E, B)
        J = [torch.zeros(nx * ny, dtype=torch.float64, device=E.device) for _ in range(3)]
        species = [push_deposit(P, sp, N, J, compat) for sp in species]
        Jp = torch.stack([j.reshape(ny, nx) for j in J])
        if P.get("nyp"):
            Jp = torch.stack([Jp[0], Jp[1], fold_jz(Jp[2], P["nyp"])])
        E1, B1 = fieldsolve(P, E, B, Jp)
        je = Jp * (E + E1) * 0.5
        if loss_mask is not None:
            je = je * loss_mask
        loss = loss + w * je.sum()
        E, B = E1, B1
        if sort_every and (step0 + n + 1) % sort_every == 0:
            new = []
            for sp in species:
                perm = torch.argsort(sp["cell"], stable=True)
                new.append({k: (v[perm] if torch.is_tensor(v) else v) for k, v in sp.items()})
            species = new
    return loss, E, B, species
