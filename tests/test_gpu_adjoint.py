"""GPU adjoint: the hand-written adjoint kernels (k_adjoint.hip) through the C
ABI and the PyTorch binding, against torch's reverse mode applied to the
differentiable restatement oracle/torch_ref.py (itself pinned to the C
oracle and to central finite differences in tests/test_torch_ref.py)."""
import numpy as np
import pytest
import torch

from cases import make_case, rel_norm
from oracle import torch_ref as tr

pytestmark = pytest.mark.gpu

PART = ("dx", "dy", "ux", "uy", "uz")


def _case():
    return make_case(nx=24, ny=16, ppc=6, seed=4, efield=0.0, bfield=0.0, vth=(0.02, 0.1))


def _ref_grads(case, A0, ud0, nsteps, ckpt_sort, compat, mask=None):
    nx, ny = case["nx"], case["ny"]
    P = dict(nx=nx, ny=ny, dx=case["dx"], dy=case["dy"], dt=case["dt"], cva=case["cva"],
             mime=case["mime"], in0=case["in0"])
    if case.get("ybound"):
        P["nyp"] = ny - 2
    x = torch.arange(nx, dtype=torch.float64)[None, :].expand(ny, nx)
    F = torch.zeros(6, ny, nx, dtype=torch.float64)
    F[1] = A0 * torch.sin(2 * np.pi * x / nx)
    F.requires_grad_(True)
    FE, FB = F[:3], F[3:]
    if P.get("nyp"):
        FE, FB = tr.wall_images(FE, FB, P["nyp"])   # as set_field does
    sps = []
    for q, s in zip((1, -1), case["species"]):
        d = dict(q=q, cell=torch.tensor(s["grid"] - 1, dtype=torch.long))
        for k in PART:
            v = torch.tensor(s[k])
            if q < 0 and k == "ux":
                v = v + ud0
            d[k] = v.clone().requires_grad_(True)
        sps.append(d)
    loss, *_ = tr.run(P, FE, FB, sps, nsteps, compat=compat, sort_every=ckpt_sort,
                      loss_mask=None if mask is None else torch.tensor(mask))
    loss.backward()
    return loss.item(), F.grad.numpy(), [{k: sp[k].grad.numpy() for k in PART} for sp in sps]


def _gpu_grads(case, A0, ud0, nsteps, ckpt, sort_every, compat, precision, mask=None):
    import pbpic_amd
    from pbpic_amd.autograd import jdote
    nx, ny = case["nx"], case["ny"]
    sim = pbpic_amd.Context(nx, ny, case["dx"], case["dy"], case["dt"], case["cva"], case["mime"],
                            case["in0"], precision=precision, compat=compat, sort_every=sort_every,
                            deposit_rho=False, ybound=case.get("ybound", 0))
    if mask is not None:
        sim.set_loss_region(mask)
    dt = torch.float64 if precision == 0 else torch.float32
    species = []
    for q, s in zip((1, -1), case["species"]):
        sim.add_species(q, len(s["grid"]))
        d = {"cell": torch.tensor(s["grid"] - 1, dtype=torch.int32, device="cuda")}
        for k in PART:
            v = torch.tensor(s[k], dtype=torch.float64)
            if q < 0 and k == "ux":
                v = v + ud0
            d[k] = v.to(dt).cuda().requires_grad_(True)
        species.append(d)
    x = torch.arange(nx, dtype=torch.float64, device="cuda")[None, :].expand(ny, nx)
    F = torch.zeros(6, ny, nx, dtype=torch.float64, device="cuda")
    F[1] = A0 * torch.sin(2 * np.pi * x / nx)
    F.requires_grad_(True)
    loss = jdote(sim, nsteps, ckpt, F, species)
    loss.backward()
    out = (loss.item(), F.grad.cpu().numpy(),
           [{k: sp[k].grad.double().cpu().numpy() for k in PART} for sp in species])
    sim.close()
    return out


@pytest.mark.parametrize("compat", [7, 0])
def test_adjoint_fp64_matches_reverse_mode(gpu, compat):
    """All gradients (6 initial field planes, 5 arrays per species) of a 10-step
    run with a checkpoint every 4 steps and a sort every 5 steps."""
    print(f"adjoint fp64 compat={compat}", flush=True)
    case = _case()
    L, gF, gP = _ref_grads(case, 0.01, 0.05, 10, 5, compat)
    Lg, gFg, gPg = _gpu_grads(case, 0.01, 0.05, 10, 4, 5, compat, 0)
    assert abs(Lg - L) <= 1e-10 * abs(L)
    for k in range(6):
        assert rel_norm(gFg[k], gF[k]) < 1e-8, k
    for s in range(2):
        for k in PART:
            assert rel_norm(gPg[s][k], gP[s][k]) < 1e-8, (s, k)


def test_adjoint_split32_near_fp64(gpu):
    """The split-precision adjoint against the FP64 reference gradient."""
    case = _case()
    L, gF, gP = _ref_grads(case, 0.01, 0.05, 10, 5, 7)
    Lg, gFg, gPg = _gpu_grads(case, 0.01, 0.05, 10, 4, 5, 7, 1)
    assert abs(Lg - L) <= 1e-4 * abs(L)
    for k in range(6):
        assert rel_norm(gFg[k], gF[k]) < 1e-3, k
    for s in range(2):
        for k in PART:
            assert rel_norm(gPg[s][k], gP[s][k]) < 1e-3, (s, k)


@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_walls_matches_reverse_mode(gpu, precision):
    """Conducting y walls with a plasma hot enough that particles reflect:
    the adjoint through the reflection, the Jz fold and the guard-row imaging
    against torch's reverse mode (field gradients with respect to the domain
    rows, as set_field images the guards)."""
    print(f"adjoint walls precision={precision}", flush=True)
    case = make_case(nx=24, ny=16, ppc=6, seed=4, efield=0.0, bfield=0.0, vth=(0.05, 0.3),
                     walls=True)
    L, gF, gP = _ref_grads(case, 0.01, 0.05, 10, 5, 7)
    Lg, gFg, gPg = _gpu_grads(case, 0.01, 0.05, 10, 4, 5, 7, precision)
    tol_l, tol = (1e-10, 1e-8) if precision == 0 else (1e-4, 1e-3)
    assert abs(Lg - L) <= tol_l * abs(L)
    for k in range(6):
        assert rel_norm(gFg[k], gF[k]) < tol, k
        # guard rows carry no gradient: rows 14, 15 of the half-row planes
        # (Ey, Ez, By, Bz); row 15 of Ex, Bx, whose row 14 is the top wall
        assert not gFg[k][14 if k in (1, 2, 4, 5) else 15:].any(), k
        if k == 0:
            assert not gFg[0][0].any() and not gFg[0][14].any()   # Ex = 0 on the walls
    for s in range(2):
        for k in PART:
            assert rel_norm(gPg[s][k], gP[s][k]) < tol, (s, k)


@pytest.mark.parametrize("walls", [False, True])
def test_adjoint_loss_region(gpu, walls):
    """The loss restricted to a region of cells (weights 1 inside a central
    box, 0.5 on its border, 0 elsewhere), the way the large Harris
    configuration restricts J.E to the diffusion region."""
    print(f"adjoint loss region walls={walls}", flush=True)
    case = make_case(nx=24, ny=16, ppc=6, seed=4, efield=0.0, bfield=0.0,
                     vth=(0.05, 0.3) if walls else (0.02, 0.1), walls=walls)
    mask = np.zeros((16, 24))
    mask[5:11, 8:16] = 0.5
    mask[6:10, 9:15] = 1.0
    L, gF, gP = _ref_grads(case, 0.01, 0.05, 10, 5, 7, mask)
    Lg, gFg, gPg = _gpu_grads(case, 0.01, 0.05, 10, 4, 5, 7, 0, mask)
    assert abs(Lg - L) <= 1e-10 * abs(L)
    for k in range(6):
        assert rel_norm(gFg[k], gF[k]) < 1e-8, k
    for s in range(2):
        for k in PART:
            assert rel_norm(gPg[s][k], gP[s][k]) < 1e-8, (s, k)


@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_full_tiles(gpu, precision):
    """131,072 particles per species (32 full 4096-particle tiles): after the
    sort at step 5 the particles are in push order, so the adjoint takes its
    push-order rank map and the replayed push its fast path, writing each
    step into the next tape slot.  10 steps, checkpoints every 4, against
    torch's reverse mode."""
    print(f"adjoint full tiles precision={precision}", flush=True)
    case = make_case(nx=64, ny=64, ppc=32, seed=11, efield=0.0, bfield=0.0, vth=(0.02, 0.1))
    L, gF, gP = _ref_grads(case, 0.01, 0.05, 10, 5, 7)
    Lg, gFg, gPg = _gpu_grads(case, 0.01, 0.05, 10, 4, 5, 7, precision)
    tol_l, tol = (1e-10, 1e-8) if precision == 0 else (1e-4, 1e-3)
    assert abs(Lg - L) <= tol_l * abs(L)
    for k in range(6):
        assert rel_norm(gFg[k], gF[k]) < tol, k
    for s in range(2):
        for k in PART:
            assert rel_norm(gPg[s][k], gP[s][k]) < tol, (s, k)


@pytest.mark.parametrize("nx,ny", [(28, 24), (40, 36)])
@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_wide_windows(gpu, precision, nx, ny):
    """Unsorted particles: every adjoint tile spans the whole domain until the
    sort at step 5.  At 28 x 24 its node window (29 x 25 nodes) fits the first
    launch (1,024 nodes split, 768 FP64) but the current-adjoint window grown
    by one node on each side (31 x 27) does not fit the LDS left beside it
    (677 and 338 nodes), so the deposit's transpose reads the current adjoint
    from L2.  At 40 x 36 the box (41 x 37 nodes) goes to the second launch
    (2,048 and 1,536 nodes), whose current adjoint is read from L2 as well.
    Against torch's reverse mode."""
    print(f"adjoint wide windows precision={precision} {nx}x{ny}", flush=True)
    case = make_case(nx=nx, ny=ny, ppc=6, seed=4, efield=0.0, bfield=0.0, vth=(0.02, 0.1),
                     shuffle=True)
    L, gF, gP = _ref_grads(case, 0.01, 0.05, 10, 5, 7)
    Lg, gFg, gPg = _gpu_grads(case, 0.01, 0.05, 10, 4, 5, 7, precision)
    tol_l, tol = (1e-10, 1e-8) if precision == 0 else (1e-4, 1e-3)
    assert abs(Lg - L) <= tol_l * abs(L)
    for k in range(6):
        assert rel_norm(gFg[k], gF[k]) < tol, k
    for s in range(2):
        for k in PART:
            assert rel_norm(gPg[s][k], gP[s][k]) < tol, (s, k)
