"""Pins the differentiable torch restatement (oracle/torch_ref.py): its forward
equals the C oracle's, and torch's reverse-mode gradient of the integrated
J.E equals central finite differences of the C oracle.  The GPU adjoint is
then checked against this gradient (tests/test_gpu_adjoint.py)."""
import numpy as np
import pytest
import torch

import oracle
from cases import make_case, oracle_for, oracle_species
from oracle import torch_ref as tr


def torch_inputs(case):
    P = dict(nx=case["nx"], ny=case["ny"], dx=case["dx"], dy=case["dy"], dt=case["dt"],
             cva=case["cva"], mime=case["mime"], in0=case["in0"])
    if case.get("ybound"):
        P["nyp"] = case["ny"] - 2
        # the fields as the oracle holds them: guard rows imaged
        o = oracle_for(dict(case, species=[]), compat=0)
        case = dict(case, fields={k: o.get_field(k) for k in case["fields"]})
        o.close()
    E = torch.tensor(np.stack([case["fields"][k] for k in ("ex", "ey", "ez")]))
    B = torch.tensor(np.stack([case["fields"][k] for k in ("bx", "by", "bz")]))
    sps = []
    for q, s in zip((1, -1), case["species"]):
        sps.append(dict(q=q, cell=torch.tensor(s["grid"] - 1, dtype=torch.long),
                        **{k: torch.tensor(s[k]) for k in ("dx", "dy", "ux", "uy", "uz")}))
    return P, E, B, sps


@pytest.mark.parametrize("compat,walls", [(7, False), (0, False), (0, True), (7, True)])
def test_forward_matches_oracle(oracle_lib, compat, walls):
    case = make_case(nx=16, ny=12, ppc=4, seed=3, vth=(0.1, 0.45) if walls else (0.05, 0.3),
                     walls=walls)
    P, E, B, sps = torch_inputs(case)
    loss, E1, B1, sp1 = tr.run(P, E, B, sps, 12, compat=compat, sort_every=5)
    o = oracle_for(case, compat=compat)
    o.step(12, sort_every=5)
    assert abs(loss.item() - o.diag()["je_int"]) <= 1e-12 * abs(loss.item())
    osp = oracle_species(o, case)
    for k in range(2):
        assert np.array_equal(sp1[k]["cell"].numpy() + 1, osp[k]["grid"])
        assert np.allclose(sp1[k]["ux"].numpy(), osp[k]["ux"], rtol=1e-11, atol=1e-14)
    assert np.allclose(E1[1].numpy(), o.get_field("ey"), rtol=1e-11, atol=1e-12)


def _c3_like(nx=24, ny=16, ppc=6, seed=4, walls=False):
    """The differentiable-PIC configuration in miniature: thermal plasma, an
    initial Ey = A sin(2 pi x/Lx), an electron drift u_d along x and a per-cell
    Bz perturbation dBz, all zero except A.  walls: conducting y walls and a
    hotter plasma, so that particles reflect during the run."""
    case = make_case(nx=nx, ny=ny, ppc=ppc, seed=seed, efield=0.0, bfield=0.0,
                     vth=(0.05, 0.3) if walls else (0.02, 0.1), walls=walls)
    return case


def _oracle_loss(case, A, ud, dbz, nsteps, compat):
    c = dict(case)
    c["fields"] = dict(case["fields"])
    x = np.arange(case["nx"])[None, :] * np.ones((case["ny"], 1))
    c["fields"]["ey"] = A * np.sin(2 * np.pi * x / case["nx"])
    c["fields"]["bz"] = dbz
    sp = [dict(s) for s in case["species"]]
    sp[1]["ux"] = sp[1]["ux"] + ud
    c["species"] = sp
    o = oracle_for(c, compat=compat)
    o.step(nsteps, sort_every=5)
    return o.diag()["je_int"]


@pytest.mark.parametrize("compat,walls", [(7, False), (0, False), (7, True)])
def test_gradient_matches_finite_differences(oracle_lib, compat, walls):
    case = _c3_like(walls=walls)
    nx, ny, nsteps = case["nx"], case["ny"], 10
    A0, ud0 = 0.01, 0.05
    P, E, B, sps = torch_inputs(case)
    A = torch.tensor(A0, dtype=torch.float64, requires_grad=True)
    ud = torch.tensor(ud0, dtype=torch.float64, requires_grad=True)
    dbz = torch.zeros((ny, nx), dtype=torch.float64, requires_grad=True)
    x = torch.arange(nx, dtype=torch.float64)[None, :].expand(ny, nx)
    E = torch.stack([E[0], A * torch.sin(2 * np.pi * x / nx), E[2]])
    B = torch.stack([B[0], B[1], dbz])
    if walls:
        # set_field images the guard rows: the free variables are the domain rows
        E, B = tr.wall_images(E, B, P["nyp"])
    sps[1] = dict(sps[1], ux=sps[1]["ux"] + ud)
    loss, *_ = tr.run(P, E, B, sps, nsteps, compat=compat, sort_every=5)
    loss.backward()
    z = np.zeros((ny, nx))
    L0 = _oracle_loss(case, A0, ud0, z, nsteps, compat)
    assert abs(loss.item() - L0) <= 1e-12 * abs(L0)
    eps = 1e-6
    fd_A = (_oracle_loss(case, A0 + eps, ud0, z, nsteps, compat) -
            _oracle_loss(case, A0 - eps, ud0, z, nsteps, compat)) / (2 * eps)
    fd_u = (_oracle_loss(case, A0, ud0 + eps, z, nsteps, compat) -
            _oracle_loss(case, A0, ud0 - eps, z, nsteps, compat)) / (2 * eps)
    assert abs(A.grad.item() - fd_A) <= 1e-4 * abs(fd_A), (A.grad.item(), fd_A)
    assert abs(ud.grad.item() - fd_u) <= 1e-4 * abs(fd_u), (ud.grad.item(), fd_u)
    for (j, i) in ((3, 5), (10, 17)) + (((0, 7), (13, 2)) if walls else ()):
        e = np.zeros((ny, nx))
        e[j, i] = eps
        fd = (_oracle_loss(case, A0, ud0, e, nsteps, compat) -
              _oracle_loss(case, A0, ud0, -e, nsteps, compat)) / (2 * eps)
        assert abs(dbz.grad[j, i].item() - fd) <= 1e-4 * abs(fd) + 1e-9, (j, i, dbz.grad[j, i].item(), fd)
