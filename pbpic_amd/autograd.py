"""PyTorch binding of the hand-written PIC adjoint.

``jdote(sim, nsteps, ckpt, fields, species)`` returns the time-integrated J.E
of ``nsteps`` steps of the PIC step on the GPU, as a differentiable scalar.
``fields`` is a (6, ny, nx) float64 CUDA tensor (Ex, Ey, Ez, Bx, By, Bz); each
entry of ``species`` is a dict with ``cell`` (int32, 0-based compact id =
grid - 1, not differentiable) and ``dx``, ``dy``, ``ux``, ``uy``, ``uz``
tensors in the context's particle precision.  The backward pass runs the
adjoint kernels of libpbpic (checkpoints every ``ckpt`` steps, recomputation
in between) and returns d(loss)/d(fields) and d(loss)/d(particle arrays);
autograd chains them to whatever produced the inputs (an initial-condition
amplitude, a drift, a per-cell field perturbation ...).
"""
import torch

PART = ("dx", "dy", "ux", "uy", "uz")


def _sync_torch():
    torch.cuda.current_stream().synchronize()


class _JdotE(torch.autograd.Function):
    @staticmethod
    def forward(fctx, sim, nsteps, ckpt, nspecies, fields, *flat):
        fields = fields.contiguous()
        cells = flat[:nspecies]
        parts = [t.contiguous() for t in flat[nspecies:]]
        _sync_torch()
        for k in range(6):
            sim.set_field_device(k, fields[k].data_ptr())
        for s in range(nspecies):
            p = parts[5 * s: 5 * s + 5]
            sim.set_particles_device(s, cells[s].numel(), cells[s].contiguous().data_ptr(),
                                     *[t.data_ptr() for t in p])
        loss = sim.adjoint_forward(nsteps, ckpt)
        fctx.sim = sim
        fctx.nspecies = nspecies
        fctx.meta = (fields.shape, fields.device, [(t.numel(), t.dtype) for t in parts])
        return torch.tensor(loss, dtype=torch.float64, device=fields.device)

    @staticmethod
    def backward(fctx, g):
        shape, dev, pmeta = fctx.meta
        fg = torch.zeros(shape, dtype=torch.float64, device=dev)
        pg = [torch.zeros(n, dtype=dt, device=dev) for n, dt in pmeta]
        _sync_torch()
        fctx.sim.adjoint_backward([fg[k].data_ptr() for k in range(6)], [t.data_ptr() for t in pg])
        g = g.to(torch.float64)
        out = [None, None, None, None, fg * g]
        out += [None] * fctx.nspecies
        out += [t * g.to(t.dtype) for t in pg]
        return tuple(out)


def jdote(sim, nsteps, ckpt, fields, species):
    """Differentiable time-integrated J.E (see the module docstring)."""
    flat = [s["cell"] for s in species] + [s[k] for s in species for k in PART]
    return _JdotE.apply(sim, int(nsteps), int(ckpt), len(species), fields, *flat)
