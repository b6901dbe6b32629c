"""torch.autograd binding of the band-energy adjoint.

E(V, R) = sum_i f_i <psi_i|H[V, R]|psi_i> for a fixed band block psi, with V
the local potential on the real grid and R the atom positions entering H
through the Gaussian projectors.  forward() loads V and the projectors at R
into the device context and evaluates E on the GPU; backward() returns

    dE/dV(r) = sum_i f_i psi_i(r)^2 / N          (pw_density)
    dE/dR_a  = -F_a,  F = nonlocal force          (pw_nonlocal_force)

both computed by the library's kernels (the density recomputes bands from the
sphere in chunks instead of keeping real-space psi), in the precision of psi
(complex128: FP64, complex64: FP32; the sums are FP64 either way).

The device context holds one set of projectors.  forward() keeps the positions
it evaluated E at, and backward() loads them again when another evaluation
has moved the context's projectors since (e.g. (E1 + E2).backward(), or a
finite-difference loop that reuses the context), so each gradient is taken at
its own forward's positions.
"""
import numpy as np
import torch


class BandEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vtot, pos, pwctx, psi, f):
        pwctx.set_vtot(vtot.detach().cpu().numpy())
        pos_np = np.array(pos.detach().cpu().numpy(), dtype=np.float64)
        pwctx.set_atoms(pos_np)
        E = pwctx.band_energy(psi, f)
        ctx.pwctx, ctx.psi, ctx.f = pwctx, psi, np.asarray(f, np.float64)
        ctx.pos = pos_np
        ctx.vshape, ctx.natom = tuple(vtot.shape), pos.shape[0]
        ctx.vdev, ctx.pdev = vtot.device, pos.device
        return torch.tensor(E, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gE):
        p = ctx.pwctx
        if p.atoms_loaded is None or not np.array_equal(p.atoms_loaded, ctx.pos):
            p.set_atoms(ctx.pos)
        dV = p.density(ctx.psi, ctx.f).reshape(ctx.vshape).to(ctx.vdev)
        npa = p.prob.nbeta // ctx.natom
        F = p.nonlocal_force(ctx.psi, ctx.f, ctx.natom, npa)
        dR = torch.from_numpy(-F).to(ctx.pdev)
        g = gE.to(torch.float64)
        return g * dV, g * dR, None, None, None


def band_energy(vtot, pos, pwctx, psi, f):
    """E(V, R) with autograd support for vtot (n3, n2, n1) and pos (natom, 3)."""
    return BandEnergy.apply(vtot, pos, pwctx, psi, f)
