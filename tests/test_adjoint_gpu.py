"""Adjoint of the band energy on the GPU: density (dE/dV) and nonlocal
forces (dE/dR) against the oracle (1e-10 relative), and the torch.autograd
binding against central finite differences of E (steps 1e-5 Ha, 1e-4 bohr)
on the 16-atom 48^3 cell with 128 bands and f = 2."""
import numpy as np
import pytest

import oracle as O
from pwdft_develop_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def r16():
    return W.rand16_48(nb=128).build()


def test_density_energy_force_match_oracle(r16):
    import torch
    from pwdft_develop_amd.ops import Context
    print("case: rand16_48 adjoint vs oracle", flush=True)
    f = np.full(r16.nb, 2.0)
    ctx = Context(r16)
    X = ctx.to_device(r16.psi)
    H = O.Hamiltonian(r16)
    rho = ctx.density(X, f).cpu().numpy()
    rho_o = H.density(r16.psi, f)
    assert np.abs(rho - rho_o).max() < 1e-11 * np.abs(rho_o).max()
    rho32 = ctx.density(ctx.to_device(r16.psi, torch.complex64), f).cpu().numpy()
    assert np.abs(rho32 - rho_o).max() < 1e-5 * np.abs(rho_o).max()
    E = ctx.band_energy(X, f)
    assert abs(E - H.band_energy(r16.psi, f)) < 1e-10 * abs(E)
    F = ctx.nonlocal_force(X, f, len(r16.atoms), 4)
    F_o = H.nonlocal_force(r16.psi, f, len(r16.atoms), 4)
    assert np.abs(F - F_o).max() < 1e-10 * max(1.0, np.abs(F_o).max())
    ctx.close()


def test_autograd_against_finite_differences(r16):
    import torch
    from pwdft_develop_amd.adjoint import band_energy
    from pwdft_develop_amd.ops import Context
    print("case: rand16_48 autograd vs central differences", flush=True)
    f = np.full(r16.nb, 2.0)
    ctx = Context(r16)
    X = ctx.to_device(r16.psi)
    V = torch.tensor(r16.vtot, dtype=torch.float64, requires_grad=True)
    R = torch.tensor(r16.atoms, dtype=torch.float64, requires_grad=True)
    E = band_energy(V, R, ctx, X, f)
    E.backward()
    gV, gR = V.grad.numpy().reshape(-1), R.grad.numpy()
    v0 = r16.vtot.reshape(-1)
    rng = np.random.default_rng(11)
    hv = 1e-5
    for r in rng.integers(0, v0.size, 4):
        vals = []
        for s in (+1, -1):
            v = v0.copy()
            v[r] += s * hv
            vals.append(band_energy(torch.tensor(v.reshape(r16.vtot.shape)), R.detach(), ctx, X, f).item())
        fd = (vals[0] - vals[1]) / (2 * hv)
        assert abs(fd - gV[r]) < 1e-6 * max(1.0, abs(gV[r])), (fd, gV[r])
    hr = 1e-4
    for a, d in ((0, 0), (7, 1), (15, 2)):
        vals = []
        for s in (+1, -1):
            pos = r16.atoms.copy()
            pos[a, d] += s * hr
            vals.append(band_energy(V.detach(), torch.tensor(pos), ctx, X, f).item())
        fd = (vals[0] - vals[1]) / (2 * hr)
        assert abs(fd - gR[a, d]) < 1e-6 * max(1.0, abs(gR[a, d])), (fd, gR[a, d])
    ctx.close()


def test_calculate_density_normalisation(r16):
    """KohnSham::CalculateDensity (hamiltonian.cpp:2453-2627): the GPU density
    is the oracle's accumulation rescaled so that sum(rho) vol / N equals
    numSpin * numOccupiedState (the reference's val), for uneven
    occupations."""
    from pwdft_develop_amd.ops import Context
    print("case: rand16_48 calculate_density", flush=True)
    occ = np.linspace(1.0, 0.0, r16.nb)
    ctx = Context(r16)
    rho, val = ctx.calculate_density(ctx.to_device(r16.psi), occ, num_spin=2, num_occupied=64)
    rho = rho.cpu().numpy()
    N, vol = r16.ngrid, r16.volume
    assert abs(val - 2 * 64) < 1e-10 * 128
    assert abs(rho.sum() * vol / N - 128) < 1e-10 * 128
    raw = O.Hamiltonian(r16).density(r16.psi, 2 * occ)
    ref = raw * (2 * 64 * N / (vol * raw.sum()))
    assert np.abs(rho - ref).max() < 1e-11 * np.abs(ref).max()
    ctx.close()
