"""Screened exact exchange on the GPU against the oracle's 3-FFT pair loop
(spinor.cpp:783-945).  FP64 1e-11 relative, FP32 1e-5."""
import numpy as np
import pytest

import oracle as O
from pwdft_develop_amd import workload as W

pytestmark = pytest.mark.gpu

MU, FRAC = 0.106, 0.25


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name,nocc", [("small", 4), ("si8_32", 16)])
def test_add_exx_matches_oracle(name, nocc):
    import torch
    from pwdft_develop_amd.ops import Context
    if name == "small":
        prob = W.Problem("exx12", (12, 12, 12), W.cubic(7.0), 5.0, W.DIAMOND[:2] * 7.0, 6).build()
    else:
        prob = W.si8_32(nb=24).build()
    print(f"case: {name} exx nocc={nocc}", flush=True)
    occ = np.linspace(1.0, 0.3, nocc)
    H = O.Hamiltonian(prob)
    div = H.set_exx(prob.ecut, MU, FRAC, prob.psi[:nocc], occ)
    ref = H.add_exx(prob.psi)
    ctx = Context(prob)
    ctx.set_exx(prob.ecut, MU, FRAC, ctx.to_device(prob.psi[:nocc]), occ)
    got = ctx.add_exx(ctx.to_device(prob.psi)).cpu().numpy()
    assert _rel(got, ref) < 1e-11
    got32 = ctx.add_exx(ctx.to_device(prob.psi, torch.complex64)).cpu().numpy()
    assert _rel(got32, ref) < 1e-5
    # MultSpinor with the hybrid term active = local + kinetic + V_x
    full = ctx.mult_spinor(ctx.to_device(prob.psi)).cpu().numpy()
    assert _rel(full, H.mult_spinor(prob.psi)) < 1e-11
    ctx.close()


def test_add_exx_layered_subgrid():
    """The 432-atom layered cell on its 96 x 96 x 40 exchange subgrid (Ecut
    2.8 Ha, the Fock sphere one index short of the x and y Nyquist planes;
    the pair kernels take the grouped y layout there), 24 occupied bands:
    two bands against the oracle."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.layered432_96(nb=24).build()
    nocc = 24
    print("case: layered432_96 exx nocc=24", flush=True)
    occ = np.linspace(1.0, 0.5, nocc)
    H = O.Hamiltonian(prob)
    H.set_exx(prob.ecut, MU, FRAC, prob.psi[:nocc], occ)
    ref = H.add_exx(prob.psi[:2])
    ctx = Context(prob)
    ctx.set_exx(prob.ecut, MU, FRAC, ctx.to_device(prob.psi[:nocc]), occ)
    got = ctx.add_exx(ctx.to_device(prob.psi)).cpu().numpy()[:2]
    assert _rel(got, ref) < 1e-11
    got32 = ctx.add_exx(ctx.to_device(prob.psi, torch.complex64)).cpu().numpy()[:2]
    assert _rel(got32, ref) < 1e-5
    ctx.close()


@pytest.mark.parametrize("pair,wfn", [("5:3:512", "3:5:1024"), ("7:5:1024", "4:3:256")])
def test_add_exx_forced_grouped_layouts(pair, wfn, monkeypatch):
    """The grouped y layouts of the pair kernels (PW_EXX_LAYOUT) and of the
    wavefunction-side slab kernels (PW_WFN_LAYOUT) forced on the si8_32
    grid with column groups and row-pair counts that are not powers of two
    and leave ragged last groups (16 row pairs in steps of 3 or 5): FP64
    against the oracle."""
    from pwdft_develop_amd.ops import Context
    prob = W.si8_32(nb=6).build()
    nocc = 5
    print(f"case: si8_32 exx layouts pair={pair} wfn={wfn}", flush=True)
    occ = np.linspace(1.0, 0.4, nocc)
    H = O.Hamiltonian(prob)
    H.set_exx(prob.ecut, MU, FRAC, prob.psi[:nocc], occ)
    ref = H.add_exx(prob.psi)
    monkeypatch.setenv("PW_EXX_PAIR", "lds")  # the LDS pair kernels, whose layouts these are
    monkeypatch.setenv("PW_EXX_LAYOUT", pair)
    monkeypatch.setenv("PW_WFN_LAYOUT", wfn)
    ctx = Context(prob)
    ctx.set_exx(prob.ecut, MU, FRAC, ctx.to_device(prob.psi[:nocc]), occ)
    got = ctx.add_exx(ctx.to_device(prob.psi)).cpu().numpy()
    assert _rel(got, ref) < 1e-11
    f = np.linspace(2.0, 0.5, prob.nb)
    rho = ctx.density(ctx.to_device(prob.psi), f).cpu().numpy()
    rho_o = H.density(prob.psi, f)
    assert _rel(rho, rho_o) < 1e-12
    ctx.close()


@pytest.mark.parametrize("name,nb,nocc", [("si8_32", 7, 5), ("rand16_48", 5, 3), ("layered432_96", 3, 4)])
def test_register_pair_kernels(name, nb, nocc, monkeypatch):
    """The register-FFT pair kernels (k_exx_pair_fwd_reg / _acc_reg: 4 lanes
    per 1D transform, the accumulator in registers over the occupied bands;
    PW_EXX_PAIR=reg forces them in FP64 too) against the LDS pair kernels
    (PW_EXX_PAIR=lds) on square planes of 32, 48 and 96 points, FP64 and
    FP32; parity of the defaults (FP32 register, FP64 LDS) with the oracle
    is the other tests'."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.WORKLOADS[name](nb=nb).build()
    print(f"case: {name} exx register vs lds pair kernels nb={nb} nocc={nocc}", flush=True)
    occ = np.linspace(1.0, 0.4, nocc)
    out = {}
    for mode in ("reg", "lds"):
        monkeypatch.setenv("PW_EXX_PAIR", mode)
        ctx = Context(prob)
        ctx.set_exx(prob.ecut, MU, FRAC, ctx.to_device(prob.psi[:nocc]), occ)
        out[mode] = (ctx.add_exx(ctx.to_device(prob.psi)).cpu().numpy(),
                     ctx.add_exx(ctx.to_device(prob.psi, torch.complex64)).cpu().numpy())
        ctx.close()
    monkeypatch.delenv("PW_EXX_PAIR", raising=False)
    assert _rel(out["reg"][0], out["lds"][0]) < 1e-12
    assert _rel(out["reg"][1], out["lds"][1]) < 2e-6


def test_density_large_cell():
    """CalculateDensity on the 180 x 180 x 72 cell: the FP64 plane's
    double-buffered y-space (265 KiB) exceeds the LDS, so the density kernel
    takes the grouped layout; FP64 and FP32 accumulation of four bands and
    the renormalised density against the oracle."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.layered432_180(nb=4).build()
    print("case: layered432_180 density fp64/fp32 nb=4", flush=True)
    f = np.array([2.0, 2.0, 1.5, 0.5])
    H = O.Hamiltonian(prob)
    rho_o = H.density(prob.psi, f)
    ctx = Context(prob)
    rho = ctx.density(ctx.to_device(prob.psi), f).cpu().numpy()
    assert _rel(rho, rho_o) < 1e-11
    rho32 = ctx.density(ctx.to_device(prob.psi, torch.complex64), f).cpu().numpy()
    assert _rel(rho32, rho_o) < 1e-5
    rho_n, val = ctx.calculate_density(ctx.to_device(prob.psi), f / 2, num_spin=2, num_occupied=3)
    ref = rho_o * (2 * 3 * prob.ngrid / (prob.volume * rho_o.sum()))
    assert _rel(rho_n.cpu().numpy(), ref) < 1e-11
    assert abs(val - 6.0) < 1e-10
    ctx.close()


def test_exx_divergence_matches_oracle():
    import ctypes as C
    from pwdft_develop_amd import _lib
    prob = W.si8_32(nb=4).build()
    H = O.Hamiltonian(prob)
    d_o = H.set_exx(prob.ecut, MU, FRAC, prob.psi[:2], np.ones(2))
    d = C.c_double()
    _lib.check(_lib.lib().pw_exx_divergence(np.asarray(prob.grid, np.int32), prob.supercell.reshape(-1).copy(),
                                            prob.ecut, MU, 1, C.byref(d)))
    assert abs(d.value - d_o) < 1e-10 * abs(d_o)


def test_ace_matches_oracle():
    """ACE build (hamiltonian.cpp:3685-3924) and apply (:3347-3430): the GPU
    operator against the oracle's on the si8_32 shape, FP64 1e-11 relative on
    H psi with ACE active for the building block and for a block outside its
    span, FP32 1e-5; dropping it restores the full exchange."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.si8_32(nb=24).build()
    nocc = 16
    print("case: si8_32 ace nb=24 nocc=16", flush=True)
    occ = np.linspace(1.0, 0.3, nocc)
    H = O.Hamiltonian(prob)
    H.set_exx(prob.ecut, MU, FRAC, prob.psi[:nocc], occ)
    full_ref = H.mult_spinor(prob.psi)
    H.build_ace(prob.psi)
    ref = H.mult_spinor(prob.psi)
    rng = np.random.default_rng(5)
    other = rng.standard_normal((8, prob.npw)) + 1j * rng.standard_normal((8, prob.npw))
    other[:, 0] = other[:, 0].real
    ref_o = H.mult_spinor(other)
    ctx = Context(prob)
    ctx.set_exx(prob.ecut, MU, FRAC, ctx.to_device(prob.psi[:nocc]), occ)
    ctx.build_ace(ctx.to_device(prob.psi))
    got = ctx.mult_spinor(ctx.to_device(prob.psi)).cpu().numpy()
    assert _rel(got, ref) < 1e-11
    assert _rel(got, full_ref) < 1e-10   # ACE reproduces V_x on its block
    got_o = ctx.mult_spinor(ctx.to_device(other)).cpu().numpy()
    assert _rel(got_o, ref_o) < 1e-11
    got32 = ctx.mult_spinor(ctx.to_device(prob.psi, torch.complex64)).cpu().numpy()
    assert _rel(got32, ref) < 1e-5
    ctx.build_ace(None)
    back = ctx.mult_spinor(ctx.to_device(prob.psi)).cpu().numpy()
    assert _rel(back, full_ref) < 1e-11
    ctx.close()
