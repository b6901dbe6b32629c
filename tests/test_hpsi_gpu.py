"""H*psi on the GPU through the C-ABI against the FP64 oracle and the dense-H
golden vectors.  Tolerances: FP64 1e-11 relative (max-norm), FP32 2e-5."""
import numpy as np
import pytest

import oracle as O
from pwdft_develop_amd import workload as W

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def ctx8(si8):
    from pwdft_develop_amd.ops import Context
    return Context(si8)


def test_apply_h_fp64_matches_oracle(ctx8, si8):
    import torch
    print(f"case: si8_32 apply_h fp64 nb={si8.nb}", flush=True)
    X = ctx8.to_device(si8.psi)
    HX = ctx8.apply_h(X)
    torch.cuda.synchronize()
    ref = O.Hamiltonian(si8).apply_h(si8.psi)
    assert _rel(HX.cpu().numpy(), ref) < 1e-11


def test_apply_h_fp64_matches_dense_golden(ctx8, golden):
    X = ctx8.to_device(golden["psi4"])
    HX = ctx8.apply_h(X).cpu().numpy()
    assert _rel(HX, golden["hpsi4"]) < 1e-11


def test_mult_spinor_and_nonlocal_split(ctx8, si8):
    H = O.Hamiltonian(si8)
    X = ctx8.to_device(si8.psi[:7])
    loc = ctx8.mult_spinor(X).cpu().numpy()
    nl = ctx8.nonlocal_multx(X).cpu().numpy()
    assert _rel(loc, H.mult_spinor(si8.psi[:7])) < 1e-11
    assert _rel(nl, H.nonlocal_multx(si8.psi[:7])) < 1e-11


def test_apply_h_fp32(ctx8, si8):
    import torch
    X = ctx8.to_device(si8.psi, torch.complex64)
    HX = ctx8.apply_h(X).cpu().numpy()
    ref = O.Hamiltonian(si8).apply_h(si8.psi)
    assert _rel(HX, ref) < 2e-5


@pytest.mark.parametrize("name,nb", [("si8_32", 13), ("rand16_48", 8), ("si64_60", 24)])
def test_fp32_register_slab_kernel(name, nb, monkeypatch):
    """The FP32 register slab kernel (one z plane of 8 bands per workgroup,
    ragged last group) against the oracle and against the LDS slab kernel
    (same arithmetic; agreement to FP32 rounding)."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.WORKLOADS[name](nb=nb).build()
    print(f"case: {name} fp32 slab kernels nb={nb}", flush=True)
    ref = O.Hamiltonian(prob).apply_h(prob.psi)
    monkeypatch.delenv("PW_SLAB", raising=False)
    creg = Context(prob)
    monkeypatch.setenv("PW_SLAB", "lds")
    clds = Context(prob)
    X = creg.to_device(prob.psi, torch.complex64)
    a = creg.apply_h(X).cpu().numpy()
    b = clds.apply_h(X).cpu().numpy()
    assert _rel(a, ref) < 2e-5
    assert _rel(a, b) < 2e-6
    creg.close()
    clds.close()


@pytest.mark.parametrize("name,nb", [("si8_32", 13), ("rand16_48", 21), ("si64_60", 40)])
def test_fp32_wide_register_slab_kernel(name, nb, monkeypatch):
    """The 16-band, 512-thread FP32 register slab variant (PW_SLAB=reg:1:16;
    ragged last group) against the oracle and against the 8-band kernel
    (same arithmetic)."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.WORKLOADS[name](nb=nb).build()
    print(f"case: {name} fp32 wide slab nb={nb}", flush=True)
    ref = O.Hamiltonian(prob).apply_h(prob.psi)
    monkeypatch.setenv("PW_SLAB", "reg:1:16")
    cw = Context(prob)
    monkeypatch.delenv("PW_SLAB", raising=False)
    c8 = Context(prob)
    X = cw.to_device(prob.psi, torch.complex64)
    a = cw.apply_h(X).cpu().numpy()
    b = c8.apply_h(X).cpu().numpy()
    assert _rel(a, ref) < 2e-5
    assert np.array_equal(a, b)
    cw.close()
    c8.close()


@pytest.mark.parametrize("name,nb", [("si8_32", 13), ("rand16_48", 21), ("si64_60", 5), ("si64_60", 203)])
def test_fp32_persistent_register_slab_kernel(name, nb, monkeypatch):
    """The FP32 register slab kernel on a persistent grid (PW_SLAB=reg:1:8:p:
    one workgroup per resident slot, looping over (plane, band group) items
    with the next item's sticks and V plane prefetched; 203 bands at 60
    points give several items per workgroup and a ragged last group)
    against the oracle and against the one-item grid (same arithmetic).
    FP64 has no persistent variant and keeps the one-item kernel."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.WORKLOADS[name](nb=nb).build()
    print(f"case: {name} persistent slab nb={nb}", flush=True)
    ref = O.Hamiltonian(prob).apply_h(prob.psi)
    monkeypatch.setenv("PW_SLAB", "reg:1:8:p")
    cp = Context(prob)
    monkeypatch.setenv("PW_SLAB", "reg:1:8")
    c1 = Context(prob)
    monkeypatch.delenv("PW_SLAB", raising=False)
    X = cp.to_device(prob.psi, torch.complex64)
    a = cp.apply_h(X).cpu().numpy()
    b = c1.apply_h(X).cpu().numpy()
    assert _rel(a, ref) < 2e-5
    assert np.array_equal(a, b)
    if nb <= 21:
        X64 = cp.to_device(prob.psi, torch.complex128)
        a64 = cp.apply_h(X64).cpu().numpy()
        assert np.array_equal(a64, c1.apply_h(X64).cpu().numpy())
        assert _rel(a64, ref) < 1e-12
    cp.close()
    c1.close()


@pytest.mark.parametrize("name,nb", [("si8_32", 13), ("rand16_48", 5), ("si64_60", 9)])
def test_register_z_kernels(name, nb, monkeypatch):
    """The register z-pencil kernels (PW_ZREG=1: each stick's z column over
    4 lanes, the sphere read and written through the (z, stick) table)
    against the oracle, FP64 and FP32, and against the LDS z kernels."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.WORKLOADS[name](nb=nb).build()
    print(f"case: {name} register z kernels nb={nb}", flush=True)
    ref = O.Hamiltonian(prob).apply_h(prob.psi)
    monkeypatch.setenv("PW_ZREG", "1")
    cz = Context(prob)
    monkeypatch.delenv("PW_ZREG", raising=False)
    cl = Context(prob)
    for dt, tol in ((torch.complex128, 1e-12), (torch.complex64, 2e-5)):
        X = cz.to_device(prob.psi, dt)
        a = cz.apply_h(X).cpu().numpy()
        b = cl.apply_h(X).cpu().numpy()
        assert _rel(a, ref) < tol
        assert _rel(a, b) < tol
    cz.close()
    cl.close()


@pytest.mark.parametrize("name,nb", [("si8_32", 32), ("si8_32", 7), ("rand16_48", 9), ("si64_60", 5)])
def test_fp64_register_slab_kernel(name, nb, monkeypatch):
    """The FP64 register slab kernel (one z plane of 4 bands per workgroup,
    ragged last group; one wave per SIMD at 60 points) against the oracle
    (FP64 parity) and against the LDS slab kernel (same arithmetic)."""
    from pwdft_develop_amd.ops import Context
    prob = W.WORKLOADS[name](nb=nb).build()
    print(f"case: {name} fp64 slab kernels nb={nb}", flush=True)
    ref = O.Hamiltonian(prob).apply_h(prob.psi)
    monkeypatch.delenv("PW_SLAB", raising=False)
    creg = Context(prob)
    monkeypatch.setenv("PW_SLAB", "lds")
    clds = Context(prob)
    X = creg.to_device(prob.psi)
    a = creg.apply_h(X).cpu().numpy()
    b = clds.apply_h(X).cpu().numpy()
    assert _rel(a, ref) < 1e-12
    assert _rel(a, b) < 1e-13
    creg.close()
    clds.close()


@pytest.mark.parametrize("name,nb,groups", [("si8_32", 7, "3"), ("rand16_48", 5, "1"), ("si64_60", 3, "4"),
                                            ("si8_32", 5, "5:3:512"), ("rand16_48", 3, "7:5:1024")])
def test_grouped_y_slab_layout(name, nb, groups, monkeypatch):
    """The LDS slab kernel's large-plane layout (one y buffer, y transforms
    in groups of g columns through the x rows' scratch), forced on small
    grids: FP64 parity with the oracle, FP32 to single-precision rounding."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.WORKLOADS[name](nb=nb).build()
    print(f"case: {name} grouped y slab g={groups} nb={nb}", flush=True)
    ref = O.Hamiltonian(prob).apply_h(prob.psi)
    monkeypatch.setenv("PW_SLAB", "lds:" + groups)
    ctx = Context(prob)
    assert _rel(ctx.apply_h(ctx.to_device(prob.psi)).cpu().numpy(), ref) < 1e-12
    assert _rel(ctx.apply_h(ctx.to_device(prob.psi, torch.complex64)).cpu().numpy(), ref) < 2e-5
    ctx.close()


def test_apply_h_large_cell():
    """180 x 180 x 72 grid, 432 atoms (1728 projectors), npw 76241: neither
    precision's y-space fits LDS twice (FP64 265 KiB, FP32 135 KiB), so the
    slab kernel takes the grouped layout -- FP32 one group of all 45 columns
    with 22 row pairs at 1024 threads, FP64 groups of 8 columns with 4 row
    pairs.  Four bands against the oracle."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.layered432_180(nb=4).build()
    print("case: layered432_180 apply_h fp64/fp32 nb=4", flush=True)
    ctx = Context(prob)
    ref = O.Hamiltonian(prob).apply_h(prob.psi)
    assert _rel(ctx.apply_h(ctx.to_device(prob.psi)).cpu().numpy(), ref) < 1e-11
    assert _rel(ctx.apply_h(ctx.to_device(prob.psi, torch.complex64)).cpu().numpy(), ref) < 2e-5
    ctx.close()


@pytest.mark.parametrize("name,nb", [("layered432_180", 1024), ("si64_60", 1100)])
def test_nonlocal_large_outputs(name, nb):
    """Projector products whose outputs span 256 or more 64 x 64 tiles take
    the 128 x 128 split-K tiles: W = beta^T X (nbeta x nb over the real
    2 npw rows), Y = (beta w) W, against a torch FP64 matmul of the same
    real-view formula (FP64 1e-12, FP32 2e-5 relative); seeded random
    blocks, ragged tile edges at 1728 and 1100."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.WORKLOADS[name](nb=nb).build(with_psi=False)
    print(f"case: {name} nonlocal nb={nb}", flush=True)
    ctx = Context(prob)
    g = torch.Generator(device="cuda").manual_seed(7)
    X = torch.randn((nb, 2 * prob.npw), dtype=torch.float64, device="cuda", generator=g)
    Xc = X.view(torch.complex128)
    beta = torch.from_numpy(prob.beta.view(np.float64).reshape(prob.nbeta, -1)).cuda()
    wgt = torch.from_numpy(prob.wgt).cuda()
    ref = (beta @ X.T).T @ (beta * wgt[:, None])
    got = ctx.nonlocal_multx(Xc).view(torch.float64)
    assert float((got - ref).abs().max() / ref.abs().max()) < 1e-12
    got32 = ctx.nonlocal_multx(Xc.to(torch.complex64)).view(torch.float32).double()
    assert float((got32 - ref).abs().max() / ref.abs().max()) < 2e-5
    ctx.close()


def test_teter(ctx8, si8):
    X = ctx8.to_device(si8.psi[:3])
    T = ctx8.teter_precond(X).cpu().numpy()
    ref = O.Hamiltonian(si8).teter(si8.psi[:3])
    assert _rel(T, ref) < 1e-14


@pytest.mark.parametrize("name", ["rand16_48", "si64_60"])
def test_apply_h_other_grids(name):
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.WORKLOADS[name](nb=8).build()
    print(f"case: {name} apply_h fp64 nb=8", flush=True)
    ctx = Context(prob)
    HX = ctx.apply_h(ctx.to_device(prob.psi)).cpu().numpy()
    ref = O.Hamiltonian(prob).apply_h(prob.psi)
    assert _rel(HX, ref) < 1e-11
    HX32 = ctx.apply_h(ctx.to_device(prob.psi, torch.complex64)).cpu().numpy()
    assert _rel(HX32, ref) < 2e-5
    ctx.close()


def test_errors_surface_as_status(ctx8, si8):
    from pwdft_develop_amd._lib import PwError
    import torch
    X = ctx8.to_device(si8.psi)
    with pytest.raises(PwError):
        ctx8.apply_h(X, out=X)            # aliasing is rejected
    big = torch.zeros((si8.nb + 1, si8.npw), dtype=torch.complex128, device="cuda")
    with pytest.raises(PwError):
        ctx8.apply_h(big)                 # more bands than the context holds
