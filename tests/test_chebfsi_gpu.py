"""ChebFSI on the GPU against the FP64 oracle.

Tolerances (FP64): Ritz values 1e-10 Ha, Lanczos bounds 1e-10, filtered
blocks 1e-10 relative; mixed (FP32 filter, FP64 Cholesky QR + RR):
Ritz values 1e-8 Ha against the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
from pwdft_develop_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx8(si8):
    from pwdft_develop_amd.ops import Context
    return Context(si8)


def test_orthogonalize(ctx8, si8):
    X0 = si8.make_psi(si8.nb, orthonormalise=False)
    X = ctx8.to_device(X0)
    ctx8.orthogonalize(X)
    Xg = X.cpu().numpy()
    Xo = X0.copy()
    buf = Xo.view(np.float64).reshape(-1).copy()
    assert O.lib().or_orthogonalize(si8.npw, si8.nb, buf) == 0
    ref = buf.view(np.complex128).reshape(X0.shape)
    assert np.abs(Xg - ref).max() < 1e-11 * np.abs(ref).max()


def test_orthogonalize_1024_bands():
    """Cholesky QR of 1024 bands (the 1024 x 1024 Gram product takes the
    128 x 128 split-K tiles, the factor rocSOLVER's potrf): orthonormal to
    1e-12 and the same span, X0 = Q R with R upper triangular."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.si64_60(nb=1024).build(with_psi=False)
    print("case: si64_60 orthogonalize nb=1024", flush=True)
    ctx = Context(prob)
    g = torch.Generator(device="cuda").manual_seed(3)
    X0 = torch.randn((1024, 2 * prob.npw), dtype=torch.float64, device="cuda", generator=g)
    Q = X0.clone()
    ctx.orthogonalize(Q.view(torch.complex128))
    assert float((Q @ Q.T - torch.eye(1024, dtype=torch.float64, device="cuda")).abs().max()) < 1e-12
    R = Q @ X0.T          # R[i, j] = q_i . x_j, zero below the diagonal
    assert float(torch.tril(R, -1).abs().max() / R.abs().max()) < 1e-12
    assert float((R.T @ Q - X0).abs().max() / X0.abs().max()) < 1e-12
    ctx.close()


def test_lanczos(ctx8, si8):
    v0 = si8.lanczos_start()
    th, b = ctx8.lanczos(8, ctx8.to_device(v0))
    th_o, b_o = O.Hamiltonian(si8).lanczos(8, v0)
    assert np.abs(th - th_o).max() < 1e-10
    assert abs(b - b_o) < 1e-10


def test_cheb_filter_fp64(ctx8, si8):
    X = ctx8.to_device(si8.psi)
    ctx8.cheb_filter(X, 10, 0.3, 12.0, -1.0)
    ref = O.Hamiltonian(si8).cheb_filter(si8.psi, 10, 0.3, 12.0, -1.0)
    got = X.cpu().numpy()
    assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()


def test_rayleigh_ritz_fp64(ctx8, si8):
    X = ctx8.to_device(si8.psi)
    _, e, r = ctx8.rayleigh_ritz(X)
    Xo, eo, ro = O.Hamiltonian(si8).rayleigh_ritz(si8.psi)
    assert np.abs(e - eo).max() < 1e-10
    assert np.abs(r - ro).max() < 1e-9
    assert np.abs(X.cpu().numpy() - Xo).max() < 1e-8


def test_chebfsi_fp64_sweeps_match_oracle_and_dense(ctx8, si8, golden):
    """5 FP64 sweeps (degree 10, 8 Lanczos steps) from the seeded block:
    Ritz values within 1e-10 Ha of the oracle's after every sweep."""
    print("case: si8_32 chebfsi 5 sweeps fp64", flush=True)
    v0 = si8.lanczos_start()
    H = O.Hamiltonian(si8)
    X = ctx8.to_device(si8.psi)
    Xo = si8.psi.copy()
    prev = prev_o = None
    for _ in range(5):
        _, e, r, b = ctx8.chebfsi_sweep(X, 10, 8, ctx8.to_device(v0), prev)
        Xo, eo, ro, bo = H.chebfsi_sweep(Xo, 10, 8, v0, prev_o)
        prev, prev_o = e, eo
        assert np.abs(e - eo).max() < 1e-10, np.abs(e - eo).max()
    # converged lowest states approach the dense-H spectrum
    assert np.abs(e[:16] - golden["evals"][:16]).max() < 1e-6


def test_chebfsi_mixed_sweep_on_warm_block(ctx8, si8):
    """SCF-like use: the block is warm (4 FP64 oracle sweeps).  One mixed
    sweep (FP32 filter, FP64 Cholesky QR + RR) against one FP64 oracle sweep
    from the same block: Ritz values of the wanted states (lowest 3/4 of
    the block) within 1e-8 Ha."""
    print("case: si8_32 mixed sweep on warm block", flush=True)
    v0 = si8.lanczos_start()
    H = O.Hamiltonian(si8)
    Xo = si8.psi.copy()
    prev = None
    for _ in range(4):
        Xo, prev, _, _ = H.chebfsi_sweep(Xo, 10, 8, v0, prev)
    Xn, en, rn, _ = H.chebfsi_sweep(Xo, 10, 8, v0, prev)
    X = ctx8.to_device(Xo)
    _, e, r, _ = ctx8.chebfsi_sweep(X, 10, 8, ctx8.to_device(v0), prev, mixed=True)
    err = np.abs(e - en)
    nwant = 3 * si8.nb // 4
    print(f"mixed-vs-fp64 Ritz err: wanted max {err[:nwant].max():.2e}, all max {err.max():.2e}")
    assert err[:nwant].max() < 1e-8


def test_chebfsi_si64_sweeps():
    """The si64_60 sweep shape on a 32-band slice.  From the seeded block:
    FP64 sweep vs oracle within 1e-10 Ha; the mixed sweep from the same cold
    block within 1e-6 Ha (FP32 rounding moves the unconverged subspace).
    After 3 FP64 oracle sweeps (warm block): mixed vs FP64 oracle within
    1e-8 Ha on the wanted (lowest 24) states."""
    from pwdft_develop_amd.ops import Context
    prob = W.si64_60(nb=32).build()
    print("case: si64_60 nb=32 chebfsi sweeps", flush=True)
    ctx = Context(prob)
    H = O.Hamiltonian(prob)
    v0 = prob.lanczos_start()
    dv0 = ctx.to_device(v0)
    Xo, eo, ro, bo = H.chebfsi_sweep(prob.psi, 10, 8, v0)
    X = ctx.to_device(prob.psi)
    _, e, r, b = ctx.chebfsi_sweep(X, 10, 8, dv0)
    assert np.abs(b - bo).max() < 1e-9
    assert np.abs(e - eo).max() < 1e-10, np.abs(e - eo).max()
    X = ctx.to_device(prob.psi)
    _, e, r, b = ctx.chebfsi_sweep(X, 10, 8, dv0, mixed=True)
    assert np.abs(e - eo).max() < 1e-6, np.abs(e - eo).max()
    prev = eo
    for _ in range(2):
        Xo, prev, _, _ = H.chebfsi_sweep(Xo, 10, 8, v0, prev)
    Xn, en, _, _ = H.chebfsi_sweep(Xo, 10, 8, v0, prev)
    X = ctx.to_device(Xo)
    _, e, r, b = ctx.chebfsi_sweep(X, 10, 8, dv0, prev, mixed=True)
    err = np.abs(e - en)
    print(f"si64 warm mixed-vs-fp64 Ritz err: wanted {err[:24].max():.2e} all {err.max():.2e}")
    assert err[:24].max() < 1e-8
    ctx.close()


@pytest.mark.parametrize("split,tol", [("3", 2e-5), ("2", 5e-5), ("0", 2e-5)])
def test_cheb_filter_fp32_projector_products(si8, split, tol, monkeypatch):
    """The FP32 degree-10 filter against the FP64 oracle filter, with the
    projector products inside the filter as products of split BF16 terms on
    MFMA (FP32 accumulation): 3 terms per operand (24 bits, six term
    products; the default), 2 terms (16 bits, three products; PW_BF16=2), or
    FP32 MFMA products (PW_BF16=0).  Relative errors within 2e-5, 5e-5 and
    2e-5 of the largest output entry."""
    import torch
    from pwdft_develop_amd.ops import Context
    print(f"case: si8_32 fp32 filter split={split}", flush=True)
    monkeypatch.setenv("PW_BF16", split)
    ctx = Context(si8)
    X = ctx.to_device(si8.psi, torch.complex64)
    ctx.cheb_filter(X, 10, 0.3, 12.0, -1.0)
    ref = O.Hamiltonian(si8).cheb_filter(si8.psi, 10, 0.3, 12.0, -1.0)
    err = float(np.abs(X.cpu().numpy() - ref).max() / np.abs(ref).max())
    print(f"fp32 filter (split {split}) relative error {err:.2e}")
    assert err < tol
    ctx.close()


@pytest.mark.parametrize("name,nb", [("si8_32", 13), ("rand16_48", 70), ("rand16_48", 128)])
def test_cheb_filter_presplit_products_ragged(name, nb, monkeypatch):
    """The FP32 filter's projector apply on operands split ahead of time
    (64-row x 256-column workgroups, partial row and column tiles: 13 and
    70 bands, 2 npw not a multiple of 64) against the same filter on FP32
    MFMA products (PW_BF16=0): 3 split terms carry FP32's 24 bits, so the two
    agree to a few FP32 roundings; and against the FP64 oracle filter."""
    import torch
    from pwdft_develop_amd.ops import Context
    prob = W.WORKLOADS[name](nb=nb).build()
    print(f"case: {name} presplit filter nb={nb}", flush=True)
    out = {}
    for split in ("3", "0"):
        monkeypatch.setenv("PW_BF16", split)
        ctx = Context(prob)
        X = ctx.to_device(prob.psi, torch.complex64)
        ctx.cheb_filter(X, 10, 0.3, 12.0, -1.0)
        out[split] = X.cpu().numpy()
        ctx.close()
    ref = O.Hamiltonian(prob).cheb_filter(prob.psi, 10, 0.3, 12.0, -1.0)
    scale = np.abs(ref).max()
    d = float(np.abs(out["3"] - out["0"]).max() / scale)
    e = float(np.abs(out["3"] - ref).max() / scale)
    print(f"split 3 vs FP32 MFMA {d:.2e}, vs oracle {e:.2e}")
    assert d < 5e-6
    assert e < 2e-5
