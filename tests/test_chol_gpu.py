"""Cholesky QR (eigensolver.cpp:386-469: Potrf 'U' + Trsm) on the GPU
against the oracle's restatement: every block size from one band to 256
(single-workgroup factor + inverse, MFMA GEMM), 300 (rocSOLVER path), an
ill-conditioned block, and LAPACK's info for a zero column."""
import numpy as np
import pytest

import oracle as O
from pwdft_develop_amd import workload as W
from pwdft_develop_amd._lib import PwError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rctx():
    from pwdft_develop_amd.ops import Context
    prob = W.rand16_48(nb=40).build()
    ctx = Context(prob, max_bands=300)
    yield prob, ctx
    ctx.close()


def _block(prob, nb, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nb, prob.npw)) + 1j * rng.standard_normal((nb, prob.npw))


def _oracle(npw, X0):
    buf = X0.view(np.float64).reshape(-1).copy()
    assert O.lib().or_orthogonalize(npw, X0.shape[0], buf) == 0
    return buf.view(np.complex128).reshape(X0.shape)


@pytest.mark.parametrize("nb", [1, 5, 31, 32, 33, 64, 100, 200, 255, 256, 300])
def test_orthogonalize_block_sizes(rctx, nb):
    prob, ctx = rctx
    print(f"case: rand16_48 orthogonalize nb={nb}", flush=True)
    X0 = _block(prob, nb, nb)
    X = ctx.to_device(X0)
    ctx.orthogonalize(X)
    ref = _oracle(prob.npw, X0)
    got = X.cpu().numpy()
    assert np.abs(got - ref).max() < 1e-11 * np.abs(ref).max()
    # orthonormal in the Re(a^H b) metric of the real (2 npw)-row view
    R = got.view(np.float64).reshape(nb, -1)
    assert np.abs(R @ R.T - np.eye(nb)).max() < 1e-12


def test_orthogonalize_ill_conditioned(rctx):
    prob, ctx = rctx
    nb = 160
    print(f"case: rand16_48 orthogonalize ill-conditioned nb={nb}", flush=True)
    X0 = _block(prob, nb, 7)
    # columns mixed by a unit upper triangle and scaled over three decades
    # (condition number ~1e3)
    rng = np.random.default_rng(8)
    C = np.triu(rng.standard_normal((nb, nb)), 1) * 0.1 + np.eye(nb)
    X0 = (C.T @ X0) * np.logspace(0, -3, nb)[:, None]
    X = ctx.to_device(X0)
    ctx.orthogonalize(X)
    ref = _oracle(prob.npw, X0)
    assert np.abs(X.cpu().numpy() - ref).max() < 1e-8 * np.abs(ref).max()


@pytest.mark.parametrize("nb,zero", [(100, 40), (256, 0), (256, 255), (300, 17)])
def test_orthogonalize_not_positive_definite(rctx, nb, zero):
    prob, ctx = rctx
    print(f"case: rand16_48 orthogonalize zero column {zero} of {nb}", flush=True)
    X0 = _block(prob, nb, 3)
    X0[zero] = 0.0
    X = ctx.to_device(X0)
    with pytest.raises(PwError, match=rf"not positive definite \(info = {zero + 1}\)"):
        ctx.orthogonalize(X)
    # the context stays usable
    Y0 = _block(prob, nb, 4)
    Y = ctx.to_device(Y0)
    ctx.orthogonalize(Y)
    ref = _oracle(prob.npw, Y0)
    assert np.abs(Y.cpu().numpy() - ref).max() < 1e-11 * np.abs(ref).max()
