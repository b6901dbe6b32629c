"""LOBPCG on the GPU against the FP64 oracle's restatement of
LOBPCGSolveReal (same algorithm, same iterates): eigenvalues 1e-10 Ha,
residual norms 1e-8, eigenvectors (sign-normalised) 1e-7."""
import numpy as np
import pytest

import oracle as O
from pwdft_develop_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("max_iter,scf_iter", [(1, 1), (3, 1), (4, 2)])
def test_lobpcg_matches_oracle(si8, max_iter, scf_iter):
    from pwdft_develop_amd.ops import Context
    print(f"case: si8_32 lobpcg iters={max_iter} scfIter={scf_iter}", flush=True)
    ctx = Context(si8)
    X = ctx.to_device(si8.psi)
    _, e, r, it = ctx.lobpcg(X, max_iter=max_iter, scf_iter=scf_iter)
    Xo, eo, ro, ito = O.Hamiltonian(si8).lobpcg(si8.psi, max_iter=max_iter, scf_iter=scf_iter)
    assert it == ito == max_iter
    assert np.abs(e - eo).max() < 1e-10
    assert np.abs(r - ro).max() < 1e-8
    assert np.abs(X.cpu().numpy() - Xo).max() < 1e-7
    ctx.close()


def test_lobpcg_converges_to_dense(si8, golden):
    from pwdft_develop_amd.ops import Context
    ctx = Context(si8)
    X = ctx.to_device(si8.psi)
    _, e, r, it = ctx.lobpcg(X, max_iter=40, tol=1e-8)
    assert np.abs(e[:28] - golden["evals"][:28]).max() < 1e-8
    ctx.close()


def test_lobpcg_rand16():
    from pwdft_develop_amd.ops import Context
    prob = W.rand16_48(nb=24).build()
    print("case: rand16_48 nb=24 lobpcg 3 iterations", flush=True)
    ctx = Context(prob)
    X = ctx.to_device(prob.psi)
    _, e, r, it = ctx.lobpcg(X, max_iter=3)
    Xo, eo, ro, ito = O.Hamiltonian(prob).lobpcg(prob.psi, max_iter=3)
    assert np.abs(e - eo).max() < 1e-10
    ctx.close()


def test_lobpcg_errors(si8):
    from pwdft_develop_amd.ops import Context
    from pwdft_develop_amd._lib import PwError
    ctx = Context(si8)
    X = ctx.to_device(si8.psi)
    with pytest.raises(PwError):
        ctx.lobpcg(X, num_eig=si8.nb + 1)
    ctx.close()
