"""PPCG on the GPU against the FP64 oracle's restatement of PPCGSolveReal
(ppcg.cpp:119-576, same algorithm): Ritz values 1e-10 Ha and residual norms
1e-8 after the same iterations, with full and ragged sub-blocks; the GPU's
small Sygvd problems go through its batched generalized eigensolver
(pw_sygvd_batched), the oracle's through its own Cholesky + Householder/QL
solver, with the same eigenvector sign rule.  Sub-blocks of up to 32
columns take the fused block update, wider ones the GEMM path."""
import numpy as np
import pytest

import oracle as O
from pwdft_develop_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("max_iter,scf_iter,sb", [(1, 1, 32), (3, 1, 32), (4, 2, 12), (3, 1, 8)])
def test_ppcg_matches_oracle(si8, max_iter, scf_iter, sb):
    from pwdft_develop_amd.ops import Context
    print(f"case: si8_32 ppcg iters={max_iter} scfIter={scf_iter} sb={sb}", flush=True)
    ctx = Context(si8)
    X = ctx.to_device(si8.psi)
    _, e, r, it = ctx.ppcg(X, max_iter=max_iter, scf_iter=scf_iter, sb_size=sb)
    Xo, eo, ro, ito = O.Hamiltonian(si8).ppcg(si8.psi, max_iter=max_iter, scf_iter=scf_iter, sb_size=sb)
    assert it == ito == max_iter
    assert np.abs(e - eo).max() < 1e-10
    assert np.abs(r - ro).max() < 1e-8
    ctx.close()


def test_ppcg_converges_to_dense(si8, golden):
    from pwdft_develop_amd.ops import Context
    print("case: si8_32 ppcg to convergence", flush=True)
    ctx = Context(si8)
    X = ctx.to_device(si8.psi)
    _, e, r, it = ctx.ppcg(X, num_eig=24, max_iter=60, tol=1e-7, sb_size=32)
    assert it < 60
    assert np.abs(e[:24] - golden["evals"][:24]).max() < 1e-11
    ctx.close()


@pytest.mark.parametrize("sb", [16, 24, 40])
def test_ppcg_rand16(sb):
    from pwdft_develop_amd.ops import Context
    prob = W.rand16_48(nb=40).build()
    print(f"case: rand16_48 nb=40 ppcg 3 iterations sb={sb}", flush=True)
    ctx = Context(prob)
    X = ctx.to_device(prob.psi)
    _, e, r, it = ctx.ppcg(X, max_iter=3, sb_size=sb)
    _, eo, ro, _ = O.Hamiltonian(prob).ppcg(prob.psi, max_iter=3, sb_size=sb)
    assert np.abs(e - eo).max() < 1e-10
    assert np.abs(r - ro).max() < 1e-8
    ctx.close()
