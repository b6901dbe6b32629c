"""The device dense symmetric eigensolver (lapack::Syevd('V','U') of the
reference's Rayleigh-Ritz and SolveGenEig) against LAPACK through numpy and
against the oracle's Householder + QL solver.

Bars, for ||A|| ~ 1: eigenvalues 1e-12 absolute; residual ||A z - lambda z||
and orthogonality ||Z^T Z - I|| 1e-11; eigenvectors equal to the oracle's
(same sign convention) 1e-9 where the spectrum is well separated."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from pwdft_develop_amd import workload as W
    from pwdft_develop_amd.ops import Context
    c = Context(W.si8_32(nb=8).build())
    yield c
    c.close()


def _solve(ctx, A):
    import torch
    w, Z = ctx.syevd(torch.from_numpy(np.ascontiguousarray(A)).to(ctx.device))
    # column-major result: row j of the torch view is eigenvector j
    return w.cpu().numpy(), Z.cpu().numpy().T


def _check(A, w, V, tol=1e-11):
    n = A.shape[0]
    scale = max(1.0, np.abs(A).max())
    ref = np.linalg.eigvalsh(A)
    assert np.abs(w - ref).max() < 1e-12 * scale * max(1, n / 64), np.abs(w - ref).max()
    assert np.all(np.diff(w) >= 0)
    assert np.abs(A @ V - V * w).max() < tol * scale * max(1, n / 64)
    assert np.abs(V.T @ V - np.eye(n)).max() < tol * max(1, n / 64)
    # sign convention: largest-magnitude component positive
    idx = np.argmax(np.abs(V), axis=0)
    assert np.all(V[idx, np.arange(n)] > 0)


def _rand_sym(n, seed):
    r = np.random.default_rng(seed).standard_normal((n, n))
    return (r + r.T) / np.sqrt(2 * n)


@pytest.mark.parametrize("n", [1, 2, 3, 17, 33, 64, 100, 193, 256, 257, 512, 768, 1024])
def test_random_symmetric(ctx, n):
    print(f"case: syevd random n={n}", flush=True)
    A = _rand_sym(n, n)
    w, V = _solve(ctx, A)
    _check(A, w, V)


@pytest.mark.parametrize("n", [768, 1024])
def test_own_solver_beyond_512(ctx, n, monkeypatch):
    """Beyond n = 512 the default is rocSOLVER's dsyevd; PW_EIGEN=own keeps
    the blocked single-workgroup tridiagonalisation, checked to the same
    bars."""
    print(f"case: syevd own solver n={n}", flush=True)
    monkeypatch.setenv("PW_EIGEN", "own")
    A = _rand_sym(n, n + 1)
    w, V = _solve(ctx, A)
    _check(A, w, V)


def test_matches_oracle_vectors(ctx):
    print("case: syevd vs oracle n=200", flush=True)
    A = _rand_sym(200, 11)
    w, V = _solve(ctx, A)
    wo, Vo = O.syev(A)
    assert np.abs(w - wo).max() < 1e-12
    gaps = np.minimum(np.r_[np.inf, np.diff(wo)], np.r_[np.diff(wo), np.inf])
    ok = gaps > 1e-3
    assert np.abs(V[:, ok] - Vo[:, ok]).max() < 1e-9


@pytest.mark.parametrize("kind", ["degenerate", "near_degenerate", "near_diagonal", "graded"])
def test_hard_spectra(ctx, kind):
    print(f"case: syevd {kind}", flush=True)
    rng = np.random.default_rng(5)
    n = 160
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    if kind == "degenerate":   # triple and six-fold clusters, as in cubic cells
        lam = np.sort(np.r_[np.repeat(-1.0, 3), np.repeat(0.25, 6), np.linspace(-0.5, 2.0, n - 9)])
        A = (Q * lam) @ Q.T
    elif kind == "near_degenerate":
        lam = np.sort(np.r_[-1.0, -1.0 + 1e-10, -1.0 + 2e-10, np.linspace(0, 1, n - 3)])
        A = (Q * lam) @ Q.T
    elif kind == "near_diagonal":   # a warm Rayleigh-Ritz matrix
        A = np.diag(np.linspace(-1.0, 3.0, n)) + 1e-5 * _rand_sym(n, 3)
    else:
        lam = np.sort(np.r_[np.logspace(-12, 0, n // 2), -np.logspace(-10, 1, n - n // 2)])
        A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    w, V = _solve(ctx, A)
    _check(A, w, V)


def test_reads_upper_triangle_only(ctx):
    print("case: syevd upper triangle", flush=True)
    A = _rand_sym(96, 21)
    B = A.copy()
    # column-major element (i, j), i > j (strictly lower) sits at torch [j, i]
    iu = np.triu_indices(96, 1)
    B[iu] = 1e3
    w, V = _solve(ctx, B)
    wa, Va = _solve(ctx, A)
    assert np.array_equal(w, wa)
    assert np.array_equal(V, Va)


@pytest.mark.parametrize("n,nbatch", [(2, 3), (17, 5), (64, 8), (96, 8), (100, 3), (256, 2)])
def test_sygvd_batched(ctx, n, nbatch):
    """pw_sygvd_batched (lapack::Sygvd itype 1, 'V', 'U' on a batch, as the
    PPCG sub-block problems): eigenvalues against scipy.linalg.eigh(A, B)
    1e-11 relative to ||A||, x^T B x = I and A x = lambda B x to 1e-10,
    only the upper triangles read (garbage planted below the diagonal)."""
    import scipy.linalg as sl
    import torch
    print(f"case: sygvd batched n={n} nbatch={nbatch}", flush=True)
    rng = np.random.default_rng(100 + n)
    As, Bs = [], []
    for b in range(nbatch):
        A = _rand_sym(n, 1000 * n + b)
        G = rng.standard_normal((n, n))
        Bm = G @ G.T / n + np.eye(n)
        As.append(A)
        Bs.append(Bm)
    # torch row-major [b, j, i] = column-major element (i, j); plant values
    # in the strictly lower triangle (i > j), which must not be read
    At = np.stack([a.T.copy() for a in As])
    Bt = np.stack([bm.T.copy() for bm in Bs])
    il = np.tril_indices(n, -1)
    for arr in (At, Bt):
        for b in range(nbatch):
            arr[b][il[1], il[0]] = 7.0
    w, Z, U, info = ctx.sygvd_batched(torch.from_numpy(At).to(ctx.device), torch.from_numpy(Bt).to(ctx.device))
    w, Z, info = w.cpu().numpy(), Z.cpu().numpy(), info.cpu().numpy()
    assert np.all(info == 0)
    for b in range(nbatch):
        A, Bm = As[b], Bs[b]
        ref = sl.eigh(A, Bm, eigvals_only=True)
        scale = max(1.0, np.abs(A).max())
        assert np.abs(w[b] - ref).max() < 1e-11 * scale * max(1, n / 64)
        X = Z[b].T  # columns = eigenvectors
        assert np.abs(X.T @ Bm @ X - np.eye(n)).max() < 1e-10 * max(1, n / 64)
        assert np.abs(A @ X - Bm @ X * w[b]).max() < 1e-10 * scale * max(1, n / 64)


def test_sygvd_batched_reports_indefinite_b(ctx):
    """A B that is not positive definite in one problem of the batch:
    info[b] = the failing pivot column (1-based), 0 for the others."""
    import torch
    n = 40
    As = np.stack([_rand_sym(n, 5 + b).T.copy() for b in range(3)])
    Bs = np.stack([np.eye(n) for _ in range(3)])
    Bs[1][12, 12] = -1.0  # column 13 (1-based) fails
    w, Z, U, info = ctx.sygvd_batched(torch.from_numpy(As).to(ctx.device), torch.from_numpy(Bs).to(ctx.device))
    assert list(info.cpu().numpy()) == [0, 13, 0]
