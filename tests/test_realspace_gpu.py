"""Real-space dual-grid H*psi on the GPU (pw_rs_*) against the oracle
(oracle/or_rs.c, pinned by test_realspace_oracle.py).

FP64 throughout, tolerance max|gpu - oracle| <= 1e-12 max|oracle| (the FFTs
and the projector sums differ only in rounding order).  At the full SiC_large
size the oracle checks one band, and linearity H(psi M) = H(psi) M checks
every band and the batch remainder.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from pwdft_develop_amd import _lib
from pwdft_develop_amd import realspace as R

pytestmark = pytest.mark.gpu

TOL = 1e-12
_CACHE = {}


def _prob(name, nb):
    key = (name, nb)
    if key not in _CACHE:
        _CACHE[key] = R.RS_WORKLOADS[name](nb=nb).build()
    return _CACHE[key]


def _oracle(p, psi, nonlocal_=True):
    if nonlocal_:
        return O.rs_mult_spinor(p.grid, p.grid_fine, p.volume, p.ekin, p.vtot, p.npts, p.nproj, p.idx,
                                p.val, p.wgt, psi)
    return O.rs_mult_spinor(p.grid, p.grid_fine, p.volume, p.ekin, p.vtot, [], [], [], [], [], psi)


def _gpu(p, psi, max_bands=None, nonlocal_=True):
    import torch
    ctx = R.RealSpaceContext(p, max_bands=max_bands, nonlocal_=nonlocal_)
    x = torch.from_numpy(np.ascontiguousarray(psi)).to("cuda:0")
    keep = x.clone()
    y = ctx.mult_spinor(x)
    torch.cuda.synchronize()
    assert torch.equal(x, keep), "psi was modified"
    out = y.cpu().numpy()
    ctx.close()
    return out


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("name,nb,max_bands", [
    ("box_rs", 6, None),
    ("box_rs", 7, 3),        # batches of 3, 3 and 1
    ("sic8_rs", 16, None),
    ("sic8_rs", 5, 2),
])
def test_rs_mult_spinor_matches_oracle(name, nb, max_bands):
    print(f"case {name} nb={nb} max_bands={max_bands}", flush=True)
    p = _prob(name, nb)
    ref = _oracle(p, p.psi)
    got = _gpu(p, p.psi, max_bands=max_bands)
    assert _rel(got, ref) < TOL


def test_rs_unsupported_length_is_refused():
    """A grid length with a prime factor above 31 (74 = 2 x 37) is refused
    with PW_ERR_UNSUPPORTED: there is no second FFT backend to fall to."""
    import ctypes as C
    from pwdft_develop_amd import _lib
    h = C.c_void_p()
    grid = np.array([74, 24, 24], np.int32)
    ekin = np.zeros((74 // 2 + 1) * 24 * 24)
    rc = _lib.lib().pw_rs_create(C.byref(h), grid, grid, 1000.0, ekin, 4, None)
    assert rc == -7 and b"prime" in _lib.lib().pw_last_error()


@pytest.mark.parametrize("name,cutoff", [("sic8_rs", 12.0), ("box_rs", 8.0)])
def test_rs_filter_against_rocfft(name, cutoff):
    """Cross-check of the own coarse transforms against rocFFT (torch.fft on
    the device, FP64): the wavefunction filter x <- C2R(mask . R2C(x)) with
    mask = [ekin <= cutoff] on the R2C half grid, 1e-12 relative."""
    import torch
    p = _prob(name, 4)
    n1, n2, n3 = p.grid
    ctx = R.RealSpaceContext(p)
    x = torch.from_numpy(p.psi.copy()).to("cuda:0")
    got = ctx.filter(x, cutoff)
    mask = torch.from_numpy((p.ekin <= cutoff).reshape(n3, n2, n1 // 2 + 1)).to("cuda:0")
    xr = torch.from_numpy(p.psi).to("cuda:0").reshape(-1, n3, n2, n1)
    ref = torch.fft.irfftn(torch.fft.rfftn(xr, dim=(1, 2, 3)) * mask, s=(n3, n2, n1), dim=(1, 2, 3))
    ctx.close()
    torch.cuda.synchronize()
    assert _rel(got.cpu().numpy(), ref.reshape(got.shape).cpu().numpy()) < 1e-12


def test_rs_local_terms_only():
    p = _prob("sic8_rs", 4)
    ref = _oracle(p, p.psi, nonlocal_=False)
    got = _gpu(p, p.psi, nonlocal_=False)
    assert _rel(got, ref) < TOL


def test_rs_repeated_points_in_one_atom():
    """An atom whose sphere lists grid points twice (a cell smaller than the
    projector box, periodtable.cpp:1063-1101 wraps without a check): the
    point-major kernel takes over from the colour-batched one."""
    p = _prob("box_rs", 5)
    offs = np.concatenate([[0], np.cumsum(p.npts)])
    voffs = np.concatenate([[0], np.cumsum(p.npts.astype(np.int64) * p.nproj)])
    idx, val, npts = [], [], p.npts.copy()
    for a in range(len(p.npts)):
        ia = p.idx[offs[a]:offs[a + 1]]
        va = p.val[voffs[a]:voffs[a + 1]].reshape(p.npts[a], p.nproj[a], order="F")
        if a == 0:
            ia = np.concatenate([ia, ia[:5]])
            va = np.vstack([va, 0.5 * va[:5]])
            npts[a] += 5
        idx.append(ia)
        val.append(va.reshape(-1, order="F"))
    idx, val = np.concatenate(idx).astype(np.int32), np.concatenate(val)
    ref = O.rs_mult_spinor(p.grid, p.grid_fine, p.volume, p.ekin, p.vtot, npts, p.nproj, idx, val, p.wgt, p.psi)
    import torch
    ctx = R.RealSpaceContext(p, nonlocal_=False)
    ctx.set_nonlocal(npts, p.nproj, idx, val, p.wgt)
    got = ctx.mult_spinor(torch.from_numpy(p.psi).to("cuda:0")).cpu().numpy()
    ctx.close()
    assert _rel(got, ref) < TOL


def test_rs_sic_large_one_band_matches_oracle():
    print("case sic216_rs one band", flush=True)
    p = _prob("sic216_rs", 2)
    ref = _oracle(p, p.psi[:1])
    got = _gpu(p, p.psi[:1])
    assert _rel(got, ref) < TOL


def test_rs_mos2_large_shape_matches_oracle():
    """The MoS2_large deck's grids: 88 x 102 x 80 coarse (primes 11 and 17),
    176 x 204 x 160 fine."""
    print("case mos2_rs two bands", flush=True)
    p = _prob("mos2_rs", 2)
    ref = _oracle(p, p.psi)
    got = _gpu(p, p.psi)
    assert _rel(got, ref) < TOL


def test_rs_sic_large_all_bands_linear():
    import torch
    print("case sic216_rs 442 bands linearity", flush=True)
    p = _prob("sic216_rs", 442)
    ctx = R.RealSpaceContext(p)
    assert ctx.info()["chunk"] < 442          # several batches and a remainder
    x = torch.from_numpy(p.psi).to("cuda:0")
    hx = ctx.mult_spinor(x)
    M = torch.from_numpy(np.random.default_rng(5).standard_normal((442, 3))).to("cuda:0")
    y = (x.T @ M).T.contiguous()              # three combinations of all bands
    hy = ctx.mult_spinor(y)
    want = (hx.T @ M).T
    torch.cuda.synchronize()
    assert (hy - want).abs().max().item() < 1e-11 * want.abs().max().item()
    # symmetry of H on two bands: <x0, H x1> = <H x0, x1>
    a = torch.dot(x[0], hx[1]).item()
    b = torch.dot(hx[0], x[1]).item()
    scale = (x[0].norm() * hx[1].norm() + hx[0].norm() * x[1].norm()).item()
    assert abs(a - b) < 1e-12 * scale
    ctx.close()


@pytest.mark.parametrize("name,cutoff", [("sic8_rs", 12.0), ("box_rs", 8.0)])
def test_rs_wavefunction_filter(name, cutoff):
    import torch
    p = _prob(name, 5)
    ref = O.rs_filter(p.grid, p.volume, p.ekin, cutoff, p.psi)
    ctx = R.RealSpaceContext(p)
    x = torch.from_numpy(p.psi.copy()).to("cuda:0")
    got = ctx.filter(x, cutoff).cpu().numpy()
    ctx.close()
    assert _rel(got, ref) < TOL


@pytest.mark.parametrize("name", ["sic8_rs", "box_rs"])
def test_rs_calculate_density(name):
    import torch
    p = _prob(name, 5)
    occ = np.array([1.0, 1.0, 0.9, 0.4, 0.1])
    rho_ref, val_ref = O.rs_density(p.grid, p.grid_fine, p.volume, p.psi, occ, 2, 5)
    ctx = R.RealSpaceContext(p, max_bands=2)     # three batches
    rho, val = ctx.calculate_density(torch.from_numpy(p.psi).to("cuda:0"), occ, 2, 5)
    got = rho.cpu().numpy()
    ctx.close()
    assert _rel(got, rho_ref) < TOL
    assert val == pytest.approx(val_ref, rel=1e-13) and val == pytest.approx(10.0, rel=1e-12)


@pytest.mark.parametrize("name", ["sic8_rs", "box_rs"])
def test_rs_teter_precond(name):
    import torch
    p = _prob(name, 5)
    a3 = np.random.default_rng(2).standard_normal(p.psi.shape)
    ref = O.rs_teter_precond(p.grid, p.volume, R.teter_of(p.ekin), p.psi, a3)
    ctx = R.RealSpaceContext(p, max_bands=2)
    y = torch.from_numpy(a3.copy()).to("cuda:0")
    ctx.teter_precond(torch.from_numpy(p.psi).to("cuda:0"), y)
    got = y.cpu().numpy()
    ctx.close()
    assert _rel(got, ref) < TOL


@pytest.mark.parametrize("name,nb,sb,iters", [("box_rs", 8, 4, 4), ("sic8_rs", 12, 4, 3), ("sic8_rs", 10, 4, 3)])
def test_rs_ppcg_matches_oracle(name, nb, sb, iters):
    """pw_rs_ppcg against the oracle's real-space PPCG, fixed iteration count
    (ragged last sub-block for nb = 10)."""
    import torch
    print(f"case rs ppcg {name} nb={nb} sb={sb} iters={iters}", flush=True)
    p = _prob(name, nb)
    Xo, eo, ro, io = O.rs_ppcg(p, p.psi, nb - 2, 1, iters, 1e-30, sb)
    ctx = R.RealSpaceContext(p)
    X = torch.from_numpy(p.psi.copy()).to("cuda:0")
    e, r, it = ctx.ppcg(X, nb - 2, 1, iters, 1e-30, sb)
    ctx.close()
    assert it == io == iters
    assert np.abs(e - eo).max() < 1e-10 * max(1.0, np.abs(eo).max())
    assert np.abs(r - ro).max() < 1e-6 * max(1e-3, ro.max())


def test_rs_ppcg_converges_to_dense():
    import torch
    p = _prob("box_rs", 8)
    H = O.rs_mult_spinor(p.grid, p.grid_fine, p.volume, p.ekin, p.vtot, p.npts, p.nproj, p.idx, p.val, p.wgt,
                         np.eye(p.ntot))
    ev = np.linalg.eigvalsh(0.5 * (H + H.T))[:8]
    ctx = R.RealSpaceContext(p)
    X = torch.from_numpy(p.psi.copy()).to("cuda:0")
    e, r, it = ctx.ppcg(X, 6, 1, 200, 1e-9, 4)
    ctx.close()
    assert it < 200
    assert np.abs(e - ev).max() < 1e-9
    assert r[:6].max() < 1e-8


def test_rs_errors():
    import torch  # noqa: F401  (the HIP runtime)
    L = _lib.lib()
    p = _prob("box_rs", 6)
    h = C.c_void_p()
    rc = L.pw_rs_create(C.byref(h), np.array([12, 15, 10], np.int32), np.array([18, 32, 20], np.int32),
                        float(p.volume), np.ascontiguousarray(p.ekin), 4, None)
    assert rc == -7 and "even" in L.pw_last_error().decode()
    rc = L.pw_rs_create(C.byref(h), np.array([12, 16, 10], np.int32), np.array([10, 32, 20], np.int32),
                        float(p.volume), np.ascontiguousarray(p.ekin), 4, None)
    assert rc == -1
    assert L.pw_rs_create(C.byref(h), np.array(p.grid, np.int32), np.array(p.grid_fine, np.int32),
                          float(p.volume), np.ascontiguousarray(p.ekin), 4, None) == 0
    try:
        import torch
        x = torch.zeros((1, p.ntot), dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        rc = L.pw_rs_mult_spinor(h, 1, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()))
        assert rc == -4 and "potential" in L.pw_last_error().decode()
        two, one, n33 = np.array([2], np.int32), np.array([1], np.int32), np.array([33], np.int32)
        bad, zero2 = np.array([0, p.ntot_fine], np.int32), np.zeros(2, np.int32)
        v2, v66, w1, w33 = np.ones(2), np.ones(66), np.ones(1), np.ones(33)
        rc = L.pw_rs_set_nonlocal(h, 1, two.ctypes.data, one.ctypes.data, bad.ctypes.data, v2.ctypes.data,
                                  w1.ctypes.data)
        assert rc == -3
        rc = L.pw_rs_set_nonlocal(h, 1, two.ctypes.data, n33.ctypes.data, zero2.ctypes.data, v66.ctypes.data,
                                  w33.ctypes.data)
        assert rc == -7
    finally:
        L.pw_rs_destroy(h)
