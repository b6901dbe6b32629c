"""The C++ adapter (reference class interfaces over the C-ABI) driven from a
C++ program: H*psi, LOBPCG, a GeneralChebyStep and PPCGSolveReal (on a
solver set up through the Hamiltonian base class) against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from pwdft_develop_amd import realspace as R
from pwdft_develop_amd import workload as W

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "adapter_driver")


def test_adapter_driver(tmp_path):
    if not os.path.exists(DRIVER):
        raise RuntimeError("tests/cpp/adapter_driver is missing: run python -m pwdft_develop_amd.build")
    prob = W.si8_32(nb=16).build()
    v0 = prob.lanczos_start()
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as fh:
        np.array(list(prob.grid) + [prob.nb, prob.nbeta, prob.npw], np.int32).tofile(fh)
        np.concatenate([prob.supercell.reshape(-1), [prob.ecut]]).astype(np.float64).tofile(fh)
        prob.vtot.astype(np.float64).tofile(fh)
        np.ascontiguousarray(prob.beta).view(np.float64).tofile(fh)
        prob.wgt.astype(np.float64).tofile(fh)
        np.ascontiguousarray(prob.psi).view(np.float64).tofile(fh)
        np.ascontiguousarray(v0).view(np.float64).tofile(fh)
    out = tmp_path / "out.bin"
    print("case: adapter driver si8_32 nb=16", flush=True)
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = np.fromfile(out, np.float64)
    n = 2 * prob.npw * prob.nb
    hpsi = res[:n].view(np.complex128).reshape(prob.nb, prob.npw)
    e_lob = res[n:n + prob.nb]
    e_cheb = res[n + prob.nb:n + 2 * prob.nb]
    e_ppcg = res[n + 2 * prob.nb:]
    H = O.Hamiltonian(prob)
    ref = H.apply_h(prob.psi)
    assert np.abs(hpsi - ref).max() < 1e-11 * np.abs(ref).max()
    Xo, eo, ro, _ = H.lobpcg(prob.psi, max_iter=3)
    assert np.abs(e_lob - eo).max() < 1e-10
    _, ec, _, _ = H.chebfsi_sweep(Xo, 10, 8, v0, eo)
    assert np.abs(e_cheb - ec).max() < 1e-9
    _, ep, _, _ = H.ppcg(prob.psi, max_iter=3, sb_size=8)
    assert e_ppcg.shape == ep.shape and np.abs(e_ppcg - ep).max() < 1e-10


def test_adapter_driver_reports_errors(tmp_path):
    bad = tmp_path / "bad.bin"
    np.array([8, 8, 8, 2, 0, 5], np.int32).tofile(bad)   # truncated input
    r = subprocess.run([DRIVER, str(bad), str(tmp_path / "o.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "ERROR" in r.stderr


@pytest.mark.parametrize("name,cutoff", [("box_rs", 0.0), ("sic8_rs", 0.0), ("sic8_rs", 12.0)])
def test_adapter_driver_realspace(tmp_path, name, cutoff):
    """Spinor::SetupReal, KohnSham::CalculateEkinR2C / SetVtotFine / SetPseudo
    and the NumTns<Real> MultSpinor overload, from C++, against the oracle;
    with set_wfn_filter on, psi is filtered on the first call only and a3 on
    every call (hamiltonian.cpp:3090-3116, 3240-3265)."""
    p = R.RS_WORKLOADS[name](nb=3).build()
    offs = np.concatenate([[0], np.cumsum(p.npts)])
    voffs = np.concatenate([[0], np.cumsum(p.npts.astype(np.int64) * p.nproj)])
    woffs = np.concatenate([[0], np.cumsum(p.nproj)])
    inp = tmp_path / "rs_in.bin"
    with open(inp, "wb") as fh:
        np.array(list(p.grid) + list(p.grid_fine) + [p.nb, len(p.npts)], np.int32).tofile(fh)
        np.stack([p.npts, p.nproj], axis=1).astype(np.int32).tofile(fh)
        p.supercell.astype(np.float64).reshape(-1).tofile(fh)
        p.vtot.astype(np.float64).tofile(fh)
        for a in range(len(p.npts)):
            p.idx[offs[a]:offs[a + 1]].astype(np.int32).tofile(fh)
            p.val[voffs[a]:voffs[a + 1]].astype(np.float64).tofile(fh)
            p.wgt[woffs[a]:woffs[a + 1]].astype(np.float64).tofile(fh)
        p.psi.astype(np.float64).tofile(fh)
    out = tmp_path / "rs_out.bin"
    print(f"case: adapter driver realspace {name}", flush=True)
    args = [DRIVER, "rs", str(inp), str(out)] + ([repr(cutoff)] if cutoff > 0 else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = np.fromfile(out, np.float64)
    n = p.nb * p.ntot
    a3, a3b = res[:n].reshape(p.nb, p.ntot), res[n:2 * n].reshape(p.nb, p.ntot)
    psi_after, ekin = res[2 * n:3 * n].reshape(p.nb, p.ntot), res[3 * n:]
    assert np.abs(ekin - p.ekin).max() < 1e-12 * p.ekin.max()
    psi_in = p.psi
    if cutoff > 0:
        psi_in = O.rs_filter(p.grid, p.volume, p.ekin, cutoff, p.psi)
    assert np.abs(psi_after - psi_in).max() < 1e-12 * np.abs(psi_in).max()
    ref = O.rs_mult_spinor(p.grid, p.grid_fine, p.volume, p.ekin, p.vtot, p.npts, p.nproj, p.idx, p.val,
                           p.wgt, psi_in)
    if cutoff > 0:
        ref = O.rs_filter(p.grid, p.volume, p.ekin, cutoff, ref)
    assert np.abs(a3 - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(a3b - ref).max() < 1e-12 * np.abs(ref).max()
