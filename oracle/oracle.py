"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, never by the product package.  Arrays are
numpy, complex blocks are complex128 (npw x nb, column-major, i.e. numpy
arrays of shape (nb, npw) in C order).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_sphere_build.restype = C.c_int
        L.or_sphere_build.argtypes = [_ip, _dp, C.c_double, _ip, _dp, C.POINTER(C.c_int), _ip, _ip]
        L.or_fock_sphere_build.argtypes = [_ip, _dp, C.c_double, _ip, _dp]
        L.or_ekin_teter.argtypes = [C.c_int, _dp, _dp, _dp]
        L.or_ekin_teter.restype = None
        L.or_fft1d.argtypes = [C.c_int, C.c_int, _dp]
        L.or_fft_c2r.argtypes = [_ip, _dp, _dp]
        L.or_fft_r2c.argtypes = [_ip, _dp, _dp]
        L.or_ham_create.restype = C.c_void_p
        L.or_ham_create.argtypes = [_ip, _dp, C.c_int, _ip, C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.or_ham_destroy.argtypes = [C.c_void_p]
        L.or_ham_destroy.restype = None
        L.or_ham_set_nonlocal.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.or_ham_set_exx.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int,
                                     C.c_int, _dp, _dp, C.POINTER(C.c_double)]
        L.or_ham_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.or_ham_build_ace.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_ppcg.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                              C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.or_sygvd.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int, _dp]
        for f in ("or_mult_spinor", "or_add_exx", "or_nonlocal_multx", "or_apply_h",
                  "or_teter_precond"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.or_gemm_tn.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.or_gemm_tn.restype = None
        L.or_potrf_upper.argtypes = [C.c_int, _dp, C.c_int]
        L.or_syev.argtypes = [C.c_int, _dp, C.c_int, _dp]
        L.or_solve_gen_eig.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.or_orthogonalize.argtypes = [C.c_int, C.c_int, _dp]
        L.or_lobpcg.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                C.c_double, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.or_lanczos.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.POINTER(C.c_double)]
        L.or_cheb_filter.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, _dp]
        L.or_rayleigh_ritz.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.or_chebfsi_sweep.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, C.c_void_p,
                                       _dp, _dp, _dp, _dp]
        L.or_density.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_double, _dp]
        L.or_band_energy.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.or_band_energy.restype = C.c_double
        L.or_nonlocal_force.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp]
        L.or_ip_c2f.argtypes = [_ip, _ip, _dp, _dp]
        L.or_ip_c2f.restype = None
        L.or_ip_f2c.argtypes = [_ip, _ip, _dp, _dp]
        L.or_ip_f2c.restype = None
        L.or_rs_mult_spinor.argtypes = [_ip, _ip, C.c_double, _dp, _dp, C.c_int, _ip, _ip, _ip, _dp,
                                        _dp, C.c_int, _dp, _dp]
        L.or_rs_filter.argtypes = [_ip, C.c_double, _dp, C.c_double, C.c_int, _dp]
        L.or_rs_teter_precond.argtypes = [_ip, C.c_double, _dp, C.c_int, _dp, _dp]
        L.or_rs_ppcg.argtypes = [_ip, _ip, C.c_double, _dp, _dp, C.c_int, _ip, _ip, _ip, _dp, _dp, C.c_int,
                                 C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp,
                                 C.POINTER(C.c_int)]
        L.or_rs_density.argtypes = [_ip, _ip, C.c_double, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp,
                                    C.POINTER(C.c_double)]
        L.or_harris_energy.argtypes = [C.c_int, C.c_int, C.c_double, _dp, C.c_void_p, _dp, _dp, _dp, C.c_int,
                                       _dp, _dp, C.c_int, C.c_int, _dp, C.POINTER(C.c_double)]
        _LIB = L
    return _LIB


HARRIS_TERMS = ("Exc", "Eself", "EIonSR", "EVdw", "Eext", "fermi", "Tbeta", "temperature")


def harris_energy(density, vxc, vhart, pseudo, vol, eig, occ, num_spin, num_occ, terms, spindensity=None):
    """SCF::CalculateHarrisEnergy (scf.cpp:1987-2140).  density, vxc:
    (ncomp, ntot) arrays (component-major: DblNumMat(ntot, ncomp) storage);
    spindensity (2, ntot) for nspin = 2; terms: dict with HARRIS_TERMS."""
    density = np.ascontiguousarray(density, dtype=np.float64)
    vxc = np.ascontiguousarray(vxc, dtype=np.float64)
    nspin, ntot = vxc.shape
    sd = None
    if spindensity is not None:
        sd = np.ascontiguousarray(spindensity, dtype=np.float64)
    t = np.array([float(terms[k]) for k in HARRIS_TERMS])
    e = C.c_double(0.0)
    _chk(lib().or_harris_energy(ntot, nspin, vol, density.reshape(-1),
                                sd.ctypes.data_as(C.c_void_p) if sd is not None else None,
                                vxc.reshape(-1), np.ascontiguousarray(vhart, dtype=np.float64),
                                np.ascontiguousarray(pseudo, dtype=np.float64), len(eig),
                                np.ascontiguousarray(eig, dtype=np.float64),
                                np.ascontiguousarray(occ, dtype=np.float64), int(num_spin), int(num_occ), t,
                                C.byref(e)), "harris_energy")
    return e.value


def _r(a):
    """complex block -> contiguous float64 view"""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64).reshape(-1)


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def sphere(n, supercell, ecut):
    n = np.asarray(n, dtype=np.int32)
    cap = (n[0] // 2 + 1) * n[1] * n[2]
    idx = np.zeros(cap, np.int32)
    gkk = np.zeros(cap)
    src = np.zeros(cap, np.int32)
    dst = np.zeros(cap, np.int32)
    npad = C.c_int(0)
    npw = lib().or_sphere_build(n, np.ascontiguousarray(supercell, dtype=np.float64).reshape(-1),
                                ecut, idx, gkk, C.byref(npad), src, dst)
    if npw <= 0:
        raise RuntimeError("sphere build failed")
    return idx[:npw].copy(), gkk[:npw].copy(), src[:npad.value].copy(), dst[:npad.value].copy()


def fock_sphere(n, supercell, ecut):
    n = np.asarray(n, dtype=np.int32)
    cap = (n[0] // 2 + 1) * n[1] * n[2]
    idx = np.zeros(cap, np.int32)
    gkk = np.zeros(cap)
    cnt = lib().or_fock_sphere_build(n, np.ascontiguousarray(supercell, dtype=np.float64).reshape(-1),
                                     ecut, idx, gkk)
    return idx[:cnt].copy(), gkk[:cnt].copy()


def ekin_teter(gkk):
    gkk = np.ascontiguousarray(gkk, dtype=np.float64)
    e = np.zeros_like(gkk)
    t = np.zeros_like(gkk)
    lib().or_ekin_teter(len(gkk), gkk, e, t)
    return e, t


def fft1d(x, sign):
    x = np.array(x, dtype=np.complex128)
    buf = _r(x).copy()
    _chk(lib().or_fft1d(len(x), sign, buf), "fft1d")
    return buf.view(np.complex128)


def fft_c2r(n, half):
    """half: complex array of shape (n3, n2, h1) -> real (n3, n2, n1)"""
    n = np.asarray(n, dtype=np.int32)
    out = np.zeros(int(n[0]) * int(n[1]) * int(n[2]))
    _chk(lib().or_fft_c2r(n, _r(half), out), "c2r")
    return out.reshape(int(n[2]), int(n[1]), int(n[0]))


def fft_r2c(n, real):
    n = np.asarray(n, dtype=np.int32)
    h1 = int(n[0]) // 2 + 1
    out = np.zeros(2 * h1 * int(n[1]) * int(n[2]))
    _chk(lib().or_fft_r2c(n, np.ascontiguousarray(real, dtype=np.float64).reshape(-1), out), "r2c")
    return out.view(np.complex128).reshape(int(n[2]), int(n[1]), h1)


class Hamiltonian:
    """Oracle Hamiltonian over one problem (grid, sphere, V, projectors, EXX)."""

    def __init__(self, prob, threads=0):
        L = lib()
        self.prob = prob
        self.npw = prob.npw
        self.h = L.or_ham_create(np.asarray(prob.grid, np.int32), prob.supercell.reshape(-1).copy(),
                                 prob.npw, prob.idx, len(prob.pad_src), prob.pad_src, prob.pad_dst,
                                 prob.ekin, prob.teter, prob.vtot.reshape(-1).copy())
        if threads:
            L.or_ham_set_threads(self.h, threads)
        if prob.nbeta > 0:
            _chk(L.or_ham_set_nonlocal(self.h, prob.nbeta, _r(prob.beta), prob.wgt), "nonlocal")

    def __del__(self):
        try:
            if self.h:
                lib().or_ham_destroy(self.h)
        except Exception:
            pass

    def set_threads(self, t):
        lib().or_ham_set_threads(self.h, int(t))

    def set_exx(self, ecut, mu, frac, psi_occ, occ, divtype=1):
        d = C.c_double(0.0)
        psi_occ = np.ascontiguousarray(psi_occ, dtype=np.complex128)
        _chk(lib().or_ham_set_exx(self.h, ecut, mu, frac, divtype, psi_occ.shape[0], _r(psi_occ),
                                  np.ascontiguousarray(occ, dtype=np.float64), C.byref(d)), "set_exx")
        return d.value

    def build_ace(self, psi):
        """KohnSham::CalculateVexxACE (hamiltonian.cpp:3685-3924) from the
        block psi; afterwards mult_spinor/apply_h use the compressed exchange.
        An empty block drops it."""
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        buf = _r(psi) if psi.size else np.zeros(1)
        _chk(lib().or_ham_build_ace(self.h, psi.shape[0], buf), "build_ace")

    def _op(self, name, psi):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        out = np.zeros_like(psi)
        o = _r(out)
        _chk(getattr(lib(), name)(self.h, psi.shape[0], _r(psi), o), name)
        return o.view(np.complex128).reshape(psi.shape)

    def mult_spinor(self, psi):
        return self._op("or_mult_spinor", psi)

    def nonlocal_multx(self, psi):
        return self._op("or_nonlocal_multx", psi)

    def apply_h(self, psi):
        return self._op("or_apply_h", psi)

    def add_exx(self, psi):
        return self._op("or_add_exx", psi)

    def teter(self, psi):
        return self._op("or_teter_precond", psi)

    def lobpcg(self, psi, num_eig=None, scf_iter=1, max_iter=3, tol=1e-20):
        psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
        nb = psi.shape[0]
        ev = np.zeros(nb)
        rv = np.zeros(nb)
        it = C.c_int(0)
        buf = _r(psi)
        _chk(lib().or_lobpcg(self.h, nb, num_eig or nb, scf_iter, max_iter, 0.0, tol, buf, ev, rv,
                             C.byref(it)), "lobpcg")
        return buf.view(np.complex128).reshape(psi.shape), ev, rv, it.value

    def ppcg(self, psi, num_eig=None, scf_iter=1, max_iter=3, tol=1e-20, sb_size=32):
        """EigenSolver::PPCGSolveReal (ppcg.cpp:119-576)."""
        psi = np.ascontiguousarray(psi, dtype=np.complex128).copy()
        nb = psi.shape[0]
        ev = np.zeros(nb)
        rv = np.zeros(nb)
        it = C.c_int(0)
        buf = _r(psi)
        _chk(lib().or_ppcg(self.h, nb, num_eig or nb, scf_iter, max_iter, 0.0, tol, sb_size, buf, ev, rv,
                           C.byref(it)), "ppcg")
        return buf.view(np.complex128).reshape(psi.shape), ev, rv, it.value

    def lanczos(self, k, v0):
        th = np.zeros(k)
        b = C.c_double(0.0)
        _chk(lib().or_lanczos(self.h, k, _r(v0), th, C.byref(b)), "lanczos")
        return th, b.value

    def cheb_filter(self, X, m, a, b, aL):
        buf = _r(np.array(X, dtype=np.complex128)).copy()
        _chk(lib().or_cheb_filter(self.h, X.shape[0], m, a, b, aL, buf), "cheb_filter")
        return buf.view(np.complex128).reshape(X.shape)

    def rayleigh_ritz(self, X):
        buf = _r(np.array(X, dtype=np.complex128)).copy()
        nb = X.shape[0]
        e = np.zeros(nb)
        r = np.zeros(nb)
        _chk(lib().or_rayleigh_ritz(self.h, nb, buf, e, r), "rayleigh_ritz")
        return buf.view(np.complex128).reshape(X.shape), e, r

    def chebfsi_sweep(self, X, degree, k, v0, prev_eig=None):
        buf = _r(np.array(X, dtype=np.complex128)).copy()
        nb = X.shape[0]
        e = np.zeros(nb)
        r = np.zeros(nb)
        bounds = np.zeros(3)
        pe = None
        if prev_eig is not None:
            pe_arr = np.ascontiguousarray(prev_eig, dtype=np.float64)
            pe = pe_arr.ctypes.data_as(C.c_void_p)
        _chk(lib().or_chebfsi_sweep(self.h, nb, degree, k, _r(v0), pe, buf, e, r, bounds),
             "chebfsi_sweep")
        return buf.view(np.complex128).reshape(X.shape), e, r, bounds

    def density(self, psi, occ, num_spin=1.0):
        rho = np.zeros(int(np.prod(self.prob.grid)))
        _chk(lib().or_density(self.h, psi.shape[0], _r(psi), np.ascontiguousarray(occ, np.float64),
                              num_spin, rho), "density")
        return rho

    def band_energy(self, psi, f):
        return lib().or_band_energy(self.h, psi.shape[0], _r(psi), np.ascontiguousarray(f, np.float64))

    def nonlocal_force(self, psi, f, natom, npa):
        out = np.zeros(3 * natom)
        _chk(lib().or_nonlocal_force(self.h, psi.shape[0], _r(psi), np.ascontiguousarray(f, np.float64),
                                     natom, npa, out), "nonlocal_force")
        return out.reshape(natom, 3)


def gemm_tn(A, B):
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    m, n1 = A.shape
    n2 = B.shape[1]
    Cm = np.zeros(n1 * n2)
    lib().or_gemm_tn(m, n1, n2, A.reshape(-1, order="F").copy(), B.reshape(-1, order="F").copy(), Cm)
    return Cm.reshape(n2, n1).T


def syev(A):
    A = np.array(A, dtype=np.float64)
    n = A.shape[0]
    buf = A.reshape(-1, order="F").copy()
    w = np.zeros(n)
    _chk(lib().or_syev(n, buf, n, w), "syev")
    return w, buf.reshape(n, n, order="F")


def sygvd(A, B):
    """lapack::Sygvd(1, 'V', 'U'): (eigenvalues, B-orthonormal eigenvectors)."""
    A = np.array(A, dtype=np.float64)
    n = A.shape[0]
    a = A.reshape(-1, order="F").copy()
    b = np.array(B, dtype=np.float64).reshape(-1, order="F").copy()
    w = np.zeros(n)
    _chk(lib().or_sygvd(n, a, n, b, n, w), "sygvd")
    return w, a.reshape(n, n, order="F")


def potrf_upper(A):
    A = np.array(A, dtype=np.float64)
    n = A.shape[0]
    buf = A.reshape(-1, order="F").copy()
    info = lib().or_potrf_upper(n, buf, n)
    return info, np.triu(buf.reshape(n, n, order="F"))


def solve_gen_eig(A, B, width):
    n = A.shape[0]
    a = np.array(A, dtype=np.float64).reshape(-1, order="F").copy()
    b = np.array(B, dtype=np.float64).reshape(-1, order="F").copy()
    e = np.zeros(n)
    rc = lib().or_solve_gen_eig(n, n, width, a, b, e)
    return rc, e[:width], a.reshape(n, n, order="F")


# ---------------------------------------------------------------- real space
def _half_shape(n):
    return (int(n[2]), int(n[1]), int(n[0]) // 2 + 1)


def ip_c2f(nc, nf, coarse):
    """IP_c2f (utility.cpp:3150-3212): coarse half spectrum (Nc2, Nc1, Nc0/2+1)
    -> fine half spectrum (Nf2, Nf1, Nf0/2+1), zeros where nothing lands."""
    nc, nf = np.asarray(nc, np.int32), np.asarray(nf, np.int32)
    out = np.zeros(_half_shape(nf), np.complex128)
    src = np.ascontiguousarray(coarse, np.complex128)
    assert src.shape == _half_shape(nc)
    ov = out.reshape(-1).view(np.float64)
    lib().or_ip_c2f(nc, nf, src.reshape(-1).view(np.float64), ov)
    return out


def ip_f2c(nc, nf, fine, coarse=None):
    """IP_f2c (utility.cpp:3215-3276): returns coarse + restriction(fine)."""
    nc, nf = np.asarray(nc, np.int32), np.asarray(nf, np.int32)
    out = np.zeros(_half_shape(nc), np.complex128) if coarse is None else \
        np.array(coarse, np.complex128, copy=True)
    src = np.ascontiguousarray(fine, np.complex128)
    assert src.shape == _half_shape(nf)
    lib().or_ip_f2c(nc, nf, out.reshape(-1).view(np.float64), src.reshape(-1).view(np.float64))
    return out


def rs_mult_spinor(nc, nf, vol, ekin, vtot, npts, nproj, idx, val, wgt, psi):
    """Real-space dual-grid a3 = H psi (spinor.cpp:376-525).  psi: (nb, ntot)
    real, band-major; ekin: coarse R2C grid; vtot: fine grid; projectors per
    atom (npts, nproj, idx, val column-major per atom, wgt)."""
    nc, nf = np.asarray(nc, np.int32), np.asarray(nf, np.int32)
    psi = np.ascontiguousarray(psi, np.float64)
    nb = psi.shape[0]
    out = np.zeros_like(psi)
    npts = np.ascontiguousarray(npts, np.int32)
    nproj = np.ascontiguousarray(nproj, np.int32)
    idx = np.ascontiguousarray(idx, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    wgt = np.ascontiguousarray(wgt, np.float64)
    if idx.size == 0:
        idx = np.zeros(1, np.int32)
    if val.size == 0:
        val = np.zeros(1)
    if wgt.size == 0:
        wgt = np.zeros(1)
    _chk(lib().or_rs_mult_spinor(nc, nf, float(vol), np.ascontiguousarray(ekin, np.float64).reshape(-1),
                                 np.ascontiguousarray(vtot, np.float64).reshape(-1), len(npts),
                                 npts if npts.size else np.zeros(1, np.int32),
                                 nproj if nproj.size else np.zeros(1, np.int32), idx, val, wgt, nb,
                                 psi.reshape(-1), out.reshape(-1)), "rs_mult_spinor")
    return out


def rs_filter(nc, vol, gkk, cutoff, psi):
    """Wavefunction ecut filter of the real MultSpinor (hamiltonian.cpp:
    3090-3116, 3240-3265) on (nb, ntot) real bands; returns a filtered copy."""
    nc = np.asarray(nc, np.int32)
    out = np.array(psi, np.float64, copy=True, order="C")
    _chk(lib().or_rs_filter(nc, float(vol), np.ascontiguousarray(gkk, np.float64), float(cutoff),
                            out.shape[0], out.reshape(-1)), "rs_filter")
    return out


def rs_density(nc, nf, vol, psi, occ, num_spin, num_occ):
    """Real-space CalculateDensity (hamiltonian.cpp:2453-2627): (rho on the
    fine grid, val)."""
    nc, nf = np.asarray(nc, np.int32), np.asarray(nf, np.int32)
    psi = np.ascontiguousarray(psi, np.float64)
    rho = np.zeros(int(np.prod(nf)))
    val = C.c_double(0.0)
    _chk(lib().or_rs_density(nc, nf, float(vol), psi.shape[0], psi.reshape(-1),
                             np.ascontiguousarray(occ, np.float64), int(num_spin), int(num_occ), rho,
                             C.byref(val)), "rs_density")
    return rho, val.value


def rs_teter_precond(nc, vol, teter, psi, a3):
    """Real Spinor::AddTeterPrecond (spinor.cpp:309-345): returns a3 + T psi."""
    nc = np.asarray(nc, np.int32)
    psi = np.ascontiguousarray(psi, np.float64)
    out = np.array(a3, np.float64, copy=True, order="C")
    _chk(lib().or_rs_teter_precond(nc, float(vol), np.ascontiguousarray(teter, np.float64), psi.shape[0],
                                   psi.reshape(-1), out.reshape(-1)), "rs_teter_precond")
    return out


def rs_ppcg(p, psi, num_eig, scf_iter, max_iter, tol, sb):
    """Real-space PPCGSolveReal (ppcg.cpp:683-1246) on a RealSpaceProblem's
    arrays: (X, eigenvalues, residual norms, iterations)."""
    nc, nf = np.asarray(p.grid, np.int32), np.asarray(p.grid_fine, np.int32)
    X = np.array(psi, np.float64, copy=True, order="C")
    nb = X.shape[0]
    ev, rv, it = np.zeros(nb), np.zeros(nb), C.c_int(0)
    _chk(lib().or_rs_ppcg(nc, nf, float(p.volume), np.ascontiguousarray(p.ekin, np.float64),
                          np.ascontiguousarray(p.vtot, np.float64), len(p.npts), np.ascontiguousarray(p.npts, np.int32),
                          np.ascontiguousarray(p.nproj, np.int32), np.ascontiguousarray(p.idx, np.int32),
                          np.ascontiguousarray(p.val, np.float64), np.ascontiguousarray(p.wgt, np.float64), nb,
                          int(num_eig), int(scf_iter), int(max_iter), float(tol), int(sb), X.reshape(-1), ev, rv,
                          C.byref(it)), "rs_ppcg")
    return X, ev, rv, it.value
