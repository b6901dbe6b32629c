"""Python test binding over the C-ABI: a device context holding one problem.

Blocks are torch tensors on the GPU, complex128 (PW_F64) or complex64
(PW_F32), shape (nb, npw) -- i.e. column-major npw x nb, the reference's
NumTns<Complex> layout.  torch only provides device memory and the stream;
every operation is a call into libpwdft_hip.so.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import PW_F32, PW_F64, check


def _dtype_code(t):
    import torch
    if t.dtype == torch.complex128:
        return PW_F64
    if t.dtype == torch.complex64:
        return PW_F32
    raise TypeError(f"unsupported dtype {t.dtype}")


class Context:
    def __init__(self, prob, max_bands=None, stream=None, device="cuda:0"):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("pwdft_develop_amd.ops needs a GPU (no CPU fallback)")
        self.torch = torch
        self.prob = prob
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        L = _lib.lib()
        self.max_bands = int(max_bands or prob.nb)
        h = C.c_void_p()
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        check(L.pw_ctx_create(C.byref(h), np.asarray(prob.grid, np.int32),
                              np.ascontiguousarray(prob.supercell, np.float64).reshape(-1),
                              prob.npw, prob.idx, len(prob.pad_src), np.ascontiguousarray(prob.pad_src),
                              np.ascontiguousarray(prob.pad_dst), prob.ekin, prob.teter,
                              self.max_bands, C.c_void_p(self.stream.cuda_stream)))
        self.h = h
        # positions of the projectors the context holds (set_atoms moves them)
        self.atoms_loaded = np.array(prob.atoms, dtype=np.float64).reshape(-1, 3) if prob.nbeta else None
        check(L.pw_set_vtot(h, np.ascontiguousarray(prob.vtot, np.float64).reshape(-1)))
        if prob.nbeta:
            check(L.pw_set_nonlocal(h, prob.nbeta, np.ascontiguousarray(prob.beta).view(np.float64).reshape(-1),
                                    np.ascontiguousarray(prob.wgt, np.float64)))

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().pw_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_long()
        check(_lib.lib().pw_ctx_info(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"npw": a.value, "nstick": b.value, "kxmax": c.value, "ngrid": d.value}

    def to_device(self, psi, dtype=None):
        t = self.torch
        dt = dtype or t.complex128
        return t.from_numpy(np.ascontiguousarray(psi)).to(device=self.device, dtype=dt).contiguous()

    def _binop(self, name, x, out=None, zero_out=False):
        t = self.torch
        assert x.is_contiguous() and x.dim() == 2 and x.shape[1] == self.prob.npw
        if out is None:
            out = t.zeros_like(x) if zero_out else t.empty_like(x)
        check(getattr(_lib.lib(), name)(self.h, _dtype_code(x), x.shape[0], C.c_void_p(x.data_ptr()),
                                        C.c_void_p(out.data_ptr())))
        return out

    def mult_spinor(self, x, out=None):
        return self._binop("pw_mult_spinor", x, out)

    def nonlocal_multx(self, x, out=None):
        return self._binop("pw_nonlocal_multx", x, out)

    def apply_h(self, x, out=None):
        return self._binop("pw_apply_h", x, out)

    def teter_precond(self, x, out=None):
        return self._binop("pw_teter_precond", x, out, zero_out=True)

    # ---- eigensolver layer (X blocks FP64 on the device, in place) ----
    def orthogonalize(self, X):
        check(_lib.lib().pw_orthogonalize(self.h, X.shape[0], C.c_void_p(X.data_ptr())))
        return X

    def lanczos(self, k, v0):
        th = np.zeros(k)
        b = C.c_double(0.0)
        check(_lib.lib().pw_lanczos(self.h, k, C.c_void_p(v0.data_ptr()), th, C.byref(b)))
        return th, b.value

    def cheb_filter(self, X, m, a, b, aL):
        check(_lib.lib().pw_cheb_filter(self.h, _dtype_code(X), X.shape[0], m, a, b, aL,
                                        C.c_void_p(X.data_ptr())))
        return X

    def rayleigh_ritz(self, X, hdtype=PW_F64):
        nb = X.shape[0]
        e, r = np.zeros(nb), np.zeros(nb)
        check(_lib.lib().pw_rayleigh_ritz(self.h, hdtype, nb, C.c_void_p(X.data_ptr()), e, r))
        return X, e, r

    def chebfsi_sweep(self, X, degree, k, v0, prev_eig=None, mixed=False):
        nb = X.shape[0]
        e, r, bounds = np.zeros(nb), np.zeros(nb), np.zeros(3)
        pe = None
        if prev_eig is not None:
            pe_arr = np.ascontiguousarray(prev_eig, dtype=np.float64)
            pe = C.c_void_p(pe_arr.ctypes.data)
        check(_lib.lib().pw_chebfsi_sweep(self.h, PW_F32 if mixed else PW_F64, nb, degree, k,
                                          C.c_void_p(v0.data_ptr()), pe, C.c_void_p(X.data_ptr()),
                                          e, r, bounds))
        return X, e, r, bounds

    def lobpcg(self, X, num_eig=None, scf_iter=1, max_iter=3, tol=1e-20):
        nb = X.shape[0]
        e, r = np.zeros(nb), np.zeros(nb)
        it = C.c_int(0)
        check(_lib.lib().pw_lobpcg(self.h, nb, num_eig or nb, scf_iter, max_iter, 0.0, tol,
                                   C.c_void_p(X.data_ptr()), e, r, C.byref(it)))
        return X, e, r, it.value

    def ppcg(self, X, num_eig=None, scf_iter=1, max_iter=3, tol=1e-20, sb_size=32):
        nb = X.shape[0]
        e, r = np.zeros(nb), np.zeros(nb)
        it = C.c_int(0)
        check(_lib.lib().pw_ppcg(self.h, nb, num_eig or nb, scf_iter, max_iter, 0.0, tol, sb_size,
                                 C.c_void_p(X.data_ptr()), e, r, C.byref(it)))
        return X, e, r, it.value

    def set_exx(self, ecut, mu, frac, psi_occ, occ, divtype=1):
        """psi_occ: device FP64 block (nocc, npw)."""
        n = 0 if psi_occ is None else psi_occ.shape[0]
        ptr = C.c_void_p(psi_occ.data_ptr()) if n else None
        check(_lib.lib().pw_set_exx(self.h, ecut, mu, frac, divtype, n, ptr,
                                    np.ascontiguousarray(occ if n else np.zeros(1), np.float64)))

    def add_exx(self, x, out=None):
        return self._binop("pw_add_exx", x, out, zero_out=True)

    def build_ace(self, psi=None):
        """Compress the exchange operator on the FP64 device block psi
        (None drops it); apply_h/mult_spinor then use the compressed form."""
        n = 0 if psi is None else psi.shape[0]
        check(_lib.lib().pw_build_ace(self.h, n, C.c_void_p(psi.data_ptr()) if n else None))

    # ---- inputs that the adjoint differentiates ----
    def set_vtot(self, vtot):
        v = np.ascontiguousarray(np.asarray(vtot, dtype=np.float64).reshape(-1))
        check(_lib.lib().pw_set_vtot(self.h, v))

    def set_atoms(self, pos):
        """Rebuild the Gaussian projector block at new atom positions (same
        recipe and couplings as the problem) and load it."""
        p = self.prob
        pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
        beta = np.zeros(2 * p.npw * p.nbeta)
        w = np.zeros(p.nbeta)
        check(_lib.lib().pw_gaussian_projectors(
            np.asarray(p.grid, np.int32), np.ascontiguousarray(p.supercell, np.float64).reshape(-1), p.npw, p.idx,
            len(pos), pos.reshape(-1), p.proj_sigma, p.lmax, np.ascontiguousarray(p.wgt_channel.reshape(-1)),
            beta, w))
        check(_lib.lib().pw_set_nonlocal(self.h, p.nbeta, beta, w))
        self.atoms_loaded = pos.copy()

    def density(self, x, f, out=None, accumulate=False):
        t = self.torch
        if out is None:
            out = t.zeros(int(np.prod(self.prob.grid)), dtype=t.float64, device=self.device)
        check(_lib.lib().pw_density(self.h, _dtype_code(x), x.shape[0], C.c_void_p(x.data_ptr()),
                                    np.ascontiguousarray(f, np.float64), C.c_void_p(out.data_ptr()),
                                    1 if accumulate else 0))
        return out

    def calculate_density(self, x, occ, num_spin=2, num_occupied=None):
        """KohnSham::CalculateDensity: renormalised density (device, N) and
        its integral sum(rho) vol / N."""
        t = self.torch
        nocc = x.shape[0] if num_occupied is None else num_occupied
        out = t.empty(int(np.prod(self.prob.grid)), dtype=t.float64, device=self.device)
        val = C.c_double()
        check(_lib.lib().pw_calculate_density(self.h, _dtype_code(x), x.shape[0], C.c_void_p(x.data_ptr()),
                                              np.ascontiguousarray(occ, np.float64), num_spin, nocc,
                                              C.c_void_p(out.data_ptr()), C.byref(val)))
        return out, val.value

    def band_energy(self, x, f):
        e = C.c_double()
        check(_lib.lib().pw_band_energy(self.h, _dtype_code(x), x.shape[0], C.c_void_p(x.data_ptr()),
                                        np.ascontiguousarray(f, np.float64), C.byref(e)))
        return e.value

    HARRIS_TERMS = ("Exc", "Eself", "EIonSR", "EVdw", "Eext", "fermi", "Tbeta", "temperature")

    def harris_energy(self, density, vxc, vhart, pseudo, vol, eig, occ, num_spin, num_occupied, terms,
                      spindensity=None):
        """pw_harris_energy (SCF::CalculateHarrisEnergy, scf.cpp:1987-2140).
        density, vxc: device float64 (ncomp, ntot) tensors (DblNumMat(ntot,
        ncomp) storage); vhart, pseudo: (ntot,); spindensity: (2, ntot) for
        nspin = 2; eig, occ: host arrays; terms: dict of HARRIS_TERMS."""
        nspin, ntot = vxc.shape
        for a in (density, vxc, vhart, pseudo) + ((spindensity,) if spindensity is not None else ()):
            if a.dtype != self.torch.float64 or not a.is_contiguous() or a.device.type != "cuda":
                raise ValueError("harris_energy: device float64 contiguous arrays expected")
        t = np.array([float(terms[k]) for k in self.HARRIS_TERMS])
        e = C.c_double()
        check(_lib.lib().pw_harris_energy(
            self.h, ntot, nspin, vol, C.c_void_p(density.data_ptr()),
            C.c_void_p(spindensity.data_ptr()) if spindensity is not None else None,
            C.c_void_p(vxc.data_ptr()), C.c_void_p(vhart.data_ptr()), C.c_void_p(pseudo.data_ptr()), len(eig),
            np.ascontiguousarray(eig, np.float64), np.ascontiguousarray(occ, np.float64), int(num_spin),
            int(num_occupied), t, C.byref(e)))
        return e.value

    def nonlocal_force(self, x, f, natom, npa):
        F = np.zeros(3 * natom)
        check(_lib.lib().pw_nonlocal_force(self.h, _dtype_code(x), x.shape[0], C.c_void_p(x.data_ptr()),
                                           np.ascontiguousarray(f, np.float64), natom, npa, F))
        return F.reshape(natom, 3)

    def syevd(self, A):
        """lapack::Syevd('V', 'U') on a device FP64 square matrix (column-
        major, i.e. A[j, i] in torch's row-major view is element (i, j)).
        Returns (eigenvalues ascending, eigenvector matrix) as new tensors."""
        t = self.torch
        n = A.shape[0]
        Z = A.detach().clone().contiguous()
        w = t.empty(n, dtype=t.float64, device=self.device)
        check(_lib.lib().pw_syevd(self.h, n, C.c_void_p(Z.data_ptr()), n, C.c_void_p(w.data_ptr())))
        return w, Z

    def sygvd_batched(self, A, B):
        """lapack::Sygvd(1, 'V', 'U') on a batch of device FP64 problems A[b] x =
        lambda B[b] x (shape (nbatch, n, n), each column-major as in syevd).
        Returns (eigenvalues (nbatch, n) ascending, eigenvectors, Cholesky
        factors, info) as new tensors."""
        t = self.torch
        nbatch, n = A.shape[0], A.shape[1]
        Z = A.detach().clone().contiguous()
        U = B.detach().clone().contiguous()
        w = t.empty((nbatch, n), dtype=t.float64, device=self.device)
        info = t.empty(nbatch, dtype=t.int32, device=self.device)
        check(_lib.lib().pw_sygvd_batched(self.h, n, C.c_void_p(Z.data_ptr()), n, n * n, C.c_void_p(U.data_ptr()), n,
                                          n * n, C.c_void_p(w.data_ptr()), n, C.c_void_p(info.data_ptr()), nbatch))
        return w, Z, U, info

    def set_graphs(self, on=True):
        """Capture/replay the Chebyshev filter as a hipGraph (default on)."""
        check(_lib.lib().pw_set_graphs(self.h, 1 if on else 0))

    def set_timing(self, on=True):
        check(_lib.lib().pw_set_timing(self.h, 1 if on else 0))

    def autotune(self, nb=None, dtype=None, cache=None):
        """pw_autotune: time the slab layouts and projector-product tilings
        for nb bands in dtype (torch complex dtype, default complex128),
        apply the fastest; cache: path of the text cache or None."""
        t = self.torch
        code = PW_F32 if dtype == t.complex64 else PW_F64
        buf = C.create_string_buffer(512)
        check(_lib.lib().pw_autotune(self.h, code, int(nb or self.max_bands),
                                     cache.encode() if cache else None, buf, 512))
        return buf.value.decode()

    def timing(self):
        """(summed ms per stage [4], number of timed calls) since set_timing."""
        ms = np.zeros(4)
        n = C.c_int(0)
        check(_lib.lib().pw_get_timing(self.h, ms, C.byref(n)))
        return ms, n.value


class Batch:
    """Many cells in one device batch (pw_batch_*): FP32 blocks of all cells
    in one complex64 device buffer, cell i's (nb_i, npw_i) block at element
    offset xoff[i]."""

    def __init__(self, probs, stream=None, device="cuda:0"):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("pwdft_develop_amd.ops needs a GPU (no CPU fallback)")
        self.torch = torch
        self.probs = list(probs)
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        P = self.probs
        cat = lambda xs, dt: np.ascontiguousarray(np.concatenate([np.asarray(x, dt).reshape(-1) for x in xs]))
        grid = cat([p.grid for p in P], np.int32)
        sc = cat([p.supercell for p in P], np.float64)
        npw = np.array([p.npw for p in P], np.int32)
        idx = cat([p.idx for p in P], np.int32)
        npad = np.array([len(p.pad_src) for p in P], np.int32)
        psrc = cat([p.pad_src for p in P] + [np.zeros(1, np.int32)], np.int32)
        pdst = cat([p.pad_dst for p in P] + [np.zeros(1, np.int32)], np.int32)
        ekin = cat([p.ekin for p in P], np.float64)
        teter = cat([p.teter for p in P], np.float64)
        vtot = cat([p.vtot for p in P], np.float64)
        nbeta = np.array([p.nbeta for p in P], np.int32)
        beta = cat([np.ascontiguousarray(p.beta).view(np.float64) for p in P] + [np.zeros(1)], np.float64)
        wgt = cat([p.wgt for p in P] + [np.zeros(1)], np.float64)
        nb = np.array([p.nb for p in P], np.int32)
        h = C.c_void_p()
        check(_lib.lib().pw_batch_create(C.byref(h), len(P), grid, sc, npw, idx, npad, psrc, pdst, ekin, teter,
                                         vtot, nbeta, beta, wgt, nb, C.c_void_p(self.stream.cuda_stream)))
        self.h = h
        self.xoff = np.concatenate([[0], np.cumsum(npw.astype(np.int64) * nb)])
        self.xtotal = int(self.xoff[-1])

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().pw_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        xt, nj, nw, ld = C.c_long(), C.c_int(), C.c_int(), C.c_int()
        check(_lib.lib().pw_batch_info(self.h, C.byref(xt), C.byref(nj), C.byref(nw), C.byref(ld)))
        return {"xtotal": xt.value, "jobs_per_step": nj.value, "workgroups": nw.value, "lds_bytes": ld.value}

    def tune(self, rp=0, zc=0, nwg=0):
        check(_lib.lib().pw_batch_tune(self.h, int(rp), int(zc), int(nwg)))

    def set_graphs(self, on=True):
        check(_lib.lib().pw_batch_set_graphs(self.h, 1 if on else 0))

    def set_timing(self, on=True):
        check(_lib.lib().pw_batch_set_timing(self.h, 1 if on else 0))

    def timing(self):
        """ms per stage (z-inverse + projections, slabs, z-forward, apply) of
        the last launch, summed over its steps."""
        ms = np.zeros(4)
        check(_lib.lib().pw_batch_timing(self.h, ms))
        return ms

    def autotune(self, cache=None):
        """pw_batch_autotune: time the work splits, keep the fastest."""
        buf = C.create_string_buffer(512)
        check(_lib.lib().pw_batch_autotune(self.h, cache.encode() if cache else None, buf, 512))
        return buf.value.decode()

    def to_device(self, blocks):
        """Host (nb_i, npw_i) complex blocks -> one complex64 device buffer."""
        t = self.torch
        flat = np.concatenate([np.asarray(b, np.complex64).reshape(-1) for b in blocks])
        return t.from_numpy(flat).to(self.device)

    def split(self, X):
        """Device buffer -> list of host (nb_i, npw_i) complex64 blocks."""
        h = X.cpu().numpy()
        return [h[self.xoff[i]:self.xoff[i + 1]].reshape(p.nb, p.npw) for i, p in enumerate(self.probs)]

    def apply_h(self, X, out=None):
        out = self.torch.empty_like(X) if out is None else out
        check(_lib.lib().pw_batch_apply_h(self.h, C.c_void_p(X.data_ptr()), C.c_void_p(out.data_ptr())))
        return out

    def cheb_filter(self, X, m, bounds, serial=False):
        """In place; bounds: (ncell, 3) array of (a, b, aL)."""
        bnd = np.ascontiguousarray(np.asarray(bounds, np.float64).reshape(-1))
        f = _lib.lib().pw_batch_cheb_filter_serial if serial else _lib.lib().pw_batch_cheb_filter
        check(f(self.h, int(m), bnd, C.c_void_p(X.data_ptr())))
        return X
