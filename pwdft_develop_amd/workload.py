"""Synthetic, seeded inputs for the H*psi / eigensolver workloads.

Random numbers come from Philox-4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11)
with Box-Muller normals, so every input is a pure function of (seed, stream,
index) and can be regenerated bit-for-bit on any host.

The physical recipe (cell, atoms, potential, projectors, wavefunctions) is the
one the workloads are defined on:
  * V(r) = sum_a -2 exp(-|r - R_a|^2 / (2 * 1.2^2)) over the 27 nearest
    periodic images, plus U[0, 0.01) Ha noise per grid point;
  * 4 projectors per atom (s, pz, px, py): r^l exp(-r^2/(2*0.8^2)) Y_lm, one
    coupling per channel drawn from +-U[0.5, 2.0) Ha;
  * psi: standard-normal real and imaginary parts on the half sphere,
    Im c(G=0) = 0 (examples/pwdft.cpp:202-208), then orthonormalised by
    Cholesky QR in the Re(a^H b) inner product.
Grid, sphere and projector blocks are built by the library's host setup
(pw_sphere_build, pw_ekin_teter, pw_gaussian_projectors).
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib

MASTER_SEED = 20261202
STREAM_PSI, STREAM_VNOISE, STREAM_WGT, STREAM_ATOMS, STREAM_LANCZOS, STREAM_BATCH = 0, 1, 2, 3, 4, 5

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32(ctr, key, rounds=10):
    """Philox-4x32 bijection.  ctr: (4, n) uint32-valued array, key: 2 ints.
    Returns a (4, n) uint32 array."""
    c = [np.asarray(x, dtype=np.uint64) & _MASK for x in ctr]
    k0 = np.uint64(key[0]) & _MASK
    k1 = np.uint64(key[1]) & _MASK
    c0, c1, c2, c3 = c
    for _ in range(rounds):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return np.stack([c0, c1, c2, c3]).astype(np.uint32)


def random_u32(seed, stream, n):
    """n uint32 values of stream `stream`: block b uses counter (b, 0, stream, 0)."""
    nblk = (n + 3) // 4
    b = np.arange(nblk, dtype=np.uint64)
    ctr = [b & _MASK, b >> np.uint64(32), np.full(nblk, stream, np.uint64), np.zeros(nblk, np.uint64)]
    out = philox4x32(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF))
    return out.T.reshape(-1)[:n]


def random_uniform(seed, stream, n):
    """U(0,1), never 0 or 1: (u + 0.5) / 2^32."""
    return (random_u32(seed, stream, n).astype(np.float64) + 0.5) * (1.0 / 4294967296.0)


def random_normal(seed, stream, n):
    """Box-Muller pairs from consecutive uniforms."""
    m = (n + 1) // 2
    u = random_uniform(seed, stream, 2 * m)
    r = np.sqrt(-2.0 * np.log(u[0::2]))
    th = 2.0 * np.pi * u[1::2]
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(th)
    z[1::2] = r * np.sin(th)
    return z[:n]


DIAMOND = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                    [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])


@dataclass
class Problem:
    name: str
    grid: tuple
    supercell: np.ndarray
    ecut: float
    atoms: np.ndarray          # (natom, 3) cartesian bohr
    nb: int
    seed: int = MASTER_SEED
    vdepth: float = -2.0
    vwidth: float = 1.2
    vnoise: float = 0.01
    proj_sigma: float = 0.8
    lmax: int = 1
    # filled by build()
    npw: int = 0
    idx: np.ndarray = field(default=None, repr=False)
    gkk: np.ndarray = field(default=None, repr=False)
    pad_src: np.ndarray = field(default=None, repr=False)
    pad_dst: np.ndarray = field(default=None, repr=False)
    ekin: np.ndarray = field(default=None, repr=False)
    teter: np.ndarray = field(default=None, repr=False)
    vtot: np.ndarray = field(default=None, repr=False)   # (n3, n2, n1)
    beta: np.ndarray = field(default=None, repr=False)   # (nbeta, npw) complex
    wgt: np.ndarray = field(default=None, repr=False)
    wgt_channel: np.ndarray = field(default=None, repr=False)
    nbeta: int = 0
    psi: np.ndarray = field(default=None, repr=False)    # (nb, npw) complex

    @property
    def volume(self):
        return abs(np.linalg.det(self.supercell))

    @property
    def ngrid(self):
        return int(np.prod(self.grid))

    def build(self, with_psi=True):
        L = _lib.lib()
        grid = np.asarray(self.grid, np.int32)
        sc = np.ascontiguousarray(self.supercell, np.float64).reshape(-1)
        npw, npad = C.c_int(0), C.c_int(0)
        _lib.check(L.pw_sphere_build(grid, sc, self.ecut, C.byref(npw), C.byref(npad), None, None, None, None))
        self.npw = npw.value
        self.idx = np.zeros(self.npw, np.int32)
        self.gkk = np.zeros(self.npw)
        self.pad_src = np.zeros(max(npad.value, 1), np.int32)
        self.pad_dst = np.zeros(max(npad.value, 1), np.int32)
        _lib.check(L.pw_sphere_build(grid, sc, self.ecut, C.byref(npw), C.byref(npad),
                                     self.idx.ctypes.data, self.gkk.ctypes.data,
                                     self.pad_src.ctypes.data, self.pad_dst.ctypes.data))
        self.pad_src = self.pad_src[:npad.value].copy()
        self.pad_dst = self.pad_dst[:npad.value].copy()
        self.ekin = np.zeros(self.npw)
        self.teter = np.zeros(self.npw)
        _lib.check(L.pw_ekin_teter(self.npw, self.gkk, self.ekin, self.teter))
        self.vtot = self.make_vtot()
        self.make_projectors()
        if with_psi:
            self.psi = self.make_psi(self.nb)
        return self

    def make_vtot(self):
        n1, n2, n3 = self.grid
        Ls = np.diag(self.supercell)
        if not np.allclose(self.supercell, np.diag(Ls)):
            raise ValueError("the synthetic potential recipe needs an orthorhombic cell")
        axes = [np.arange(n) * (Ls[d] / n) for d, n in enumerate(self.grid)]
        V = np.zeros((n3, n2, n1))
        inv2s2 = 1.0 / (2.0 * self.vwidth ** 2)
        for R in self.atoms:
            g = []
            for d in range(3):
                dx = axes[d][:, None] - R[d] - np.array([-1.0, 0.0, 1.0])[None, :] * Ls[d]
                g.append(np.exp(-dx * dx * inv2s2).sum(axis=1))
            V += self.vdepth * g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
        V += self.vnoise * random_uniform(self.seed, STREAM_VNOISE, V.size).reshape(V.shape)
        return V

    def make_projectors(self):
        natom = len(self.atoms)
        nch = self.lmax + 1
        u = random_uniform(self.seed, STREAM_WGT, 2 * natom * nch)
        mag = 0.5 + 1.5 * u[0::2]
        sgn = np.where(u[1::2] < 0.5, -1.0, 1.0)
        self.wgt_channel = (sgn * mag).reshape(natom, nch)
        self.nbeta = natom * nch * nch
        beta = np.zeros(2 * self.npw * self.nbeta)
        w = np.zeros(self.nbeta)
        _lib.check(_lib.lib().pw_gaussian_projectors(
            np.asarray(self.grid, np.int32), np.ascontiguousarray(self.supercell, np.float64).reshape(-1),
            self.npw, self.idx, natom, np.ascontiguousarray(self.atoms, np.float64).reshape(-1),
            self.proj_sigma, self.lmax, np.ascontiguousarray(self.wgt_channel.reshape(-1)), beta, w))
        self.beta = beta.view(np.complex128).reshape(self.nbeta, self.npw)
        self.wgt = w

    def make_psi(self, nb, stream=STREAM_PSI, orthonormalise=True):
        z = random_normal(self.seed, stream, 2 * self.npw * nb)
        psi = z.view(np.complex128).reshape(nb, self.npw).copy()
        psi[:, 0] = psi[:, 0].real
        if orthonormalise:
            psi = cholesky_qr(psi)
        return psi

    def lanczos_start(self):
        """Deterministic Lanczos start vector (stream 4)."""
        return self.make_psi(1, stream=STREAM_LANCZOS, orthonormalise=False)


def cholesky_qr(psi):
    """psi (nb, npw) complex -> orthonormal in Re(a^H b) (eigensolver.cpp:386-469)."""
    Xr = psi.view(np.float64).reshape(psi.shape[0], -1)   # (nb, 2npw)
    S = Xr @ Xr.T
    U = np.linalg.cholesky(S).T                            # S = U^T U
    Y = np.linalg.solve(U.T, Xr)                           # rows of (X U^-1)^T = U^-T X^T
    return np.ascontiguousarray(Y).view(np.complex128).reshape(psi.shape)


def cubic(L):
    return np.diag([L, L, L]).astype(np.float64)


def si8_32(nb=32):
    """8-atom diamond cell, side 10.26 bohr, 32^3 grid, Ecut 10 Ha."""
    L = 10.26
    return Problem("si8_32", (32, 32, 32), cubic(L), 10.0, DIAMOND * L, nb)


def si64_60(nb=256):
    """2x2x2 diamond supercell (64 atoms), side 20.52 bohr, 60^3 grid, Ecut 10 Ha."""
    L0 = 10.26
    cells = np.array([[i, j, k] for i in range(2) for j in range(2) for k in range(2)], float)
    atoms = ((DIAMOND[None, :, :] + cells[:, None, :]) * L0).reshape(-1, 3)
    return Problem("si64_60", (60, 60, 60), cubic(2 * L0), 10.0, atoms, nb)


def random_atoms(n, L, min_sep, seed, stream=STREAM_ATOMS):
    """Rejection-sampled positions with a minimum periodic separation."""
    out = []
    u = random_uniform(seed, stream, 3 * 200000)
    k = 0
    while len(out) < n:
        if k + 3 > len(u):
            raise RuntimeError("could not place atoms")
        p = u[k:k + 3] * L
        k += 3
        ok = True
        for q in out:
            d = p - q
            d -= L * np.round(d / L)
            if np.dot(d, d) < min_sep * min_sep:
                ok = False
                break
        if ok:
            out.append(p)
    return np.array(out)


def rand16_48(nb=128):
    """16 random atoms (minimum separation 3.5 bohr), side 16.86 bohr, 48^3 grid."""
    L = 16.86
    return Problem("rand16_48", (48, 48, 48), cubic(L), 10.0,
                   random_atoms(16, L, 3.5, MASTER_SEED), nb)


LAYERED_CELL = (63.22, 63.22, 25.29)


def layered432_atoms():
    """432 atoms in three layers of a 12 x 12 square net at z = 1/6, 1/2,
    5/6 of the 63.22 x 63.22 x 25.29 bohr cell, the middle layer shifted by
    half a site."""
    Lx, _, Lz = LAYERED_CELL
    a = Lx / 12
    atoms = []
    for layer, zf in enumerate((1 / 6, 1 / 2, 5 / 6)):
        sh = 0.5 * a if layer % 2 else 0.0
        for i in range(12):
            for j in range(12):
                atoms.append([i * a + sh, j * a + sh, zf * Lz])
    return np.array(atoms)


def layered432_180(nb=1024):
    """The large single cell: orthorhombic 63.22 x 63.22 x 25.29 bohr, 180 x
    180 x 72 grid (Ecut 10 Ha, npw 76241), 432 atoms, 1024 bands."""
    return Problem("layered432_180", (180, 180, 72), np.diag(LAYERED_CELL).astype(np.float64), 10.0,
                   layered432_atoms(), nb)


def layered432_96(nb=432):
    """The same cell on the 96 x 96 x 40 exchange subgrid at Ecut 2.8 Ha:
    the Fock sphere (|G|^2/2 < 4 Ecut) reaches |kx|, |ky| = 47 and |kz| =
    19, one index short of the subgrid's Nyquist planes (from 2.85 Ha on it
    reaches them); 432 bands (the occupied count of the exchange apply)."""
    return Problem("layered432_96", (96, 96, 40), np.diag(LAYERED_CELL).astype(np.float64), 2.8,
                   layered432_atoms(), nb)


BATCH_GRIDS = (24, 30, 36, 40, 48, 60, 72)


def batch_cells(ncell=48, seed=MASTER_SEED, min_sep=2.0):
    """Uneven cells of the batched workload: per cell a cubic grid n drawn
    from BATCH_GRIDS, side L = n pi / sqrt(80) bohr (so that Ecut 10 Ha
    gives that grid by the reference's rule), 2..32 atoms at random
    positions (minimum separation min_sep), 16..128 bands; the potential,
    projector and wavefunction recipe of the other workloads with a seed of
    its own per cell.  Returns unbuilt Problems."""
    u = random_uniform(seed, STREAM_BATCH, 3 * ncell).reshape(ncell, 3)
    cells = []
    for i in range(ncell):
        n = BATCH_GRIDS[min(len(BATCH_GRIDS) - 1, int(u[i, 0] * len(BATCH_GRIDS)))]
        natom = 2 + min(30, int(u[i, 1] * 31))
        nb = 16 + min(112, int(u[i, 2] * 113))
        L = n * np.pi / np.sqrt(80.0)
        cseed = seed + 1000 + i
        cells.append(Problem(f"batch{i:02d}_{n}", (n, n, n), cubic(L), 10.0,
                             random_atoms(natom, L, min_sep, cseed), nb, seed=cseed))
    return cells


WORKLOADS = {"si8_32": si8_32, "si64_60": si64_60, "rand16_48": rand16_48, "layered432_180": layered432_180,
             "layered432_96": layered432_96}
