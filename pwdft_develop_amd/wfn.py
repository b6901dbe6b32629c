"""Wavefunction restart files in the reference's serialize layout (Python
mirror of host/wfn_io.hpp for the test binding).

Layout (include/utility.hpp:761-830, 1088-1094, 1373-1381, 1439-1452):
Int = 4-byte int, Real = 8-byte double, Complex = 16 bytes; NumVec/IntNumVec
= Int m + m values; NumTns = Int m, n, p + values column-major.  Records as
written by SCF::Iterate's complex-wavefunction branch (src/SCF/scf.cpp:
1115-1142): WFN_<index> = [CpxNumTns npw x ncom x nband][DblNumVec occ];
IDX_<index> = [IntNumVec idxCoarseCut].  File names "<name>_<index>" as
SeparateWrite(name, os, index) (src/utility.cpp:499-510).
"""
import numpy as np

_I = np.dtype("<i4")


def file_name(name, index):
    return f"{name}_{index}"


def _vec(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
    return np.array([a.size], _I).tobytes() + a.tobytes()


def _tns(psi):
    """psi: (nband, ncom, npw) array, i.e. the column-major npw x ncom x
    nband tensor in row-major storage."""
    nband, ncom, npw = psi.shape
    return np.array([npw, ncom, nband], _I).tobytes() + np.ascontiguousarray(psi, np.complex128).tobytes()


def write_restart(wfn_name, idx_name, index, psi, occ, idx):
    """psi: (nband, npw) or (nband, ncom, npw) complex coefficients."""
    psi = np.asarray(psi)
    if psi.ndim == 2:
        psi = psi[:, None, :]
    with open(file_name(wfn_name, index), "wb") as f:
        f.write(_tns(psi) + _vec(occ, np.float64))
    with open(file_name(idx_name, index), "wb") as f:
        f.write(_vec(idx, _I))


class _Stream:
    def __init__(self, data):
        self.b, self.p = data, 0

    def take(self, dtype, count):
        n = np.dtype(dtype).itemsize * count
        if self.p + n > len(self.b):
            raise ValueError("wfn: file shorter than its records")
        out = np.frombuffer(self.b, dtype=dtype, count=count, offset=self.p)
        self.p += n
        return out

    def vec(self, dtype):
        m = int(self.take(_I, 1)[0])
        if m < 0:
            raise ValueError("wfn: negative vector length")
        return self.take(dtype, m).copy()


def read_restart(wfn_name, idx_name, index):
    """Returns (psi (nband, ncom, npw) complex, occ, idx)."""
    with open(file_name(wfn_name, index), "rb") as f:
        s = _Stream(f.read())
    npw, ncom, nband = (int(x) for x in s.take(_I, 3))
    if min(npw, ncom, nband) < 0:
        raise ValueError("wfn: negative tensor dimension")
    psi = s.take(np.complex128, npw * ncom * nband).reshape(nband, ncom, npw).copy()
    occ = s.vec(np.float64)
    if s.p != len(s.b):
        raise ValueError("wfn: trailing bytes after the occupation rates")
    with open(file_name(idx_name, index), "rb") as f:
        idx = _Stream(f.read()).vec(_I)
    if idx.size != npw:
        raise ValueError("wfn: index file does not match the coefficients")
    return psi, occ, idx
