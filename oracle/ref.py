"""ctypes front end of oracle/_ref (the reference's own functions, compiled
from /root/reference by Makefile.ref).  TEST INFRASTRUCTURE ONLY.

The reference reads its input from ``patin.dat`` in the current directory
(basis.c:251), so :func:`load` copies nothing: callers write a patin.dat into
a scratch directory and pass it here; the loader changes into that directory
only for the duration of ``init_Basis``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libpatref_harness.so")
REF_DIR = "/root/reference"

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def available() -> bool:
    return os.path.exists(LIB) or os.path.isdir(REF_DIR)


def build() -> bool:
    if not os.path.isdir(REF_DIR):
        return os.path.exists(LIB)
    subprocess.run(["make", "-s", "-f", "Makefile.ref", "-C", HERE], check=True)
    return True


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.ref_init_basis.argtypes = []
        L.ref_norb.restype = C.c_int
        L.ref_shell_coefs.argtypes = [_dp]
        L.ref_one_electron_ints.argtypes = [_dp, _dp, _dp]
        L.ref_rys_roots.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.ref_dpm_ints.argtypes = [_dp]
        L.ref_set_center.argtypes = [C.c_int, _dp]
        L.ref_nuclear_repulsion.restype = C.c_double
        L.ref_formMDD.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        L.ref_formMDR.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double, _dp]
        L.ref_eri_block.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.ref_spack_diag.argtypes = [C.c_int, _dp, _dp, _dp]
        L.ref_diag_fock.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.ref_check_phase.argtypes = [C.c_int, _dp]
        L.momints.argtypes = []
        _lib = L
    return _lib


def load(workdir: str):
    """init_Basis() on workdir/patin.dat (one basis per process)."""
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        lib().ref_init_basis()
    finally:
        os.chdir(cwd)


def rys_roots(n, T):
    """Reference roots as u = t^2/(1-t^2) and weights."""
    u, w = np.zeros(max(n, 11)), np.zeros(max(n, 11))
    lib().ref_rys_roots(n, T, u, w)
    return u[:n].copy(), w[:n].copy()


def eri_block(I, J, K, L, nst):
    out = np.zeros(int(np.prod([nst[I], nst[J], nst[K], nst[L]])))
    lib().ref_eri_block(I, J, K, L, out)
    return out.reshape(nst[I], nst[J], nst[K], nst[L])


def one_electron_ints(norb):
    no2 = norb * (norb + 1) // 2
    S, T, H = np.zeros(no2), np.zeros(no2), np.zeros(no2)
    lib().ref_one_electron_ints(S, T, H)
    return S, T, H


def shell_coefs(nprim):
    out = np.zeros(nprim)
    lib().ref_shell_coefs(out)
    return out


def formMDD(l1, l2, abi, ax, bx):
    out = np.zeros((l1 + 1) * (l2 + 1) * (l1 + l2 + 1))
    lib().ref_formMDD(out, l1, l2, abi, ax, bx)
    return out.reshape(l1 + 1, l2 + 1, l1 + l2 + 1)


def formMDR(L, sr, t, w, pq):
    out = np.zeros((L + 1) ** 3)
    lib().ref_formMDR(out, L, sr, t, w, np.ascontiguousarray(pq, dtype=np.float64))
    return out.reshape(L + 1, L + 1, L + 1)


def diag_fock(S, F, nocc):
    """(eigenvalues, C row-major, packed P) of the reference's diagonalisation step."""
    no2 = len(S)
    n = int(round((np.sqrt(8 * no2 + 1) - 1) / 2))
    ev, Cm, P = np.zeros(n), np.zeros(n * n), np.zeros(no2)
    lib().ref_diag_fock(n, np.ascontiguousarray(S), np.ascontiguousarray(F), nocc, ev, Cm, P)
    return ev, Cm.reshape(n, n), P


MOMINTS_DTYPE = np.dtype([("ir", "<i4"), ("jr", "<i4"), ("dip", "<f8", 3), ("quad", "<f8", 6)])


def momints(workdir: str):
    """The reference's momints() (momints.c:94-190) run in workdir; returns
    the records of the MOMINTS.DAT it writes there."""
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        lib().momints()
    finally:
        os.chdir(cwd)
    return np.fromfile(os.path.join(workdir, "MOMINTS.DAT"), dtype=MOMINTS_DTYPE)
