"""patmol_amd -- MI355X (gfx950) two-electron hot path of Patmol.

Python front end over the two in-tree shared libraries:

* ``libpatmol_hip.so``  -- HIP kernels and the reference-compatible C ABI
  (init_ERIFS, two_electron_ints, rhf_form_gmatrix, uhf_form_gmatrix,
  dpm_ints, clean_ERIFS) plus extensions (device-resident G build, stats,
  ERI export, work-list export, options);
* ``libpatmol_host.so`` -- the caller side: basis model behind the
  reference's basis.h accessors, one-electron integrals and the RHF/UHF
  SCF drivers.

There is no CPU fallback: :func:`lib` raises if the libraries are missing,
and every call checks the library's status code.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

from . import systems  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
# PATMOL_AMD_LIB_DIR points at another build of the two libraries (the
# autotuner's candidate builds); by default the in-tree ones
_LIB_DIR = os.environ.get("PATMOL_AMD_LIB_DIR") or HERE
LIB_HIP = os.path.join(_LIB_DIR, "libpatmol_hip.so")
LIB_HOST = os.path.join(_LIB_DIR, "libpatmol_host.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class PatmolError(RuntimeError):
    pass


NOALLOC_MAX_NORB = 256   # PATMOL_HIP_NOALLOC_MAX_NORB (include/patmol_hip.h)


class HipStats(C.Structure):
    _fields_ = [("nshell", C.c_int), ("norb", C.c_int), ("start", C.c_int),
                ("pairs", C.c_int64), ("shell_quartets_total", C.c_int64),
                ("shell_quartets_kept", C.c_int64), ("prim_quartets_kept", C.c_int64),
                ("flops_per_build", C.c_double), ("work_items", C.c_int64),
                ("work_items_fp32", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ScfResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("ediff", C.c_double), ("pdiff", C.c_double),
                ("enuc", C.c_double), ("eelec", C.c_double), ("iterations", C.c_int),
                ("converged", C.c_int), ("gbuild_seconds", C.c_double),
                ("diag_seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class HipScfResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("eelec", C.c_double), ("enuc", C.c_double), ("ediff", C.c_double),
                ("pdiff", C.c_double), ("iterations", C.c_int), ("converged", C.c_int),
                ("gbuild_ms", C.c_double), ("other_ms", C.c_double), ("total_ms", C.c_double),
                ("energy0", C.c_double), ("purify_fallbacks", C.c_int), ("full_builds", C.c_int),
                ("delta_builds", C.c_int), ("items", C.c_double), ("items_fp32", C.c_double),
                ("items_skipped", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build(jobs: int = 8) -> None:
    """Compile both libraries in-tree (hipcc cross-compiles gfx950 without a GPU)."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", os.path.join(HERE, "csrc")], check=True)


_host = None
_hip = None


def lib():
    """(host, hip) ctypes handles; raises if the libraries are not built."""
    global _host, _hip
    if _host is None:
        for p in (LIB_HIP, LIB_HOST):
            if not os.path.exists(p):
                raise PatmolError(f"{p} is missing: run patmol_amd.build() (no CPU fallback exists)")
        H = C.CDLL(LIB_HOST, mode=C.RTLD_GLOBAL)
        G = C.CDLL(LIB_HIP, mode=C.RTLD_GLOBAL)
        H.patmol_basis_load.restype = C.c_int
        H.patmol_basis_load.argtypes = [C.c_int, _dp, _dp, C.c_int, _ip, _ip, _ip, _dp, _dp,
                                        C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]
        H.patmol_basis_add_dpm_center.restype = C.c_int
        H.number_of_orbitals.restype = C.c_int
        H.number_of_shells.restype = C.c_int
        H.nuclear_repulsion.restype = C.c_double
        H.set_center_position.argtypes = [C.c_int, _dp]
        H.patmol_host_one_electron_ints.argtypes = [_dp, _dp, _dp]
        H.patmol_rhf_scf.argtypes = [_dp, _dp, C.POINTER(ScfResult)]
        H.patmol_uhf_scf.argtypes = [_dp, _dp, C.POINTER(ScfResult)]
        G.init_ERIFS.restype = C.c_void_p
        G.init_ERIFS.argtypes = [C.c_int, C.c_int, C.c_char_p]
        G.clean_ERIFS.argtypes = [C.c_void_p]
        G.two_electron_ints.argtypes = [C.c_void_p, C.c_int]
        G.rhf_form_gmatrix.argtypes = [C.c_void_p, _dp, _dp]
        G.uhf_form_gmatrix.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        G.dpm_ints.argtypes = [_dp]
        G.one_electron_ints.argtypes = [_dp, _dp, _dp]
        G.patmol_hip_last_error.restype = C.c_int
        G.patmol_hip_error_string.restype = C.c_char_p
        G.patmol_hip_set_strict.argtypes = [C.c_int]
        G.patmol_hip_set_schwarz_threshold.argtypes = [C.c_double]
        G.patmol_hip_set_precision.argtypes = [C.c_int, C.c_double]
        G.patmol_hip_set_fp32_policy.argtypes = [C.c_int]
        G.patmol_hip_set_jk_engine.argtypes = [C.c_int]
        G.patmol_hip_rhf_gmatrix_device.restype = C.c_int
        G.patmol_hip_rhf_gmatrix_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        G.patmol_hip_rhf_gmatrix_device_noalloc.restype = C.c_int
        G.patmol_hip_rhf_gmatrix_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        G.patmol_hip_stats.restype = C.c_int
        G.patmol_hip_stats.argtypes = [C.c_void_p, C.POINTER(HipStats)]
        G.patmol_hip_class_stats.restype = C.c_int
        G.patmol_hip_class_stats.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(np.int64), _dp]
        G.patmol_hip_pair_schwarz.restype = C.c_int
        G.patmol_hip_pair_schwarz.argtypes = [C.c_void_p, _dp]
        G.patmol_hip_jk_engine_mask.restype = C.c_int
        G.patmol_hip_jk_engine_mask.argtypes = [C.c_void_p]
        G.patmol_hip_work_items.restype = C.c_int64
        G.patmol_hip_work_items.argtypes = [C.c_void_p] + [C.c_void_p] * 7
        G.patmol_hip_eri_blocks.restype = C.c_int
        G.patmol_hip_eri_blocks.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
        G.patmol_hip_eri_blocks_prec.restype = C.c_int
        G.patmol_hip_eri_blocks_prec.argtypes = [C.c_void_p, C.c_int, _ip, _dp, C.c_int]
        G.patmol_hip_last_timing.restype = C.c_int
        G.patmol_hip_last_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), _dp]
        G.patmol_hip_set_timing.argtypes = [C.c_int]
        _odp = C.c_void_p  # optional double arrays (None or array.ctypes.data)
        G.patmol_hip_rhf_scf.restype = C.c_int
        G.patmol_hip_rhf_scf.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_int,
                                         C.c_double, _odp, _odp, _odp, _odp, C.POINTER(HipScfResult)]
        G.patmol_hip_uhf_scf.restype = C.c_int
        G.patmol_hip_uhf_scf.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double,
                                         C.c_int, C.c_double, _odp, _odp, _odp, _odp, _odp, _odp,
                                         C.POINTER(HipScfResult)]
        G.patmol_hip_set_scf_density.argtypes = [C.c_int]
        G.patmol_hip_set_scf_incremental.argtypes = [C.c_int]
        G.patmol_hip_diag_fock.restype = C.c_int
        G.patmol_hip_diag_fock.argtypes = [_dp, _dp, C.c_int, _odp, _odp, _odp]
        G.patmol_hip_uhf_gmatrix_device.restype = C.c_int
        G.patmol_hip_write_erints.restype = C.c_int64
        G.patmol_hip_write_erints.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
        G.patmol_hip_rhf_gmatrix_erints.restype = C.c_int
        G.patmol_hip_rhf_gmatrix_erints.argtypes = [C.c_void_p, C.c_int64, _dp, _dp]
        G.patmol_hip_uhf_gmatrix_erints.restype = C.c_int
        G.patmol_hip_uhf_gmatrix_erints.argtypes = [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp]
        G.patmol_hip_rhf_gmatrix_erint_file.restype = C.c_int
        G.patmol_hip_rhf_gmatrix_erint_file.argtypes = [C.c_char_p, _dp, _dp]
        G.patmol_hip_uhf_gmatrix_erint_file.restype = C.c_int
        G.patmol_hip_uhf_gmatrix_erint_file.argtypes = [C.c_char_p, _dp, _dp, _dp, _dp]
        G.patmol_hip_moment_ints.restype = C.c_int
        G.patmol_hip_moment_ints.argtypes = [_dp, _dp]
        G.patmol_hip_write_momints.restype = C.c_int
        G.patmol_hip_write_momints.argtypes = [C.c_char_p]
        G.patmol_hip_uhf_gmatrix_device.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        H.scf_convergence.restype = C.c_double
        H.patmol_rhf_moments.argtypes = [_dp, _dp, _dp, _dp]
        H.patmol_uhf_moments.argtypes = [_dp, _dp, _dp, _dp, _dp]
        H.patmol_rhf_finite_field.argtypes = [C.c_double, _dp, C.POINTER(ScfResult)]
        H.patmol_rhf_gradient.restype = C.c_int
        H.patmol_rhf_gradient.argtypes = [C.c_void_p, _dp, _dp]
        H.patmol_rhf_dpm_gradient.restype = C.c_int
        H.patmol_rhf_dpm_gradient.argtypes = [C.c_void_p, _dp, _dp]
        H.patmol_host_onee_gradient.restype = C.c_int
        H.patmol_host_onee_gradient.argtypes = [_dp, _dp, _dp]
        H.patmol_nuclear_repulsion_gradient.argtypes = [_dp]
        G.patmol_hip_eri_gradient.restype = C.c_int
        G.patmol_hip_eri_gradient.argtypes = [C.c_void_p, _dp, _dp]
        G.patmol_hip_onee_gradient.restype = C.c_int
        G.patmol_hip_onee_gradient.argtypes = [_dp, _dp, _dp]
        G.patmol_hip_dpm_gradient.restype = C.c_int
        G.patmol_hip_dpm_gradient.argtypes = [_dp, _dp]
        G.patmol_hip_batch_new.restype = C.c_void_p
        G.patmol_hip_batch_free.argtypes = [C.c_void_p]
        G.patmol_hip_batch_add.restype = C.c_int
        G.patmol_hip_batch_add.argtypes = [C.c_void_p]
        G.patmol_hip_batch_finalize.restype = C.c_int
        G.patmol_hip_batch_finalize.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        G.patmol_hip_batch_gmatrix.restype = C.c_int
        G.patmol_hip_batch_gmatrix.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        G.patmol_hip_batch_buffers.restype = C.c_int
        G.patmol_hip_batch_buffers.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
        G.patmol_hip_batch_launch.restype = C.c_int
        G.patmol_hip_batch_launch.argtypes = [C.c_void_p, C.c_void_p]
        G.patmol_hip_batch_info.restype = C.c_int
        G.patmol_hip_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        G.patmol_hip_batch_rhf_scf.restype = C.c_int
        G.patmol_hip_batch_rhf_scf.argtypes = [C.c_void_p, _dp, _dp, _ip, _dp, C.c_int, C.c_double, C.c_int, _dp, _ip,
                                               _ip, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        H.patmol_dpm_calculation.restype = C.c_int
        H.patmol_dpm_calculation.argtypes = [C.c_int, C.c_int, _dp, C.c_void_p, C.POINTER(ScfResult)]
        _host, _hip = H, G
    return _host, _hip


def _check(what: str = ""):
    _, G = lib()
    code = G.patmol_hip_last_error()
    if code:
        raise PatmolError(f"{what}: error {code}: {G.patmol_hip_error_string().decode()}")


def exported_symbols():
    """Names the C ABI declares (include/patmol_hip.h)."""
    return ["init_ERIFS", "clean_ERIFS", "two_electron_ints", "rhf_form_gmatrix", "uhf_form_gmatrix",
            "dpm_ints", "one_electron_ints", "patmol_hip_last_error", "patmol_hip_error_string", "patmol_hip_set_strict",
            "patmol_hip_set_schwarz_threshold", "patmol_hip_set_precision", "patmol_hip_set_fp32_policy",
            "patmol_hip_set_jk_engine",
            "patmol_hip_rhf_gmatrix_device", "patmol_hip_rhf_gmatrix_device_noalloc", "patmol_hip_stats", "patmol_hip_class_stats",
            "patmol_hip_pair_schwarz", "patmol_hip_jk_engine_mask",
            "patmol_hip_work_items", "patmol_hip_eri_blocks", "patmol_hip_eri_blocks_prec", "patmol_hip_last_timing",
            "patmol_hip_set_timing", "patmol_hip_rhf_scf", "patmol_hip_uhf_scf", "patmol_hip_diag_fock",
            "patmol_hip_uhf_gmatrix_device", "patmol_hip_set_scf_density", "patmol_hip_set_scf_incremental", "momints", "patmol_hip_moment_ints", "patmol_hip_write_momints",
            "patmol_hip_write_erints", "patmol_hip_rhf_gmatrix_erints", "patmol_hip_rhf_gmatrix_erint_file",
            "patmol_hip_uhf_gmatrix_erints", "patmol_hip_uhf_gmatrix_erint_file", "patmol_hip_eri_gradient",
            "patmol_hip_onee_gradient",
            "patmol_hip_dpm_gradient",
            "patmol_hip_batch_new", "patmol_hip_batch_add", "patmol_hip_batch_finalize", "patmol_hip_batch_gmatrix",
            "patmol_hip_batch_buffers", "patmol_hip_batch_launch", "patmol_hip_batch_info", "patmol_hip_batch_free",
            "patmol_hip_batch_rhf_scf"]


def load_system(sys_: "systems.System") -> None:
    """Make `sys_` the current basis (the reference's global basis model)."""
    H, _ = lib()
    q, xyz, npr, lsh, cen, al, co = sys_.flat()
    rc = H.patmol_basis_load(len(q), q, xyz, len(npr), npr, lsh, cen, al, co, sys_.nelec, sys_.mult,
                             sys_.maxits, sys_.scf_eps, sys_.scf_a)
    if rc:
        raise PatmolError(f"patmol_basis_load failed ({rc})")


def set_options(tau: Optional[float] = None, precision: Optional[int] = None,
                mixed_threshold: float = 1e-6, strict: Optional[bool] = None, timing: Optional[bool] = None,
                scf_density: Optional[int] = None, scf_incremental: Optional[int] = None,
                md_engine: Optional[int] = None):
    """md_engine: classes whose FP64 items run on the McMurchie-Davidson
    kernel, a mask (1 (dd|dd), 2 (dd|dp), 4 (dp|dp), 8 (dd|pp); 0 none, -1 the
    library's default (dd|dd) and (dd|dp)); (dd|dp) stays on its Rys kernel
    for a basis whose (dd|dp) work items are mostly kept, unless 16 is added
    (EriStore.engine_mask tells what runs)."""
    _, G = lib()
    if md_engine is not None:
        G.patmol_hip_set_jk_engine(int(md_engine))
    if scf_density is not None:
        G.patmol_hip_set_scf_density(int(scf_density))
    if scf_incremental is not None:
        G.patmol_hip_set_scf_incremental(int(scf_incremental))
    if tau is not None:
        G.patmol_hip_set_schwarz_threshold(tau)
    if precision is not None:
        G.patmol_hip_set_precision(precision, mixed_threshold)
    if strict is not None:
        G.patmol_hip_set_strict(int(strict))
    if timing is not None:
        G.patmol_hip_set_timing(int(timing))


def one_electron_ints():
    """S, T, H (packed) of the current basis, computed on the GPU."""
    H, G = lib()
    n = H.number_of_orbitals()
    no2 = n * (n + 1) // 2
    S, T, Hm = np.zeros(no2), np.zeros(no2), np.zeros(no2)
    G.one_electron_ints(S, T, Hm)
    _check("one_electron_ints")
    return S, T, Hm


def host_one_electron_ints():
    """The same integrals from the host library's CPU code."""
    H, _ = lib()
    n = H.number_of_orbitals()
    no2 = n * (n + 1) // 2
    S, T, Hm = np.zeros(no2), np.zeros(no2), np.zeros(no2)
    H.patmol_host_one_electron_ints(S, T, Hm)
    return S, T, Hm


class EriStore:
    """An ERIFS handle of the current basis (the reference's integral store,
    here device-resident pair tables and screened work lists)."""

    def __init__(self, start: int = 0, name: str = "MINTS.DAT"):
        H, G = lib()
        self.norb = H.number_of_orbitals()
        self.nshell = H.number_of_shells()
        self.h = G.init_ERIFS(0, self.norb, name.encode())
        _check("init_ERIFS")
        G.two_electron_ints(self.h, start)
        _check("two_electron_ints")

    def rebuild(self, start: int = 0) -> None:
        """two_electron_ints again on this handle (after the basis moved,
        e.g. set_center_position of the DPM centre)."""
        _, G = lib()
        G.two_electron_ints(self.h, start)
        _check("two_electron_ints")

    def close(self):
        if getattr(self, "h", None) and _hip is not None:
            _hip.clean_ERIFS(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def rhf_gmatrix(self, P: np.ndarray, G: Optional[np.ndarray] = None) -> np.ndarray:
        _, L = lib()
        G = np.zeros_like(P) if G is None else G
        L.rhf_form_gmatrix(self.h, np.ascontiguousarray(P, dtype=np.float64), G)
        _check("rhf_form_gmatrix")
        return G

    def uhf_gmatrix(self, PA, PB):
        _, L = lib()
        GA, GB = np.zeros_like(PA), np.zeros_like(PB)
        L.uhf_form_gmatrix(self.h, np.ascontiguousarray(PA), np.ascontiguousarray(PB), GA, GB)
        _check("uhf_form_gmatrix")
        return GA, GB

    def rhf_gmatrix_device(self, d_P: int, d_G: int, stream: int = 0) -> None:
        _, L = lib()
        rc = L.patmol_hip_rhf_gmatrix_device(self.h, C.c_void_p(d_P), C.c_void_p(d_G), C.c_void_p(stream))
        if rc:
            _check("patmol_hip_rhf_gmatrix_device")

    def rhf_gmatrix_device_noalloc(self, d_P: int, d_G: int, stream: int = 0) -> None:
        """The allocation-free form of :meth:`rhf_gmatrix_device` (FP64,
        bases of at most NOALLOC_MAX_NORB functions)."""
        _, L = lib()
        rc = L.patmol_hip_rhf_gmatrix_device_noalloc(self.h, C.c_void_p(d_P), C.c_void_p(d_G), C.c_void_p(stream))
        if rc:
            _check("patmol_hip_rhf_gmatrix_device_noalloc")

    def stats(self) -> dict:
        _, L = lib()
        s = HipStats()
        if L.patmol_hip_stats(self.h, C.byref(s)):
            _check("patmol_hip_stats")
        return s.as_dict()

    def class_stats(self):
        """(kept quartets, algorithmic FLOPs) per class, arrays of 36."""
        _, L = lib()
        q = np.zeros(36, np.int64)
        f = np.zeros(36)
        if L.patmol_hip_class_stats(self.h, q, f):
            _check("patmol_hip_class_stats")
        return q, f

    @staticmethod
    def class_kernel_name(bra_pc: int, ket_pc: int, fp32: bool = False, md_mask: int = 3) -> str:
        """Symbol prefix (as profilers print it) of the J/K kernel that runs
        class (bra_pc | ket_pc); pair class pc = l(l+1)/2 + l'.  Mirrors the
        dispatch rule of launch_jk_t in eri_kernels_impl.h (md_mask: the
        classes on the McMurchie-Davidson kernel, the library's default)."""
        ls = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
        (la, lb), (lc, ld) = ls[bra_pc], ls[ket_pc]
        r = "float" if fp32 else "double"
        bit = {(2, 2, 2, 2): 1, (2, 2, 2, 1): 2, (2, 1, 2, 1): 4, (2, 2, 1, 1): 8}.get((la, lb, lc, ld), 0)
        if not fp32 and (bit & md_mask):
            return f"void pm::jk_md_kernel<{la}, {lb}, {lc}, {ld}>"
        if (la, lb, lc, ld) == (2, 2, 2, 2):
            return f"void pm::jk_kernel<{la}, {lb}, {lc}, {ld}, {r}>"
        return f"void pm::jk_kernel3<{la}, {lb}, {lc}, {ld}, {r}>"

    def engine_mask(self) -> int:
        """Classes whose FP64 items run on the McMurchie-Davidson kernel for
        this store (patmol_hip_jk_engine_mask; bits as set_options'
        md_engine), for class_kernel_name's md_mask."""
        _, L = lib()
        m = L.patmol_hip_jk_engine_mask(self.h)
        if m < 0:
            _check("patmol_hip_jk_engine_mask")
            raise PatmolError("patmol_hip_jk_engine_mask failed")
        return m

    def pair_schwarz(self) -> np.ndarray:
        _, L = lib()
        q = np.zeros(self.nshell * (self.nshell + 1) // 2)
        if L.patmol_hip_pair_schwarz(self.h, q):
            _check("patmol_hip_pair_schwarz")
        return q

    def work_items(self):
        _, L = lib()
        n = L.patmol_hip_work_items(self.h, None, None, None, None, None, None, None)
        if n < 0:
            _check("patmol_hip_work_items")
        t4 = np.zeros(4 * max(n, 1), np.int32)
        c2 = np.zeros(2 * max(n, 1), np.int32)
        b = np.zeros(max(n, 1))
        so = np.zeros(self.nshell, np.int32)
        tf = np.zeros(self.nshell, np.int32)
        ts = np.zeros(self.nshell, np.int32)
        nt = C.c_int()
        L.patmol_hip_work_items(self.h, t4.ctypes.data, c2.ctypes.data, b.ctypes.data, so.ctypes.data,
                                tf.ctypes.data, ts.ctypes.data, C.addressof(nt))
        k = nt.value
        return dict(tiles=t4[:4 * n].reshape(n, 4), classes=c2[:2 * n].reshape(n, 2), bounds=b[:n],
                    sorted_orig=so, tile_first=tf[:k], tile_size=ts[:k])

    def eri_blocks(self, quartets: Sequence[Sequence[int]], nst: Sequence[int], fp32: bool = False):
        """Contracted ERI blocks of the listed shell quartets (original shell
        indices); fp32: in the FP32 arithmetic of the mixed J/K kernels."""
        _, L = lib()
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.int32).reshape(-1))
        nq = len(q) // 4
        sizes = [int(nst[a] * nst[b] * nst[c] * nst[d]) for a, b, c, d in np.asarray(quartets).reshape(-1, 4)]
        out = np.zeros(max(sum(sizes), 1))
        if L.patmol_hip_eri_blocks_prec(self.h, nq, q, out, int(fp32)):
            _check("patmol_hip_eri_blocks_prec")
        res, o = [], 0
        for (a, b, c, d), s in zip(np.asarray(quartets).reshape(-1, 4), sizes):
            res.append(out[o:o + s].reshape(nst[a], nst[b], nst[c], nst[d]))
            o += s
        return res

    def write_erints(self, path: str, start: int = 0) -> int:
        """Every canonical integral with |v| > 1e-14 as ERINT records in the
        reference's order; returns the record count."""
        _, L = lib()
        n = L.patmol_hip_write_erints(self.h, start, path.encode())
        if n < 0:
            _check("patmol_hip_write_erints")
        return n

    def last_timing(self):
        _, L = lib()
        tot = C.c_double()
        cls = np.zeros(36)
        L.patmol_hip_last_timing(self.h, C.byref(tot), cls)
        return tot.value, cls


def rhf_scf():
    """Run the SCF driver on the current basis; returns (result, P, orbital energies)."""
    H, _ = lib()
    n = H.number_of_orbitals()
    P = np.zeros(n * (n + 1) // 2)
    ev = np.zeros(n)
    res = ScfResult()
    rc = H.patmol_rhf_scf(P, ev, C.byref(res))
    if rc:
        _check("patmol_rhf_scf")
        raise PatmolError(f"patmol_rhf_scf failed ({rc})")
    return res.as_dict(), P, ev


def uhf_scf():
    H, _ = lib()
    n = H.number_of_orbitals()
    PA, PB = np.zeros(n * (n + 1) // 2), np.zeros(n * (n + 1) // 2)
    res = ScfResult()
    rc = H.patmol_uhf_scf(PA, PB, C.byref(res))
    if rc:
        _check("patmol_uhf_scf")
        raise PatmolError(f"patmol_uhf_scf failed ({rc})")
    return res.as_dict(), PA, PB


def _ptr(a):
    return None if a is None else a.ctypes.data


def _scf_inputs():
    H, _ = lib()
    return dict(n=H.number_of_orbitals(), ne=H.number_of_electrons(), mult=H.spin_multiplicity(),
                maxits=H.max_scf_iterations(), eps=H.scf_convergence(), accel=H.scf_acceleration(),
                enuc=H.nuclear_repulsion())


def rhf_scf_device(store: Optional["EriStore"] = None, guess: Optional[np.ndarray] = None,
                   want_c: bool = False, dpm: bool = False):
    """The whole RHF iteration on the GPU (patmol_hip_rhf_scf) for the
    current basis; returns (result dict, P, orbital energies[, C]).  With
    dpm, H carries the DPM potential of the positron at skip_center()
    (dpm_ints), as one grid point of the DPM loop has it."""
    _, G = lib()
    a = _scf_inputs()
    n = a["n"]
    S, T, Hm = one_electron_ints()
    if dpm:
        G.dpm_ints(Hm)
        _check("dpm_ints")
    own = store is None
    st = EriStore() if own else store
    P = np.zeros(n * (n + 1) // 2)
    ev = np.zeros(n)
    Cm = np.zeros((n, n)) if want_c else None
    g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64)
    res = HipScfResult()
    rc = G.patmol_hip_rhf_scf(st.h, None, S, Hm, a["ne"] // 2, a["maxits"], a["eps"], a["accel"], a["enuc"],
                              _ptr(g), _ptr(P), _ptr(ev), _ptr(Cm), C.byref(res))
    if own:
        st.close()
    if rc:
        _check("patmol_hip_rhf_scf")
        raise PatmolError(f"patmol_hip_rhf_scf failed ({rc})")
    out = (res.as_dict(), P, ev)
    return out + (Cm,) if want_c else out


def uhf_scf_device(store: Optional["EriStore"] = None):
    """The whole UHF iteration on the GPU (patmol_hip_uhf_scf); returns
    (result dict, PA, PB, alpha and beta orbital energies)."""
    _, G = lib()
    a = _scf_inputs()
    n = a["n"]
    noccb = (a["ne"] - (a["mult"] - 1)) // 2
    nocca = a["ne"] - noccb
    S, T, Hm = one_electron_ints()
    own = store is None
    st = EriStore() if own else store
    PA, PB = np.zeros(n * (n + 1) // 2), np.zeros(n * (n + 1) // 2)
    eA, eB = np.zeros(n), np.zeros(n)
    res = HipScfResult()
    rc = G.patmol_hip_uhf_scf(st.h, None, S, Hm, nocca, noccb, a["maxits"], a["eps"], a["accel"], a["enuc"],
                              None, None, _ptr(PA), _ptr(PB), _ptr(eA), _ptr(eB), C.byref(res))
    if own:
        st.close()
    if rc:
        _check("patmol_hip_uhf_scf")
        raise PatmolError(f"patmol_hip_uhf_scf failed ({rc})")
    return res.as_dict(), PA, PB, eA, eB


def diag_fock(S: np.ndarray, F: np.ndarray, nocc: int):
    """One device diagonalisation of packed F in the basis orthogonalised by
    packed S: (eigenvalues, C row-major, packed P)."""
    _, G = lib()
    n = int(round((np.sqrt(8 * len(S) + 1) - 1) / 2))
    ev, Cm, P = np.zeros(n), np.zeros((n, n)), np.zeros(len(S))
    rc = G.patmol_hip_diag_fock(np.ascontiguousarray(S), np.ascontiguousarray(F), nocc, _ptr(ev), _ptr(Cm), _ptr(P))
    if rc:
        _check("patmol_hip_diag_fock")
    return ev, Cm, P


def moment_ints():
    """Dipole (3, no2) and quadrupole (6, no2) integrals of the current basis
    about the origin, computed on the GPU."""
    H, G = lib()
    n = H.number_of_orbitals()
    no2 = n * (n + 1) // 2
    d, q = np.zeros(3 * no2), np.zeros(6 * no2)
    if G.patmol_hip_moment_ints(d, q):
        _check("patmol_hip_moment_ints")
    return d.reshape(3, no2), q.reshape(6, no2)


MOMINTS_DTYPE = np.dtype([("ir", "<i4"), ("jr", "<i4"), ("dip", "<f8", 3), ("quad", "<f8", 6)])


def write_momints(path: str) -> None:
    _, G = lib()
    if G.patmol_hip_write_momints(path.encode()):
        _check("patmol_hip_write_momints")


def rhf_moments(P: np.ndarray):
    """rhf_analyzeMoments: (electronic, nuclear, total) moments, 9 values each
    (x, y, z, xx, xy, xz, yy, yz, zz), for half density P."""
    H, _ = lib()
    e, n, t = np.zeros(9), np.zeros(9), np.zeros(9)
    if H.patmol_rhf_moments(np.ascontiguousarray(P, dtype=np.float64), e, n, t):
        _check("patmol_rhf_moments")
        raise PatmolError("patmol_rhf_moments failed")
    return e, n, t


def uhf_moments(PA: np.ndarray, PB: np.ndarray):
    H, _ = lib()
    e, n, t = np.zeros(9), np.zeros(9), np.zeros(9)
    if H.patmol_uhf_moments(np.ascontiguousarray(PA, dtype=np.float64), np.ascontiguousarray(PB, dtype=np.float64),
                            e, n, t):
        _check("patmol_uhf_moments")
        raise PatmolError("patmol_uhf_moments failed")
    return e, n, t


def rhf_finite_field(efx: float):
    """rhf_finite_field_calculation on the device SCF: (alfa[5], ground-state
    result); alfa = xx, yy, zz, mean, and the reference's alfa2."""
    H, _ = lib()
    a = np.zeros(5)
    gs = ScfResult()
    rc = H.patmol_rhf_finite_field(efx, a, C.byref(gs))
    if rc:
        _check("patmol_rhf_finite_field")
        raise PatmolError(f"patmol_rhf_finite_field failed ({rc})")
    return a, gs.as_dict()


ERINT_DTYPE = np.dtype({"names": ["val", "i", "j", "k", "l"],
                        "formats": ["<f8", "<u2", "<u2", "<u2", "<u2"],
                        "offsets": [0, 8, 10, 12, 14], "itemsize": 16})


def _records(recs):
    r = np.ascontiguousarray(recs)
    if r.dtype != ERINT_DTYPE:
        raise PatmolError("ERINT records must have dtype patmol_amd.ERINT_DTYPE")
    return r


def rhf_gmatrix_erints(recs, P, G=None, path: Optional[str] = None):
    """G += 2J[P] - K[P] from ERINT records (an array of ERINT_DTYPE) or from
    a record file at `path`, on the GPU."""
    _, L = lib()
    P = np.ascontiguousarray(P, dtype=np.float64)
    G = np.zeros_like(P) if G is None else G
    if path is not None:
        rc = L.patmol_hip_rhf_gmatrix_erint_file(path.encode(), P, G)
    else:
        r = _records(recs)
        rc = L.patmol_hip_rhf_gmatrix_erints(r.ctypes.data, len(r), P, G)
    if rc:
        _check("patmol_hip_rhf_gmatrix_erints")
    return G


def uhf_gmatrix_erints(recs, PA, PB, path: Optional[str] = None):
    _, L = lib()
    PA = np.ascontiguousarray(PA, dtype=np.float64)
    PB = np.ascontiguousarray(PB, dtype=np.float64)
    GA, GB = np.zeros_like(PA), np.zeros_like(PB)
    if path is not None:
        rc = L.patmol_hip_uhf_gmatrix_erint_file(path.encode(), PA, PB, GA, GB)
    else:
        r = _records(recs)
        rc = L.patmol_hip_uhf_gmatrix_erints(r.ctypes.data, len(r), PA, PB, GA, GB)
    if rc:
        _check("patmol_hip_uhf_gmatrix_erints")
    return GA, GB


class DpmPoint(C.Structure):
    _fields_ = [("vpol", C.c_double), ("alfa", C.c_double), ("e_first", C.c_double), ("e_conv", C.c_double),
                ("iterations", C.c_int), ("converged", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def add_dpm_center() -> None:
    """Add the DPM centre (charge +1) and its 7 shells to the current basis
    (basis.c:220-241, 280-284); they take part once dpm_calculation (or
    dpm_augment) makes them visible."""
    H, _ = lib()
    if H.patmol_basis_add_dpm_center():
        raise PatmolError("patmol_basis_add_dpm_center failed")


def dpm_calculation(points, uhf: bool = False):
    """The DPM grid loop (rhf_dpm_calculation, rhfscf.c:217-263, or its UHF
    form) on the device SCF for the current basis, whose DPM centre has been
    added (:func:`add_dpm_center`) but not made visible.  points: (npts, 3)
    positron positions in bohr.  Returns (per-point dicts with vpol, alfa,
    e_first, e_conv, iterations, converged; ground-state result).  The basis
    is left augmented, as the reference leaves it."""
    H, _ = lib()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    out = (DpmPoint * max(1, len(pts)))()
    gs = ScfResult()
    rc = H.patmol_dpm_calculation(int(uhf), len(pts), pts.reshape(-1), C.addressof(out), C.byref(gs))
    if rc:
        _check("patmol_dpm_calculation")
        raise PatmolError(f"patmol_dpm_calculation failed ({rc})")
    return [out[i].as_dict() for i in range(len(pts))], gs.as_dict()


def rhf_gradient(P: np.ndarray, store: Optional["EriStore"] = None) -> np.ndarray:
    """dE/dR (ncen x 3, hartree/bohr) of the RHF energy at the converged half
    density P of the current basis: two-electron part on the GPU over
    `store`'s screened quartets (or a store built for the call), one-electron
    and nuclear repulsion parts on the host with W = P F P."""
    H, _ = lib()
    nc = H.number_of_centers()
    g = np.zeros(3 * nc)
    rc = H.patmol_rhf_gradient(None if store is None else store.h, np.ascontiguousarray(P, dtype=np.float64), g)
    if rc:
        _check("patmol_rhf_gradient")
        raise PatmolError(f"patmol_rhf_gradient failed ({rc})")
    return g.reshape(nc, 3)


def rhf_dpm_gradient(P: np.ndarray, store: Optional["EriStore"] = None) -> np.ndarray:
    """dE/dR (ncen x 3) of the DPM-augmented RHF energy of one grid point
    (basis after dpm_augment, positron at skip_center; its row is dE/dQ) at
    the converged half density P: rhf_gradient's terms with H - V and the
    DPM potential's derivative (patmol_hip_dpm_gradient).  store must cover
    the whole augmented basis, or None."""
    H, _ = lib()
    nc = H.number_of_centers()
    g = np.zeros(3 * nc)
    rc = H.patmol_rhf_dpm_gradient(None if store is None else store.h, np.ascontiguousarray(P, dtype=np.float64), g)
    if rc:
        _check("patmol_rhf_dpm_gradient")
        raise PatmolError(f"patmol_rhf_dpm_gradient failed ({rc})")
    return g.reshape(nc, 3)


def dpm_gradient(P: np.ndarray) -> np.ndarray:
    """The DPM potential's part alone (patmol_hip_dpm_gradient): d/dR of
    -2 sum_ab P_ab V_ab, ncen x 3."""
    H, G = lib()
    nc = H.number_of_centers()
    g = np.zeros(3 * nc)
    if G.patmol_hip_dpm_gradient(np.ascontiguousarray(P, dtype=np.float64), g):
        _check("patmol_hip_dpm_gradient")
    return g.reshape(nc, 3)


def eri_gradient(P: np.ndarray, store: "EriStore") -> np.ndarray:
    """The two-electron part alone (patmol_hip_eri_gradient), ncen x 3."""
    H, G = lib()
    nc = H.number_of_centers()
    g = np.zeros(3 * nc)
    if G.patmol_hip_eri_gradient(store.h, np.ascontiguousarray(P, dtype=np.float64), g):
        _check("patmol_hip_eri_gradient")
    return g.reshape(nc, 3)


def onee_gradient(P: np.ndarray, W: np.ndarray) -> np.ndarray:
    """The one-electron part 2 sum P dH/dR - 2 sum W dS/dR on the GPU
    (patmol_hip_onee_gradient), ncen x 3."""
    H, G = lib()
    nc = H.number_of_centers()
    g = np.zeros(3 * nc)
    if G.patmol_hip_onee_gradient(np.ascontiguousarray(P, dtype=np.float64), np.ascontiguousarray(W, dtype=np.float64), g):
        _check("patmol_hip_onee_gradient")
    return g.reshape(nc, 3)


def host_onee_gradient(P: np.ndarray, W: np.ndarray) -> np.ndarray:
    """2 sum P dH/dR - 2 sum W dS/dR (packed P, W) for the current basis,
    computed on the host, ncen x 3."""
    H, _ = lib()
    nc = H.number_of_centers()
    g = np.zeros(3 * nc)
    if H.patmol_host_onee_gradient(np.ascontiguousarray(P, dtype=np.float64), np.ascontiguousarray(W, dtype=np.float64), g):
        raise PatmolError("patmol_host_onee_gradient failed")
    return g.reshape(nc, 3)


def nuclear_repulsion_gradient() -> np.ndarray:
    H, _ = lib()
    nc = H.number_of_centers()
    g = np.zeros(3 * nc)
    H.patmol_nuclear_repulsion_gradient(g)
    return g.reshape(nc, 3)


class Batch:
    """Many independent molecules in one G build (patmol_hip_batch_*): add
    each system (its basis is loaded, its S and H computed on the GPU and its
    screened tables snapshotted), finalize (one work queue per class, FP32
    split, the G build captured as a HIP graph), then G builds or the RHF of
    every molecule iterating together."""

    def __init__(self):
        _, G = lib()
        self.h = G.patmol_hip_batch_new()
        self.systems, self.S, self.H = [], [], []

    def add(self, system) -> int:
        _, G = lib()
        load_system(system)
        S, _, Hm = one_electron_ints()
        idx = G.patmol_hip_batch_add(self.h)
        if idx < 0:
            _check("patmol_hip_batch_add")
        self.systems.append(system)
        self.S.append(S)
        self.H.append(Hm)
        return idx

    def finalize(self, precision: int = 0, threshold: float = 1e-6, dmax: float = 1.0) -> None:
        _, G = lib()
        if G.patmol_hip_batch_finalize(self.h, int(precision), threshold, dmax):
            _check("patmol_hip_batch_finalize")

    def offsets(self):
        o = [0]
        for s in self.systems:
            o.append(o[-1] + s.norb * (s.norb + 1) // 2)
        return o

    def gmatrix(self, P, use_graph: bool = True) -> np.ndarray:
        _, L = lib()
        P = np.ascontiguousarray(P, dtype=np.float64)
        G = np.zeros_like(P)
        if L.patmol_hip_batch_gmatrix(self.h, P, G, int(use_graph)):
            _check("patmol_hip_batch_gmatrix")
        return G

    def buffers(self):
        _, L = lib()
        p, g = C.c_void_p(), C.c_void_p()
        if L.patmol_hip_batch_buffers(self.h, C.byref(p), C.byref(g)):
            _check("patmol_hip_batch_buffers")
        return p.value, g.value

    def launch(self, stream: int = 0) -> None:
        _, L = lib()
        if L.patmol_hip_batch_launch(self.h, C.c_void_p(stream)):
            _check("patmol_hip_batch_launch")

    def info(self) -> dict:
        _, L = lib()
        nm, n2, it, i32 = C.c_int(), C.c_int64(), C.c_int64(), C.c_int64()
        L.patmol_hip_batch_info(self.h, C.byref(nm), C.byref(n2), C.byref(it), C.byref(i32))
        return {"molecules": nm.value, "no2_total": n2.value, "work_items": it.value, "work_items_fp32": i32.value}

    def rhf_scf(self, maxits: int = 200, eps: float = 1e-10, accel: int = 3):
        """(energies, iterations, converged, P concatenated, total ms, G-build ms)."""
        _, L = lib()
        nm = len(self.systems)
        S, Hm = np.concatenate(self.S), np.concatenate(self.H)
        nocc = np.array([s.nelec // 2 for s in self.systems], dtype=np.int32)
        enuc = np.array([s.nuclear_repulsion() for s in self.systems])
        e, it, cv = np.zeros(nm), np.zeros(nm, np.int32), np.zeros(nm, np.int32)
        P = np.zeros(len(S))
        tot, gms = C.c_double(), C.c_double()
        rc = L.patmol_hip_batch_rhf_scf(self.h, S, Hm, nocc, enuc, maxits, eps, accel, e, it, cv, P.ctypes.data,
                                        C.byref(tot), C.byref(gms))
        if rc:
            _check("patmol_hip_batch_rhf_scf")
            raise PatmolError(f"patmol_hip_batch_rhf_scf failed ({rc})")
        return e, it, cv, P, tot.value, gms.value

    def close(self):
        if getattr(self, "h", None) and _hip is not None:
            _hip.patmol_hip_batch_free(self.h)
            self.h = None

    def __del__(self):
        self.close()
