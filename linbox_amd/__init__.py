"""linbox_amd: MI355X (gfx950) FFT polynomial-matrix multiplication over Z/pZ.

Python binding of the C ABI in include/lbx/lbx.h (liblbx.so, built by
linbox_amd/build_lib.py).  The host-side classes in linbox_amd.polymat mirror
LinBox's PolynomialMatrix / PolynomialMatrixMulDomain interface; everything
computes on the GPU through the C ABI.  There is no CPU fallback: loading
fails loudly when the library is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbx.so")

LBX_OK = 0
STATUS_NAMES = {
    0: "LBX_OK", 1: "LBX_ERR_BAD_FFT_PRIME", 2: "LBX_ERR_K_TOO_LARGE", 3: "LBX_ERR_K_NOT_POSITIVE",
    4: "LBX_ERR_NOT_PRIMITIVE_ROOT", 5: "LBX_ERR_SIZE_MISMATCH", 6: "LBX_ERR_NOT_ENOUGH_PRIMES",
    7: "LBX_ERR_INVALID_ARG", 8: "LBX_ERR_PRIME_TOO_LARGE", 9: "LBX_ERR_HIP", 10: "LBX_ERR_OUT_OF_MEMORY",
    11: "LBX_ERR_UNSUPPORTED", 12: "LBX_ERR_NO_DEVICE", 13: "LBX_ERR_WORKSPACE_TOO_SMALL",
    14: "LBX_ERR_CHECK_FAILED",
}
POLFIRST, MATFIRST, MATROWFIRST = 0, 1, 2
METHOD_AUTO, METHOD_PLAIN_U64, METHOD_INT8_MFMA = 0, 1, 2

# Every symbol include/lbx/lbx.h declares.
EXPORTED = [
    "lbx_init", "lbx_destroy", "lbx_strerror", "lbx_trim", "lbx_polymat_mul_u32", "lbx_polymat_mul_fft_u32",
    "lbx_ntt_forward_u32", "lbx_ntt_inverse_u32", "lbx_polymat_mul_mp", "lbx_polymat_mul_batch",
    "lbx_polymat_midproduct_u32", "lbx_polymat_midproduct_naive_u32", "lbx_order_basis_pm_u32",
    "lbx_order_basis_m_u32", "lbx_polymat_copy_u32", "lbx_profile_enable", "lbx_profile_get",
    "lbx_polymat_mul_u32_noalloc", "lbx_batch_tuning", "lbx_polymat_check_mul_u32", "lbx_set_check",
    "lbx_polymat_check_midproduct_u32",
    "lbx_polymat_mul_zz",
]


class LinboxError(RuntimeError):
    """Raised for a non-zero lbx_status; .status holds the code."""

    def __init__(self, status, message):
        super().__init__("%s (%s)" % (message, STATUS_NAMES.get(status, status)))
        self.status = status


class Polmat(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("layout", ctypes.c_int), ("rows", ctypes.c_size_t),
                ("cols", ctypes.c_size_t), ("size", ctypes.c_size_t), ("storage_size", ctypes.c_size_t),
                ("shift", ctypes.c_size_t)]


class Job(ctypes.Structure):
    _fields_ = [("p", ctypes.c_uint32), ("m", ctypes.c_size_t), ("k", ctypes.c_size_t), ("n", ctypes.c_size_t),
                ("sa", ctypes.c_size_t), ("sb", ctypes.c_size_t), ("a", ctypes.c_void_p), ("b", ctypes.c_void_p),
                ("c", ctypes.c_void_p)]


_lib = None


def build():
    import importlib
    return importlib.import_module(__name__ + ".build_lib").build()


def lib():
    """Load liblbx.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("linbox_amd: %s is missing; run linbox_amd/build_lib.py (no CPU fallback exists)"
                               % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, u32, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int
        PM = ctypes.POINTER(Polmat)
        L.lbx_init.argtypes = [ctypes.POINTER(vp), i32]
        L.lbx_destroy.argtypes = [vp]
        L.lbx_trim.argtypes = [vp]
        L.lbx_strerror.argtypes = [i32]
        L.lbx_strerror.restype = ctypes.c_char_p
        L.lbx_polymat_mul_u32.argtypes = [vp, u32, PM, PM, PM, sz, i32, vp]
        L.lbx_polymat_mul_u32_noalloc.argtypes = [vp, u32, PM, PM, PM, sz, i32, vp]
        L.lbx_polymat_mul_fft_u32.argtypes = [vp, u32, ctypes.c_uint, sz, sz, sz, vp, vp, vp, i32, vp]
        L.lbx_ntt_forward_u32.argtypes = [vp, u32, ctypes.c_uint, u32, vp, sz, ctypes.POINTER(u32), vp]
        L.lbx_ntt_inverse_u32.argtypes = [vp, u32, ctypes.c_uint, u32, vp, sz, ctypes.POINTER(u32), vp]
        L.lbx_polymat_mul_mp.argtypes = [vp, vp, i32, PM, PM, PM, sz, i32, vp]
        L.lbx_polymat_mul_batch.argtypes = [vp, ctypes.POINTER(Job), sz, i32, vp]
        L.lbx_polymat_midproduct_u32.argtypes = [vp, u32, PM, PM, PM, i32, sz, sz, i32, vp]
        L.lbx_polymat_midproduct_naive_u32.argtypes = [vp, u32, PM, PM, PM, sz, sz, vp]
        SZP = ctypes.POINTER(sz)
        L.lbx_order_basis_pm_u32.argtypes = [vp, u32, PM, PM, sz, SZP, SZP, i32, vp]
        L.lbx_order_basis_m_u32.argtypes = [vp, u32, PM, PM, sz, SZP, SZP, vp]
        L.lbx_polymat_copy_u32.argtypes = [vp, PM, PM, sz, sz, sz, vp]
        L.lbx_profile_enable.argtypes = [vp, i32]
        L.lbx_profile_get.argtypes = [vp, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_float), i32]
        L.lbx_profile_get.restype = i32
        L.lbx_batch_tuning.argtypes = [vp]
        L.lbx_polymat_check_mul_u32.argtypes = [vp, u32, PM, PM, PM, sz, u32, ctypes.POINTER(i32), vp]
        L.lbx_polymat_check_midproduct_u32.argtypes = [vp, u32, PM, PM, PM, sz, sz, ctypes.POINTER(i32), vp]
        L.lbx_set_check.argtypes = [vp, i32]
        L.lbx_polymat_mul_zz.argtypes = [vp, PM, i32, PM, PM, i32, sz, i32, vp]
        L.lbx_batch_tuning.restype = ctypes.c_char_p
        for name in EXPORTED:
            if name not in ("lbx_strerror", "lbx_profile_get", "lbx_batch_tuning"):
                getattr(L, name).restype = i32
        _lib = L
    return _lib


def check(status):
    if status != LBX_OK:
        raise LinboxError(status, lib().lbx_strerror(status).decode())


class Context:
    """An lbx_ctx on one device (default 0)."""

    def __init__(self, device=0):
        self.handle = ctypes.c_void_p()
        check(lib().lbx_init(ctypes.byref(self.handle), device))

    def close(self):
        if self.handle:
            lib().lbx_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trim(self):
        """Release the cached device workspace and the pinned host buffers (lbx_trim)."""
        check(lib().lbx_trim(self.handle))

    def profile(self, enable=True):
        """enable: False/0 off, True/1 stage times, 2 stage and per-kernel times."""
        check(lib().lbx_profile_enable(self.handle, int(enable)))

    def set_check(self, on=True):
        """CHECK_MATPOL_MUL as a switch: check every whole product at a random point
        (a mismatch raises LBX_ERR_CHECK_FAILED)."""
        check(lib().lbx_set_check(self.handle, 1 if on else 0))

    def stage_times(self):
        labels = (ctypes.c_char_p * 32)()
        ms = (ctypes.c_float * 32)()
        n = lib().lbx_profile_get(self.handle, labels, ms, 32)
        return [(labels[i].decode(), ms[i]) for i in range(n)]


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


from .polymat import (PolynomialMatrix, PolynomialMatrixMulDomain, PolynomialMatrixFFTMulDomain,  # noqa: E402
                      PolynomialMatrixFFTPrimeMulDomain, PolynomialMatrixNaiveMulDomain,
                      PolynomialMatrixThreePrimesFFTMulDomain, FFT, ntt_forward, ntt_inverse, mul_mp, mul_batch,
                      OrderBasis, check_mul, check_midproduct, mul_zz)

__all__ = ["Context", "LinboxError", "PolynomialMatrix", "PolynomialMatrixMulDomain", "PolynomialMatrixFFTMulDomain",
           "PolynomialMatrixFFTPrimeMulDomain", "PolynomialMatrixNaiveMulDomain",
           "PolynomialMatrixThreePrimesFFTMulDomain", "FFT", "ntt_forward", "ntt_inverse", "mul_mp", "mul_batch", "OrderBasis", "check_mul", "check_midproduct", "mul_zz",
           "POLFIRST", "MATFIRST", "MATROWFIRST", "METHOD_AUTO", "METHOD_PLAIN_U64", "METHOD_INT8_MFMA", "lib",
           "build", "default_context", "EXPORTED"]
