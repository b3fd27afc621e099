"""Device buffers and thin wrappers over the extended C ABI.

Host data is numpy (column-major arrays are passed as Fortran-ordered numpy
arrays or flat vectors); device memory comes from rt_malloc.  No torch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib

PREC = {"fp64": 0, "split_fp32": 1, "bf16x3": 2, "ozaki_bf16": 3, "ozaki_int8": 4}


def _as_c_vector(a: np.ndarray, dtype=np.float64) -> np.ndarray:
    """Flat view in column-major (Fortran) element order, contiguous."""
    a = np.asarray(a, dtype=dtype)
    return np.ascontiguousarray(a.reshape(-1, order="F"))


class DeviceArray:
    """A flat device allocation of `n` float64 (or other dtype) elements."""

    def __init__(self, n: int, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().rt_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value or 0

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=np.float64) -> "DeviceArray":
        v = _as_c_vector(a, dtype)
        d = cls(v.size, dtype)
        if v.size:
            check(lib().rt_memcpy_h2d(d.ptr, v.ctypes.data, d.nbytes))
        return d

    def to_host(self, shape=None) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        if self.n:
            check(lib().rt_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        if shape is not None:
            out = out.reshape(shape, order="F")
        return out

    def zero(self):
        if self.nbytes:
            check(lib().rt_memset(self.ptr, 0, self.nbytes))
        return self

    def free(self):
        if self.ptr:
            check(lib().rt_free(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            if self.ptr:
                lib().rt_free(self.ptr)
                self.ptr = 0
        except Exception:
            pass

    @property
    def vp(self):
        return C.c_void_p(self.ptr)


def synchronize():
    check(lib().rt_synchronize())


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(lib().rt_device_name(buf, 256))
    return buf.value.decode()


def device_count() -> int:
    n = C.c_int(0)
    check(lib().rt_device_count(C.byref(n)))
    return n.value


class Stream:
    def __init__(self):
        p = C.c_void_p()
        check(lib().rt_stream_create(C.byref(p)))
        self.handle = p.value

    def synchronize(self):
        check(lib().rt_stream_synchronize(self.handle))

    def destroy(self):
        if self.handle:
            check(lib().rt_stream_destroy(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().rt_event_create(C.byref(p)))
        self.handle = p.value

    def record(self, stream=None):
        check(lib().rt_event_record(self.handle, stream))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0)
        check(lib().rt_event_elapsed_ms(self.handle, end.handle, C.byref(ms)))
        return float(ms.value)


def kernel_timer(on: bool):
    check(lib().rt_kernel_timer_enable(1 if on else 0))


def kernel_timer_reset():
    check(lib().rt_kernel_timer_reset())


def kernel_time(prefix: str):
    tot = C.c_double(0)
    cnt = C.c_int64(0)
    check(lib().rt_kernel_timer_get(prefix.encode(), C.byref(tot), C.byref(cnt)))
    return tot.value, cnt.value


# ---- generators --------------------------------------------------------------
def gen_ri3ao(d: DeviceArray, nbas: int, naux: int, seed: int, decay: float = 16.0,
              stream=None):
    assert d.n >= nbas * nbas * naux
    check(lib().rt_gen_ri3ao_dev(d.ptr, nbas, naux, seed, decay, stream))


def gen_normal(d: DeviceArray, n: int, seed: int, stream=None):
    assert d.n >= n
    check(lib().rt_gen_normal_dev(d.ptr, n, seed, stream))


# ---- ops -----------------------------------------------------------------------
def ao2mo(C_: DeviceArray, ri3ao: DeviceArray, ri3mo: DeviceArray, ns, nbas, naux,
          stream=None):
    assert C_.n >= nbas * ns and ri3ao.n >= nbas * nbas * naux and ri3mo.n >= naux * ns * ns
    check(lib().rt_ao2mo_dev(C_.ptr, ri3ao.ptr, ri3mo.ptr, ns, nbas, naux, stream))


def ao2mo_occ(C_: DeviceArray, nbas, nmo, nocc, ri3ao: DeviceArray, naux, B: DeviceArray,
              precision="fp64", stream=None):
    assert C_.n >= nbas * nmo and ri3ao.n >= nbas * nbas * naux and B.n >= nocc * nmo * naux
    check(lib().rt_ao2mo_occ_dev(C_.ptr, nbas, nmo, nocc, ri3ao.ptr, naux, B.ptr,
                                 PREC[precision], stream))


def density(C_: DeviceArray, nbas, nocc, D: DeviceArray, stream=None):
    assert C_.n >= nbas * nocc and D.n >= nbas * nbas
    check(lib().rt_density_dev(C_.ptr, nbas, nocc, D.ptr, stream))


def rij(ri3ao: DeviceArray, nbas, naux, D: DeviceArray, J: DeviceArray,
        gamma: DeviceArray | None = None, want_energy=True, stream=None):
    assert ri3ao.n >= nbas * nbas * naux and D.n >= nbas * nbas and J.n >= nbas * nbas
    e = C.c_double(0)
    check(lib().rt_rij_dev(ri3ao.ptr, nbas, naux, D.ptr, J.ptr,
                           gamma.ptr if gamma is not None else None,
                           C.byref(e) if want_energy else None, stream))
    return e.value if want_energy else None


def rik(B: DeviceArray, nocc, nmo, naux, K: DeviceArray, precision="fp64",
        want_energy=True, stream=None):
    assert B.n >= nocc * nmo * naux and K.n >= nmo * nmo
    e = C.c_double(0)
    check(lib().rt_rik_dev(B.ptr, nocc, nmo, naux, K.ptr,
                           C.byref(e) if want_energy else None, PREC[precision], stream))
    return e.value if want_energy else None


def exchange_energy(B: DeviceArray, nocc, nmo, naux, stream=None) -> float:
    e = C.c_double(0)
    check(lib().rt_exchange_energy_dev(B.ptr, nocc, nmo, naux, C.byref(e), stream))
    return e.value


def upper_unpack(packed: DeviceArray, packed_len, full: DeviceArray, stream=None):
    check(lib().rt_upper_unpack_dev(packed.ptr, packed_len, full.ptr, stream))


def upper_pack(full: DeviceArray, n, packed: DeviceArray, stream=None):
    check(lib().rt_upper_pack_dev(full.ptr, n, packed.ptr, stream))


def rifull_to_matfull_symm(ri: DeviceArray, nao, naux, out: DeviceArray, stream=None):
    check(lib().rt_rifull_to_matfull_symm_dev(ri.ptr, nao, naux, out.ptr, stream))


def upper_dim(packed_len: int) -> int:
    return int(lib().rt_upper_dim(packed_len))


def dgemm(opa, opb, m, n, k, alpha, A: DeviceArray, lda, B: DeviceArray, ldb, beta,
          Cm: DeviceArray, ldc, stream=None):
    check(lib().rt_dgemm_dev(opa.encode(), opb.encode(), m, n, k, alpha, A.ptr, lda,
                             B.ptr, ldb, beta, Cm.ptr, ldc, stream))


# ---- level-2 / rank-k primitives (A7) -------------------------------------------
def dgemv(trans, m, n, alpha, A: DeviceArray, lda, x: DeviceArray, beta, y: DeviceArray,
          incx=1, incy=1, stream=None):
    check(lib().rt_dgemv_dev(trans.encode(), m, n, alpha, A.ptr, lda, x.ptr, incx, beta,
                             y.ptr, incy, stream))


def dsyrk(uplo, trans, n, k, alpha, A: DeviceArray, lda, beta, Cm: DeviceArray, ldc,
          stream=None):
    check(lib().rt_dsyrk_dev(uplo.encode(), trans.encode(), n, k, alpha, A.ptr, lda, beta,
                             Cm.ptr, ldc, stream))


def einsum_ip_ip_p(A: DeviceArray, a_rows, a_cols, B: DeviceArray, b_rows, b_cols,
                   out: DeviceArray, stream=None):
    check(lib().rt_einsum_ip_ip_p_dev(A.ptr, a_rows, a_cols, B.ptr, b_rows, b_cols,
                                      out.ptr, stream))


# ---- FP64 aux-metric functions (A11) ---------------------------------------------
def dpotrf(uplo, n, A: DeviceArray, lda, stream=None) -> int:
    """Returns LAPACK's info (0 = success)."""
    info = C.c_int(0)
    check(lib().rt_dpotrf_dev(uplo.encode(), n, A.ptr, lda, C.byref(info), stream))
    return info.value


def cholesky_decompose_inverse(uplo, n, A: DeviceArray, lda, out: DeviceArray, ldo,
                               stream=None):
    check(lib().rt_cholesky_decompose_inverse_dev(uplo.encode(), n, A.ptr, lda, out.ptr,
                                                  ldo, stream))


def matrix_power(n, A: DeviceArray, lda, p, threshold, out: DeviceArray, ldo, stream=None):
    check(lib().rt_matrix_power_dev(n, A.ptr, lda, p, threshold, out.ptr, ldo, stream))


def special_dgemm(T: DeviceArray, xa, ya, za, sx, lx, sz, lz, B: DeviceArray, rb, cb, srb,
                  lrb, scb, lcb, alpha, beta, stream=None):
    """special_dgemm_f_01 semantics on device buffers (in place on T)."""
    check(lib().rt_special_dgemm_dev(T.ptr, xa, ya, za, sx, lx, sz, lz, B.ptr, rb, cb, srb, lrb,
                                     scb, lcb, alpha, beta, stream))


def rifull_transpose(src: DeviceArray, n0, n1, n2, order: str, out: DeviceArray, stream=None):
    """RIFull::transpose_<order> of the column-major [n0, n1, n2] tensor."""
    check(lib().rt_rifull_transpose_dev(src.ptr, n0, n1, n2, order.encode(), out.ptr, stream))


def dsyev(jobz, n, A: DeviceArray, lda, w: DeviceArray, stream=None) -> int:
    """Eigenvalues (ascending) into w, eigenvectors into A for jobz 'V'.
    Returns LAPACK's info."""
    info = C.c_int(0)
    check(lib().rt_dsyev_dev(jobz.encode(), n, A.ptr, lda, w.ptr, C.byref(info), stream))
    return info.value


def dspgvx(n, Ap: DeviceArray, Bp: DeviceArray, num_orb, Z: DeviceArray, w: DeviceArray,
           stream=None) -> int:
    """A x = lambda B x on packed-upper A, B; returns LAPACK's info."""
    info = C.c_int(0)
    check(lib().rt_dspgvx_dev(n, Ap.ptr, Bp.ptr, num_orb, Z.ptr, w.ptr, C.byref(info), stream))
    return info.value


def dspevx(n, Ap: DeviceArray, Z: DeviceArray, w: DeviceArray, stream=None) -> int:
    """All eigenpairs of a packed-upper matrix; returns the number found."""
    nf = C.c_int(0)
    check(lib().rt_dspevx_dev(n, Ap.ptr, Z.ptr, w.ptr, C.byref(nf), stream))
    return nf.value


def fock_rhf_packed(n, H: DeviceArray, J: DeviceArray, K: DeviceArray, F: DeviceArray,
                    stream=None):
    check(lib().rt_fock_rhf_packed_dev(n, H.ptr, J.ptr, K.ptr, F.ptr, stream))


def dsteqr_host(compz, d, e, z=None):
    """Host DSTEQR (the tridiagonal stage of dsyev) on numpy arrays, in place:
    d (n,), e (n-1,), z (n, n) Fortran order for compz 'I'/'V'.  Returns info."""
    n = d.shape[0]
    info = C.c_int(0)
    zp = z.ctypes.data if z is not None else None
    ldz = z.shape[0] if z is not None else max(n, 1)
    check(lib().rt_dsteqr_host(compz.encode(), n, d.ctypes.data, e.ctypes.data if n > 1 else None,
                               zp, ldz, C.byref(info)))
    return info.value
