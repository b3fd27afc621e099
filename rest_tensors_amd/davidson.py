"""Davidson solver, RI linear-response operator and learned preconditioner
through the C ABI (rt_davidson_*, rt_casida_*, rt_mlp_*)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib
from .device import DeviceArray

PRECOND = {"diag_ref": 0, "diag_root": 1, "mlp": 2, "aux": 3}


class Record(C.Structure):
    _fields_ = [("r", C.c_void_p), ("x", C.c_void_p), ("e_k", C.POINTER(C.c_double)),
                ("e_0", C.POINTER(C.c_double)), ("root", C.POINTER(C.c_int)),
                ("iter", C.POINTER(C.c_int)), ("cap", C.c_int64), ("count", C.c_int64),
                ("xs", C.c_void_p), ("xs_cap", C.c_int64), ("xs_used", C.c_int64),
                ("xs_off", C.POINTER(C.c_int64)), ("xs_cnt", C.POINTER(C.c_int))]


class Params(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxcyc", C.c_int), ("maxspace", C.c_int),
                ("lindep", C.c_double), ("nroots", C.c_int), ("precond", C.c_int),
                ("mlp", C.c_void_p), ("record", C.POINTER(Record)), ("verbose", C.c_int),
                ("aux_op", C.c_void_p), ("aux_ctx", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("matvecs", C.c_int), ("restarts", C.c_int),
                ("nroots_found", C.c_int), ("mlp_fallbacks", C.c_int),
                ("seconds", C.c_double), ("max_residual", C.c_double)]


@dataclass
class DavidsonResult:
    conv: list
    e: np.ndarray
    x: DeviceArray | None
    stats: dict
    status: int


class Recorder:
    """Owns the device/host buffers of an rt_davidson_record; with
    subspace_doubles > 0 it also keeps the subspace of every recorded cycle."""

    def __init__(self, n: int, cap: int, subspace_doubles: int = 0):
        self.n, self.cap = n, cap
        self.r = DeviceArray(n * cap)
        self.x = DeviceArray(n * cap)
        self.e_k = np.zeros(cap)
        self.e_0 = np.zeros(cap)
        self.root = np.zeros(cap, dtype=np.int32)
        self.iter = np.zeros(cap, dtype=np.int32)
        self.xs = DeviceArray(subspace_doubles) if subspace_doubles > 0 else None
        self.xs_off = np.full(cap, -1, dtype=np.int64)
        self.xs_cnt = np.zeros(cap, dtype=np.int32)
        self.rec = Record(self.r.ptr, self.x.ptr,
                          self.e_k.ctypes.data_as(C.POINTER(C.c_double)),
                          self.e_0.ctypes.data_as(C.POINTER(C.c_double)),
                          self.root.ctypes.data_as(C.POINTER(C.c_int)),
                          self.iter.ctypes.data_as(C.POINTER(C.c_int)), cap, 0,
                          self.xs.ptr if self.xs is not None else None,
                          subspace_doubles, 0,
                          self.xs_off.ctypes.data_as(C.POINTER(C.c_int64)),
                          self.xs_cnt.ctypes.data_as(C.POINTER(C.c_int)))

    @property
    def count(self) -> int:
        return int(self.rec.count)


def solve(matvec_fn, ctx, x0: DeviceArray, hdiag: DeviceArray, n: int, nroots: int,
          tol=1e-5, maxcyc=50, maxspace=12, lindep=1e-14, precond="diag_ref", mlp=None,
          recorder: Recorder | None = None, want_x=True, verbose=0, stream=None,
          allow_notconv=False, aux=None) -> DavidsonResult:
    """aux: an object with .apply_fn and .handle (e.g. ShiftInvert) used as the
    auxiliary correction operator."""
    p = Params(tol, maxcyc, maxspace, lindep, nroots, PRECOND[precond],
               mlp.handle if mlp is not None else None,
               C.pointer(recorder.rec) if recorder is not None else None, verbose,
               aux.apply_fn if aux is not None else None,
               aux.handle if aux is not None else None)
    conv = (C.c_int * nroots)()
    e = (C.c_double * nroots)()
    xo = DeviceArray(n * nroots) if want_x else None
    st = Stats()
    status = lib().rt_davidson_solve_dev(matvec_fn, ctx, x0.ptr, hdiag.ptr, n,
                                         C.byref(p), conv, e,
                                         xo.ptr if xo is not None else None,
                                         C.byref(st), stream)
    if status != 0 and not (allow_notconv and status == 4):
        check(status)
    stats = {f: getattr(st, f) for f, _ in Stats._fields_}
    return DavidsonResult([bool(c) for c in conv], np.array(e[:]), xo, stats, status)


class CasidaOp:
    """A = diag(d) + 2 scale^2 B B^T on device (B: n x naux, leading dim ldb)."""

    def __init__(self, B: DeviceArray, ldb: int, n: int, naux: int, scale: float,
                 d: DeviceArray, offset: int = 0):
        self.keep = (B, d)
        h = C.c_void_p()
        check(lib().rt_casida_op_create(B.ptr + 8 * offset, ldb, n, naux, scale, d.ptr,
                                        C.byref(h)))
        self.handle = h.value
        self.n = n
        self.matvec_fn = C.cast(lib().rt_casida_matvec, C.c_void_p)

    def apply(self, X: DeviceArray, Y: DeviceArray, k: int, stream=None):
        check(lib().rt_casida_matvec(self.handle, X.ptr, Y.ptr, self.n, k, stream))

    def hdiag(self, out: DeviceArray, stream=None):
        check(lib().rt_casida_hdiag(self.handle, out.ptr, stream))

    def __del__(self):
        try:
            if self.handle:
                lib().rt_casida_op_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class ShiftInvert:
    """T = (A - sigma)^-1 R for a CasidaOp, sigma = min(d) - shift (Woodbury)."""

    def __init__(self, op: CasidaOp, shift: float = 0.05, stream=None):
        self.keep = op
        h = C.c_void_p()
        check(lib().rt_casida_si_create(op.handle, shift, C.byref(h), stream))
        self.handle = h.value
        self.n = op.n
        self.apply_fn = C.cast(lib().rt_casida_si_apply, C.c_void_p)

    def apply(self, R: DeviceArray, T: DeviceArray, k: int, stream=None):
        check(lib().rt_casida_si_apply(self.handle, R.ptr, T.ptr, self.n, k, stream))

    def info(self) -> dict:
        sig, err = C.c_double(), C.c_double()
        check(lib().rt_casida_si_info(self.handle, C.byref(sig), C.byref(err)))
        return {"sigma": sig.value, "residual": err.value}

    def __del__(self):
        try:
            if self.handle:
                lib().rt_casida_si_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _dptr(a):
    return a.ptr if a is not None else None


def _subspaces(rec):
    """(Q, q_off, q_cnt) ctypes arguments of a Recorder (or Nones)."""
    if rec is None or rec.xs is None:
        return None, None, None
    return (rec.xs.ptr, rec.xs_off.ctypes.data_as(C.POINTER(C.c_int64)),
            rec.xs_cnt.ctypes.data_as(C.POINTER(C.c_int)))


class Mlp:
    def __init__(self, seed: int = 1):
        h = C.c_void_p()
        check(lib().rt_mlp_create(seed, C.byref(h)))
        self.handle = h.value
        self.nparam = lib().rt_mlp_num_params()

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.nparam, dtype=np.float32)
        check(lib().rt_mlp_get_params(self.handle, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        check(lib().rt_mlp_set_params(self.handle, p.ctypes.data_as(C.POINTER(C.c_float))))

    def apply(self, R, X, hdiag, ek, e0, n, ns, T, S=None, stream=None):
        ek = np.ascontiguousarray(ek, dtype=np.float64)
        e0 = np.ascontiguousarray(e0, dtype=np.float64)
        check(lib().rt_mlp_apply_dev(self.handle, R.ptr, X.ptr, _dptr(S), hdiag.ptr,
                                     ek.ctypes.data_as(C.POINTER(C.c_double)),
                                     e0.ctypes.data_as(C.POINTER(C.c_double)), n, ns,
                                     T.ptr, stream))

    def loss_grad(self, R, X, hdiag, ek, e0, tau, wts, n, ns, S=None, recorder=None,
                  q=None, stream=None):
        """q: explicit (Q DeviceArray, q_off int64[], q_cnt int32[]) subspaces;
        recorder: take them from a Recorder."""
        ek = np.ascontiguousarray(ek, dtype=np.float64)
        e0 = np.ascontiguousarray(e0, dtype=np.float64)
        wts = np.ascontiguousarray(wts, dtype=np.float64)
        Q, qo, qc = self._q(recorder, q)
        loss = C.c_double(0)
        grad = np.zeros(self.nparam)
        check(lib().rt_mlp_loss_grad_dev(
            self.handle, R.ptr, X.ptr, _dptr(S), hdiag.ptr,
            ek.ctypes.data_as(C.POINTER(C.c_double)),
            e0.ctypes.data_as(C.POINTER(C.c_double)), tau.ptr,
            wts.ctypes.data_as(C.POINTER(C.c_double)), n, ns, Q, qo, qc, C.byref(loss),
            grad.ctypes.data_as(C.POINTER(C.c_double)), stream))
        return loss.value, grad

    def train(self, R, X, hdiag, ek, e0, tau, wts, n, ns, steps, batch, lr, seed=7,
              S=None, recorder=None, q=None, stream=None):
        ek = np.ascontiguousarray(ek, dtype=np.float64)
        e0 = np.ascontiguousarray(e0, dtype=np.float64)
        wts = np.ascontiguousarray(wts, dtype=np.float64)
        Q, qo, qc = self._q(recorder, q)
        hist = np.zeros(max(steps, 1))
        check(lib().rt_mlp_train_dev(
            self.handle, R.ptr, X.ptr, _dptr(S), hdiag.ptr,
            ek.ctypes.data_as(C.POINTER(C.c_double)),
            e0.ctypes.data_as(C.POINTER(C.c_double)), tau.ptr,
            wts.ctypes.data_as(C.POINTER(C.c_double)), n, ns, Q, qo, qc, steps, batch, lr,
            seed, hist.ctypes.data_as(C.POINTER(C.c_double)), stream))
        return hist[:steps]

    @staticmethod
    def _q(recorder, q):
        if q is not None:
            Qa, off, cnt = q
            off = np.ascontiguousarray(off, dtype=np.int64)
            cnt = np.ascontiguousarray(cnt, dtype=np.int32)
            Mlp._keep_q = (off, cnt)
            return (Qa.ptr, off.ctypes.data_as(C.POINTER(C.c_int64)),
                    cnt.ctypes.data_as(C.POINTER(C.c_int)))
        return _subspaces(recorder)

    def __del__(self):
        try:
            if self.handle:
                lib().rt_mlp_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def mlp_targets(X: DeviceArray, root: np.ndarray, ns: int, Xfinal: DeviceArray, n: int,
                tau: DeviceArray, min_overlap=0.3, recorder=None, stream=None) -> np.ndarray:
    """Targets from the converged Xfinal; with a recorder that kept subspaces,
    tau = x* minus its projection on each sample's subspace."""
    root = np.ascontiguousarray(root, dtype=np.int32)
    wts = np.zeros(ns)
    Q, qo, qc = _subspaces(recorder)
    check(lib().rt_mlp_targets_dev(X.ptr, root.ctypes.data_as(C.POINTER(C.c_int)), ns,
                                   Xfinal.ptr, n, min_overlap, Q, qo, qc, tau.ptr,
                                   wts.ctypes.data_as(C.POINTER(C.c_double)), stream))
    return wts
