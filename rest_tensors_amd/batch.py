"""Batched many-molecule pipeline through the C ABI (rt_batch_*)."""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from ._lib import check, lib


class MolDesc(C.Structure):
    _fields_ = [("nbas", C.c_int64), ("naux", C.c_int64), ("nocc", C.c_int64),
                ("ri3ao", C.c_void_p), ("C", C.c_void_p), ("D", C.c_void_p),
                ("J", C.c_void_p), ("K", C.c_void_p), ("B", C.c_void_p)]


class Batch:
    """mols: list of dicts with DeviceArrays ri3ao, C, D, J, K, B and ints
    nbas, naux, nocc."""

    def __init__(self, mols):
        self.mols = mols
        arr = (MolDesc * len(mols))()
        for i, m in enumerate(mols):
            arr[i] = MolDesc(m["nbas"], m["naux"], m["nocc"], m["ri3ao"].ptr, m["C"].ptr,
                             m["D"].ptr, m["J"].ptr, m["K"].ptr, m["B"].ptr)
        h = C.c_void_p()
        check(lib().rt_batch_create(arr, len(mols), C.byref(h)))
        self.handle = h.value

    def autotune(self, reps=2, stream=None):
        check(lib().rt_batch_autotune(self.handle, reps, stream))

    STAGES = {"density": 0, "ao2mo_s1": 1, "ao2mo_s2": 2, "k_partials": 3}

    def set_tiles(self, stage: str, bm: int, bn: int, waves_m: int, grid_mul: int = 2):
        check(lib().rt_batch_set_stage(self.handle, self.STAGES[stage], bm, bn, waves_m,
                                       grid_mul))

    def reload(self, stream=None):
        """Refresh the padded ri3ao copies after the molecules' ri3ao changed."""
        check(lib().rt_batch_reload(self.handle, stream))

    def run(self, stream=None):
        check(lib().rt_batch_run(self.handle, stream))

    def capture(self, stream):
        check(lib().rt_batch_graph_capture(self.handle, stream))

    def launch(self, stream):
        check(lib().rt_batch_graph_launch(self.handle, stream))

    def energies(self, stream=None):
        n = len(self.mols)
        ej, ex = np.zeros(n), np.zeros(n)
        check(lib().rt_batch_energies(self.handle, ej.ctypes.data_as(C.POINTER(C.c_double)),
                                      ex.ctypes.data_as(C.POINTER(C.c_double)), stream))
        return ej, ex

    def info(self) -> dict:
        buf = C.create_string_buffer(4096)
        check(lib().rt_batch_info(self.handle, buf, 4096))
        return json.loads(buf.value.decode())

    def __del__(self):
        try:
            if self.handle:
                lib().rt_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
