"""Batches of Hydro handles advanced together by grouped launches (m70h_batch_*)."""
from __future__ import annotations

import ctypes

from . import _lib


class Batch:
    """Advances many independent Hydro objects (uneven grids) with one launch per stage of
    the step; the per-step launch sequence is replayed from a hipGraph when `graph`."""

    def __init__(self, hydros, graph: bool = True):
        self._L = _lib.load()
        self.hydros = list(hydros)
        arr = (ctypes.c_void_p * len(self.hydros))(*[h.handle for h in self.hydros])
        b = ctypes.c_void_p()
        rc = self._L.m70h_batch_create(len(self.hydros), arr, ctypes.byref(b))
        if rc != _lib.M70H_OK:
            raise _lib.M70HError(rc, self._L.m70h_batch_last_error(None).decode())
        self._b = b
        self.set_graph(graph)

    def _chk(self, rc):
        if rc != _lib.M70H_OK:
            raise _lib.M70HError(rc, self._L.m70h_batch_last_error(self._b).decode())

    def set_rows_block(self, threads: int):
        self._chk(self._L.m70h_batch_set_rows_block(self._b, int(threads)))

    def set_graph(self, on: bool):
        self._chk(self._L.m70h_batch_set_graph(self._b, 1 if on else 0))

    @property
    def stream(self) -> int:
        p = ctypes.c_void_p()
        self._chk(self._L.m70h_batch_stream(self._b, ctypes.byref(p)))
        return p.value or 0

    def run(self, n: int):
        """Run(n) of every handle (forcing must have been set on each)."""
        self._chk(self._L.m70h_batch_run(self._b, int(n)))

    def run_async(self, n: int):
        self._chk(self._L.m70h_batch_run_async(self._b, int(n)))

    def synchronize(self):
        self._chk(self._L.m70h_batch_synchronize(self._b))

    def close(self):
        if getattr(self, "_b", None):
            self._L.m70h_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
