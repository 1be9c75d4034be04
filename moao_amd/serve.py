"""Python handle on the serving engine (moao_serve_* in include/moao_hip.h).

    with Server(slots=4) as srv:
        t = srv.submit(sys_file, atm_file, target=0, mode=MOAO_FP64, want_R=True)
        res = srv.wait(t)          # dict: status, n, m, slot, footprint, times, R (m x n)

submit() returns at once (the request is queued and admitted by HBM footprint as device slots
free up); wait() blocks until that request is done.  ctypes releases the GIL during the C
calls, so other Python threads may keep submitting while one waits.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from .lib import (MOAO_FP64, ServeConfig, ServeCounters, ServeRequest, ServeResult, check,
                  lib)


class Server:
    def __init__(self, slots: int = 4, hbm_budget: int = 0, max_bypass: int = 0):
        cfg = ServeConfig(slots, hbm_budget, max_bypass)
        h = ctypes.c_void_p()
        check(lib().moao_serve_create(ctypes.byref(cfg), ctypes.byref(h)), "moao_serve_create")
        self._h = h
        self._mu = threading.Lock()
        self._out = {}   # ticket -> (R array or None, request keep-alive)

    def submit(self, sys_file: str, atm_file: str, target: int = 0, mode: int = MOAO_FP64,
               tol: float = 1e-11, max_iter: int = 5, want_R: bool = True,
               n_m: tuple | None = None) -> int:
        """Queue one refine-step request; returns its ticket.  R (m x n, column-major) is
        allocated here when want_R; n_m may give its shape, else it is read from the system
        (the engine refuses an R whose shape is not the system's)."""
        R = None
        if want_R:
            if n_m is None:
                from .tomo import Tomo
                t = Tomo(sys_file, atm_file)
                n_m = (t.nmeas(), t.nmeasts())
            n, m = n_m
            R = np.zeros((m, n), order="F")
        sf, af = sys_file.encode(), atm_file.encode()
        req = ServeRequest(sf, af, target, mode, tol, max_iter,
                           R.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if R is not None
                           else None, R.shape[0] if R is not None else 0,
                           R.shape[1] if R is not None else 0)
        ticket = ctypes.c_long()
        check(lib().moao_serve_submit(self._h, ctypes.byref(req), ctypes.byref(ticket)),
              "moao_serve_submit")
        with self._mu:
            self._out[ticket.value] = (R, (sf, af, req))
        return ticket.value

    def wait(self, ticket: int, raise_on_error: bool = True) -> dict:
        res = ServeResult()
        rc = lib().moao_serve_wait(self._h, ticket, ctypes.byref(res))
        with self._mu:
            R, _ = self._out.pop(ticket, (None, None))
        if raise_on_error:
            check(rc, f"request {ticket}")
        out = {k: getattr(res, k) for k, _ in ServeResult._fields_}
        out["status"] = rc
        out["latency_s"] = res.t_end - res.t_submit
        out["queue_s"] = res.t_start - res.t_submit
        out["R"] = R if rc == 0 else None
        return out

    def stats(self) -> dict:
        c = ServeCounters()
        check(lib().moao_serve_stats(self._h, ctypes.byref(c)), "moao_serve_stats")
        return {k: getattr(c, k) for k, _ in ServeCounters._fields_}

    def close(self):
        if self._h:
            lib().moao_serve_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
