"""Python handle on the library's host-side system model (moao_tomo in include/moao_hip.h).

The model itself is C++ (moao_amd/csrc/tomo.cpp); this class only owns the handle and exposes
its vectors as numpy arrays.  Mirrors Tomo_struct<double> (common/tomo_struct/tomo_struct.hpp).
"""
from __future__ import annotations

import ctypes

import numpy as np

from .lib import TomoView, check, lib


class Tomo:
    def __init__(self, sys_file: str, atm_file: str):
        h = ctypes.c_void_p()
        check(lib().moao_tomo_create(sys_file.encode(), atm_file.encode(), ctypes.byref(h)),
              "moao_tomo_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            lib().moao_tomo_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def update_atm(self, atm_file: str):
        check(lib().moao_tomo_update_atm(self._h, atm_file.encode()), "moao_tomo_update_atm")

    def set_target(self, t: int):
        check(lib().moao_tomo_set_target(self._h, t), "moao_tomo_set_target")

    def set_target_coord(self, ax: float, ay: float, to_rad: bool = False):
        check(lib().moao_tomo_set_target_coord(self._h, ax, ay, int(to_rad)),
              "moao_tomo_set_target_coord")

    def nmeas(self) -> int:
        return int(lib().moao_tomo_nmeas(self._h))

    def nmeasts(self) -> int:
        return int(lib().moao_tomo_nmeasts(self._h))

    def dims(self) -> dict:
        out = (ctypes.c_long * 6)()
        check(lib().moao_tomo_dims(self._h, out), "moao_tomo_dims")
        return dict(zip(("nW", "nLgs", "nNgs", "nTarget", "nLayer", "nssp"), list(out)))

    def field(self, name: str) -> np.ndarray:
        n = check(lib().moao_tomo_field(self._h, name.encode(), None, 0), "moao_tomo_field")
        a = np.zeros(n)
        lib().moao_tomo_field(self._h, name.encode(),
                              a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
        return a

    def view(self) -> TomoView:
        """Host-pointer view into this model (valid until the next update/set call)."""
        v = TomoView()
        check(lib().moao_tomo_get_view(self._h, ctypes.byref(v)), "moao_tomo_get_view")
        return v
