"""Mirror of the reference's ``linalg`` loop-order vocabulary
(src/linalg.hpp:11-37).  The GPU kernels do not depend on the loop order;
it is validated and carried so callers written against the reference work
unchanged."""
from __future__ import annotations

import ctypes as C
import enum

from . import _lib


class Order(enum.IntEnum):
    """linalg::Order, same enumerators in the same order (src/linalg.hpp:11-19)."""
    ijk = 0
    ikj = 1
    jik = 2
    jki = 3
    kij = 4
    kji = 5


def string2order(order_str: str) -> Order:
    """linalg::string2order (src/linalg.hpp:21-37); raises
    RuntimeError("Invalid loop order") like the reference's throw (:36)."""
    out = C.c_int(-1)
    st = _lib.lib().femtc_string2order(order_str.encode(), C.byref(out))
    if st == _lib.ERR_INVALID_ORDER_STRING:
        raise RuntimeError("Invalid loop order")
    _lib.check(st, "string2order")
    return Order(out.value)
