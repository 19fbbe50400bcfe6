"""ctypes driver of the reference's own src/lq.c, compiled from the reference
sources into oracle/_ref/liblq.so by oracle/Makefile (`make -C oracle ref`).

Only used where the reference sources exist (this container): by
gen_lq_golden.py to write the committed fixtures, and by the CPU tests as an
extra live cross-check.  lq.h's public interface is driven as a client would:
lqCreateDatabase / lqInitClientProxy / lqUpdateForNewLocation /
lqMapOverAllObjectsInLocality / lqBinForLocation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parents[2] / "oracle" / "_ref" / "liblq.so"


class LqClientProxy(C.Structure):
    pass


# include/OpenSteer/lq.h:118-136
LqClientProxy._fields_ = [
    ("prev", C.POINTER(LqClientProxy)),
    ("next", C.POINTER(LqClientProxy)),
    ("bin", C.POINTER(C.POINTER(LqClientProxy))),
    ("object", C.c_void_p),
    ("x", C.c_float),
    ("y", C.c_float),
    ("z", C.c_float),
]

CALLBACK = C.CFUNCTYPE(None, C.c_void_p, C.c_float, C.c_void_p)


def available() -> bool:
    return LIB.exists()


class LqDB:
    def __init__(self, origin, size, div):
        self.L = C.CDLL(str(LIB))
        L = self.L
        L.lqCreateDatabase.restype = C.c_void_p
        L.lqCreateDatabase.argtypes = [C.c_float] * 6 + [C.c_int] * 3
        L.lqDeleteDatabase.argtypes = [C.c_void_p]
        L.lqInitClientProxy.argtypes = [C.POINTER(LqClientProxy), C.c_void_p]
        L.lqUpdateForNewLocation.argtypes = [C.c_void_p, C.POINTER(LqClientProxy), C.c_float, C.c_float, C.c_float]
        L.lqMapOverAllObjectsInLocality.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float,
                                                    CALLBACK, C.c_void_p]
        L.lqBinForLocation.restype = C.c_void_p
        L.lqBinForLocation.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float]
        self.origin, self.size, self.div = origin, size, div
        self.db = L.lqCreateDatabase(*[float(v) for v in origin], *[float(v) for v in size], *[int(d) for d in div])
        # Address of bin 0 and of the "other" bin, found through lqBinForLocation.
        self.bin0 = L.lqBinForLocation(self.db, *[float(o) for o in origin])
        self.other = L.lqBinForLocation(self.db, float(origin[0]) - 1.0, float(origin[1]), float(origin[2]))
        self.proxies = None

    def bin_index(self, x, y, z) -> int:
        """Linear bin index lq assigns to a point, or -1 for the "other" bin."""
        b = self.L.lqBinForLocation(self.db, float(x), float(y), float(z))
        if b == self.other:
            return -1
        return (b - self.bin0) // C.sizeof(C.c_void_p)

    def insert(self, pts: np.ndarray) -> None:
        n = pts.shape[0]
        self.proxies = (LqClientProxy * n)()
        for i in range(n):
            self.L.lqInitClientProxy(C.byref(self.proxies[i]), C.c_void_p(i + 1))
            self.L.lqUpdateForNewLocation(self.db, C.byref(self.proxies[i]), float(pts[i, 0]), float(pts[i, 1]),
                                          float(pts[i, 2]))

    def radius_query(self, q, r) -> list:
        found = []

        def cb(obj, d2, state):
            found.append(int(obj) - 1)

        f = CALLBACK(cb)
        self.L.lqMapOverAllObjectsInLocality(self.db, float(q[0]), float(q[1]), float(q[2]), float(r), f, None)
        return sorted(found)

    def close(self):
        if self.db:
            self.L.lqDeleteDatabase(self.db)
            self.db = None
