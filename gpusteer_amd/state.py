"""Host-side agent / wall state as numpy struct-of-arrays.

The layout mirrors gs_agents / gs_walls of include/gpusteer.h: one float32
array per field (AgentData, include/OpenSteer/AgentData.h:13-42, plus the
smoothed acceleration of SimpleVehicle) and a uint32 bits array.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict

import numpy as np

AGENT_FIELDS = (
    "px", "py", "pz", "fx", "fy", "fz", "speed", "ax", "ay", "az",
    "sx", "sy", "sz", "ux", "uy", "uz", "gx", "gy", "gz",
    "radius", "max_speed", "max_force", "mass",
)
WALL_FIELDS = ("sx", "sy", "sz", "ex", "ey", "ez", "nx", "ny", "nz")


@dataclass
class AgentState:
    """Struct-of-arrays agent state (float32 unless dtype says otherwise)."""

    arrays: Dict[str, np.ndarray] = field(default_factory=dict)
    bits: np.ndarray = None

    @property
    def n(self) -> int:
        return 0 if self.bits is None else int(self.bits.shape[0])

    def __getitem__(self, k: str) -> np.ndarray:
        return self.bits if k == "bits" else self.arrays[k]

    def copy(self, dtype=None) -> "AgentState":
        dt = dtype or self.arrays["px"].dtype
        return AgentState({k: np.array(v, dtype=dt, copy=True) for k, v in self.arrays.items()},
                          np.array(self.bits, dtype=np.uint32, copy=True))

    def pos(self) -> np.ndarray:
        return np.stack([self.arrays["px"], self.arrays["py"], self.arrays["pz"]], axis=1)

    def fwd(self) -> np.ndarray:
        return np.stack([self.arrays["fx"], self.arrays["fy"], self.arrays["fz"]], axis=1)

    @staticmethod
    def empty(n: int, dtype=np.float32) -> "AgentState":
        return AgentState({k: np.zeros(n, dtype=dtype) for k in AGENT_FIELDS}, np.zeros(n, dtype=np.uint32))

    def subset(self, idx: np.ndarray) -> "AgentState":
        return AgentState({k: np.ascontiguousarray(v[idx]) for k, v in self.arrays.items()},
                          np.ascontiguousarray(self.bits[idx]))


@dataclass
class WallState:
    arrays: Dict[str, np.ndarray]

    @property
    def n(self) -> int:
        return int(self.arrays["sx"].shape[0])

    def mid(self) -> np.ndarray:
        a = self.arrays
        f = np.float32(0.5)
        return np.stack([f * (a["sx"] + a["ex"]), f * (a["sy"] + a["ey"]), f * (a["sz"] + a["ez"])], axis=1)


def basis_from_forward(fx, fy, fz):
    """Side and up vectors for unit forward vectors, with world up (0,1,0):
    side = normalize(forward x up0), up = side x forward (right-handed, as
    SimpleVehicle's regenerateOrthonormalBasisUF, SimpleVehicle.h:190-217)."""
    f = np.stack([fx, fy, fz], axis=1).astype(np.float64)
    up0 = np.array([0.0, 1.0, 0.0])
    s = np.cross(f, up0)
    ln = np.linalg.norm(s, axis=1, keepdims=True)
    bad = ln[:, 0] < 1e-6
    s[bad] = np.cross(f[bad], np.array([1.0, 0.0, 0.0]))
    ln[bad] = np.linalg.norm(s[bad], axis=1, keepdims=True)
    s = s / ln
    u = np.cross(s, f)
    return s.astype(np.float32), u.astype(np.float32)
