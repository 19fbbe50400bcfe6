"""Test helpers: lattice-tie generators and state comparison."""
import numpy as np

STATE_FIELDS = ("px", "py", "pz", "fx", "fy", "fz", "speed", "ax", "ay", "az", "sx", "sy", "sz", "ux", "uy", "uz",
                "gx", "gy", "gz")


def assert_state_equal(a, b, fields=STATE_FIELDS, what=""):
    """Bit-exact (value) equality of agent states (+0 == -0)."""
    for f in fields:
        x, y = a.arrays[f], b.arrays[f]
        if not np.array_equal(x, y):
            bad = np.nonzero(x != y)[0]
            i = bad[0]
            raise AssertionError(f"{what}field {f}: {bad.size} agents differ, first {i}: {x[i]!r} vs {y[i]!r}")
    assert np.array_equal(a.bits, b.bits), f"{what}bits differ"


def lattice_points(n_side: int, spacing: float = 1.0, origin=(0.0, 0.0, 0.0), dims=3):
    """Integer lattice: every point has many exactly equidistant neighbours,
    which exercises the lower-index tie rule."""
    r = np.arange(n_side, dtype=np.float32) * np.float32(spacing)
    if dims == 2:
        x, z = np.meshgrid(r, r, indexing="ij")
        pts = np.stack([x.ravel(), np.zeros(x.size, np.float32), z.ravel()], axis=1)
    else:
        x, y, z = np.meshgrid(r, r, r, indexing="ij")
        pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    return (pts + np.asarray(origin, np.float32)).astype(np.float32)
