"""Writes tests/golden/lq_golden.npz from the reference's own src/lq.c.

Run in a container that has the reference sources (the fixture is committed,
so the tests do not need them):

    make -C oracle ref && python tests/golden/gen_lq_golden.py

Contents (all inputs seeded and stored with the outputs):
  grid_origin, grid_size, grid_div   super-brick of the bin lattice
  points [N,3] float32               stored objects (some outside the lattice)
  bin [N] int32                      lqBinForLocation bin index (-1 = "other")
  queries [Q,3] float32, radii [Q]   locality queries
  nbr_offsets [Q+1], nbr_members     lqMapOverAllObjectsInLocality results
                                     (object indices, ascending)
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import lqref  # noqa: E402


def main() -> None:
    if not lqref.available():
        raise SystemExit("oracle/_ref/liblq.so missing: run `make -C oracle ref` where the reference exists")
    g = np.random.default_rng(20260412)
    origin = np.array([-55.0, -55.0, -55.0], np.float32)
    size = np.array([110.0, 110.0, 110.0], np.float32)
    div = np.array([10, 10, 10], np.int32)
    inside = g.uniform(-54.0, 54.0, size=(1500, 3))
    outside = g.uniform(-80.0, 80.0, size=(300, 3))
    # Points exactly on bin faces exercise the truncation rule.
    faces = np.stack([np.full(40, -55.0 + 11.0 * 3), g.uniform(-50, 50, 40), g.uniform(-50, 50, 40)], axis=1)
    pts = np.concatenate([inside, outside, faces]).astype(np.float32)
    db = lqref.LqDB(origin, size, div)
    bins = np.array([db.bin_index(*p) for p in pts], np.int32)
    db.insert(pts)
    queries = np.concatenate([g.uniform(-60, 60, size=(60, 3)), pts[:20]]).astype(np.float32)
    radii = g.uniform(2.0, 25.0, size=queries.shape[0]).astype(np.float32)
    offs = [0]
    members = []
    for q, r in zip(queries, radii):
        m = db.radius_query(q, r)
        members.extend(m)
        offs.append(len(members))
    db.close()
    np.savez_compressed(HERE / "lq_golden.npz", grid_origin=origin, grid_size=size, grid_div=div, points=pts,
                        bin=bins, queries=queries, radii=radii, nbr_offsets=np.array(offs, np.int64),
                        nbr_members=np.array(members, np.int64))
    print(f"wrote {HERE / 'lq_golden.npz'}: {pts.shape[0]} points, {queries.shape[0]} queries, "
          f"{len(members)} neighbours")


if __name__ == "__main__":
    main()
