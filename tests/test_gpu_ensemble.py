"""GPU parity of the batched ensemble: every scenario of a mixed ensemble
(Boids flocks and choke-point crowds of uneven size, both force modes) must
equal, bit for bit, the same scenario stepped alone through the
single-instance calls -- which the other tests pin to the oracle."""
import numpy as np
import pytest

from gpusteer_amd import Ensemble, Sequential, scenarios
from helpers import assert_state_equal

pytestmark = pytest.mark.gpu


def _compare(ens, seq, specs, what):
    for i, spec in enumerate(specs):
        ng = 1 if spec[0] == "boids" else 2
        for g in range(ng):
            tag = f"{what} scenario {i} ({spec[0]}) group {g}: "
            np.testing.assert_array_equal(ens.download_knn(i, g), seq.download_knn(i, g), err_msg=tag)
            assert_state_equal(ens.download(i, g), seq.download(i, g), what=tag)


@pytest.mark.parametrize("mode", [0, 1])
def test_ensemble_equals_sequential(ctx, mode):
    specs = scenarios.ensemble_mix(n_scenarios=6, n_min=300, n_max=3000, workload=11, mode=mode)
    print(f"ensemble mode {mode}: {[(s[0], s[1].n if s[0] == 'boids' else 2 * s[1].groups[0].n) for s in specs]}",
          flush=True)
    ens = Ensemble(ctx, specs)
    seq = Sequential(ctx, specs)
    # 1 eager first step + graph pairs + one eager remainder.
    ens.step(8)
    seq.step(8)
    _compare(ens, seq, specs, f"after 8 steps (mode {mode})")
    ens.step(3, eager=True)
    seq.step(3)
    _compare(ens, seq, specs, f"after 11 steps (mode {mode})")
    ens.close()
    seq.close()


def test_ensemble_graph_equals_eager(ctx):
    specs = scenarios.ensemble_mix(n_scenarios=4, n_min=500, n_max=2000, workload=12)
    a, b = Ensemble(ctx, specs), Ensemble(ctx, specs)
    a.step(10)
    b.step(10, eager=True)
    for i, spec in enumerate(specs):
        for g in range(1 if spec[0] == "boids" else 2):
            assert_state_equal(a.download(i, g), b.download(i, g), what=f"scenario {i} group {g}: ")
    a.close()
    b.close()


def test_ensemble_reupload_restarts(ctx):
    """Uploading a state again starts from it (the next step is unseeded)."""
    specs = scenarios.ensemble_mix(n_scenarios=2, n_min=800, n_max=800, workload=13)
    ens = Ensemble(ctx, specs)
    ens.step(5)
    for i, spec in enumerate(specs):
        groups = [spec[1]] if spec[0] == "boids" else list(spec[1].groups)
        for g, st in enumerate(groups):
            ens.upload(i, g, st)
    seq = Sequential(ctx, specs)
    ens.step(4)
    seq.step(4)
    _compare(ens, seq, specs, "after re-upload")
    ens.close()
    seq.close()


def test_ensemble_per_agent_constants_and_steering(ctx):
    """The counting sort leaves in place the fields every population holds
    one value for (radius, speed and force limits, mass) and, after the
    first step, the steering and bits (zero after every step).  Agents
    uploaded with varying radii and non-zero steering and bits must still
    equal the sequential calls; a re-upload that makes the radii uniform
    and then varied again changes the set the captured steps move."""
    specs = scenarios.ensemble_mix(n_scenarios=4, n_min=600, n_max=1500, workload=14)
    g = np.random.default_rng(5)

    def varied(st):
        st = st.copy()
        a = st.arrays
        a["radius"][:] = g.uniform(0.3, 0.7, st.n).astype(np.float32)
        for f in ("gx", "gy", "gz"):
            a[f][:] = g.uniform(-1, 1, st.n).astype(np.float32)
        st.bits[:] = g.integers(0, 1 << 12, st.n).astype(np.uint32)
        return st

    def states(spec, fn):
        return [fn(spec[1])] if spec[0] == "boids" else [fn(x) for x in spec[1].groups]

    ens = Ensemble(ctx, specs)
    seq = Sequential(ctx, specs)
    for round_, fn in enumerate((varied, lambda st: st.copy(), varied)):
        for i, spec in enumerate(specs):
            for gi, st in enumerate(states(spec, fn)):
                ens.upload(i, gi, st)
                seq.upload(i, gi, st)
        ens.step(7)
        seq.step(7)
        _compare(ens, seq, specs, f"upload round {round_}")
    ens.close()
    seq.close()


def test_ensemble_workload_subset_equals_oracle(ctx, orc):
    """The batched-ensemble workload itself: all 256 scenarios of the mix
    (agent counts log-uniform over 1,024-65,536, about 3.9 M agents; even
    scenarios Boids flocks with jittered weights, odd ones choke-point crowds
    with gaps 7.5 ... 32.5), 21 steps (one eager step, then ten replays of
    the captured two-step graph), lattices at the bench's cell factor.  A
    stratified subset is compared with the oracle directly, bit for bit
    (KNN lists and state): the smallest and largest flock, the smallest and
    largest crowd, and the first crowd of every gap width."""
    specs = scenarios.ensemble_mix(cell_factor=1.6)
    boids = [i for i, s in enumerate(specs) if s[0] == "boids"]
    crowds = [i for i, s in enumerate(specs) if s[0] == "choke"]
    size = {i: (s[1].n if s[0] == "boids" else 2 * s[1].groups[0].n) for i, s in enumerate(specs)}
    pick = {min(boids, key=size.get), max(boids, key=size.get), min(crowds, key=size.get),
            max(crowds, key=size.get)}
    gaps = {}
    for i in crowds:
        w = float(specs[i][1].walls.arrays["ex"][2])  # the left piece's end cap sits at x = -gap / 2
        gaps.setdefault(w, i)
    assert len(gaps) == 13
    pick |= set(gaps.values())
    print(f"ensemble workload: {len(specs)} scenarios, {sum(size.values())} agents; oracle subset "
          f"{sorted((i, specs[i][0], size[i]) for i in pick)}", flush=True)
    refs = {}
    for i in pick:
        s = specs[i]
        refs[i] = s[1].copy() if s[0] == "boids" else [g.copy() for g in s[1].groups]
    ens = Ensemble(ctx, specs)
    steps = 21
    ens.step(steps)
    for i in sorted(pick):
        s = specs[i]
        print(f"oracle: scenario {i} ({s[0]}, {size[i]} agents), {steps} steps", flush=True)
        if s[0] == "boids":
            for _ in range(steps):
                r_idx, _ = orc.boids_step(refs[i], s[2])
            tag = f"scenario {i} (boids, {size[i]}): "
            np.testing.assert_array_equal(ens.download_knn(i, 0), r_idx, err_msg=tag)
            assert_state_equal(ens.download(i, 0), refs[i], what=tag)
        else:
            for _ in range(steps):
                r = orc.choke_step(refs[i], s[1])
            for g in range(2):
                tag = f"scenario {i} (crowd, {size[i]}) group {g}: "
                np.testing.assert_array_equal(ens.download_knn(i, g), r[g][0], err_msg=tag)
                assert_state_equal(ens.download(i, g), refs[i][g], what=tag)
    ens.close()


def test_ensemble_crowds_meeting_equal_oracle(ctx, orc):
    """Crowds whose groups overlap: group 1 starts shifted 40 units toward
    group 0, so part of each group has the other within the cap and part lies
    beyond the other group's bounding box (those queries skip the search).
    Three crowds of uneven size and a flock, 8 steps (graph pairs), each
    crowd compared with the oracle bit for bit (own lists and state; the
    other-group lists enter the state through separation)."""
    specs = []
    for i, (n, gap) in enumerate(((1200, 10.0), (3000, 15.0), (5000, 25.0))):
        sc = scenarios.choke(n_total=n, gap=gap, workload=40 + i)
        sc.groups[1].arrays["pz"] += np.float32(40.0)
        c = scenarios.choke_cells_for(n, cell_factor=1.6)
        sc.grid_cells = (c, 1, c)
        specs.append(("choke", sc))
    st, p = scenarios.boids_sphere(n=1500)
    specs.insert(1, ("boids", st, p))
    refs = {i: [g.copy() for g in s[1].groups] for i, s in enumerate(specs) if s[0] == "choke"}
    ens = Ensemble(ctx, specs)
    steps = 8
    ens.step(steps)
    met = 0
    for i, r_groups in refs.items():
        sc = specs[i][1]
        for _ in range(steps):
            r = orc.choke_step(r_groups, sc)
        met += int(np.sum(r[0][1] != 0xFFFFFFFF))
        for g in range(2):
            tag = f"crowd {i} group {g}: "
            np.testing.assert_array_equal(ens.download_knn(i, g), r[g][0], err_msg=tag)
            assert_state_equal(ens.download(i, g), r_groups[g], what=tag)
    assert met > 0  # the groups do meet within the cap
    ens.close()
