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
