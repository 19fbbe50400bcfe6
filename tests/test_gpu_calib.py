"""GPU adjoint of the calibration model (gs_calib_*).  The oracle runs the
same model in fp32 (its final state must equal the GPU's bit for bit), which
records every branch decision and KNN list; the FP64 model replays them for
the loss and for central finite differences of the 14 parameters, so both
sides differentiate the branch the GPU took."""
import numpy as np
import pytest

from gpusteer_amd import Calibration, Ensemble, scenarios

pytestmark = pytest.mark.gpu


def _specs(n_scen, n_total, workload):
    specs = []
    for i in range(n_scen):
        sc = scenarios.choke(n_total=n_total, workload=workload + i)
        sc.mode = 1
        specs.append(("choke", sc))
    return specs


# h = 1e-3 as specified where the branch is well conditioned; over longer
# horizons the crowd aligns, so alignment's normalize(mean forward - forward)
# and cohesion's normalize(centroid - position) act on vectors near zero and a
# central difference of +-1e-3 can straddle that point (measured: at 12 steps
# the FP64 differences give 2.15 for d/dw_ali[1] at h = 1e-3 but 2.8778 at
# h = 1e-4 and at h = 1e-6, where the adjoint gives 2.8772), so the longer
# cases use h = 1e-4 (FP64 round-off at that step is ~1e-9 of the loss).
@pytest.mark.parametrize("steps,every,h", [(8, 3, 1e-3), (12, 12, 1e-4), (20, 7, 1e-4)])
def test_adjoint_matches_finite_differences(ctx, orc, steps, every, h):
    cfg = scenarios.CalibConfig(steps=steps, checkpoint_every=every)
    specs = _specs(2, 200, 7)
    th = scenarios.calib_params()
    ens = Ensemble(ctx, specs)
    cal = Calibration(ens, cfg)
    loss = cal.forward(th)
    grad = cal.backward()
    # Oracle: the fp32 model (the GPU's arithmetic) records the branches and
    # KNN lists; its final state must equal the GPU's bit for bit.  The FP64
    # model replays them for the loss and the central differences.
    base, fd = 0.0, np.zeros(14)
    for i, (_, sc) in enumerate(specs):
        c = orc.C4(sc, cfg)
        _, fs32 = c.run(th, 1, want_state=True, precision=32)
        n0, n1 = sc.groups[0].n, sc.groups[1].n
        for g, (off, n) in enumerate(((0, n0), (13 * n0, n1))):
            gpu = cal.final_state(i, g)
            for j, f in enumerate(orc.C4_STATE):
                np.testing.assert_array_equal(gpu.arrays[f], fs32[off + j * n: off + (j + 1) * n].astype(np.float32),
                                              err_msg=f"scenario {i} group {g} field {f}")
        b, g = c.fd_grad(th, h, record_precision=32)
        c.close()
        base += b
        fd += g
    base /= len(specs)
    fd /= len(specs)
    print(f"loss gpu {loss:.6f} oracle {base:.6f}")
    print("grad gpu", np.array2string(grad, precision=5))
    print("grad fd ", np.array2string(fd, precision=5))
    assert abs(loss - base) <= 1e-5 * abs(base)
    scale = np.abs(fd)
    err = np.abs(grad - fd)
    bad = err > 1e-3 * scale + 1e-6 * np.max(scale)
    assert not bad.any(), f"components {np.nonzero(bad)[0]}: gpu {grad[bad]} fd {fd[bad]}"
    cal.close()
    ens.close()


def test_forward_is_repeatable_and_backward_needs_forward(ctx):
    from gpusteer_amd._capi import GpuSteerError

    cfg = scenarios.CalibConfig(steps=5, checkpoint_every=2)
    ens = Ensemble(ctx, _specs(1, 128, 40))
    cal = Calibration(ens, cfg)
    with pytest.raises(GpuSteerError):
        cal.backward()
    th = scenarios.calib_params()
    a = cal.forward(th)
    b = cal.forward(th)
    assert a == b
    cal.close()
    ens.close()


def test_loss_and_gradient_are_bitwise_reproducible(ctx):
    """Every sum of the forward and backward is taken in a fixed order: the
    loss and the 14 gradient components of 4 crowds of 4,000 agents over 30
    steps are identical bit for bit across repeated backward passes, repeated
    forward passes and a second Calibration of the same ensemble."""
    cfg = scenarios.CalibConfig(steps=30, checkpoint_every=10)
    specs = _specs(4, 4000, 80)
    th = scenarios.calib_params()
    runs = []
    for _ in range(2):
        ens = Ensemble(ctx, specs)
        cal = Calibration(ens, cfg)
        for _ in range(2):
            loss = cal.forward(th)
            grads = [cal.backward() for _ in range(3)]
            for g in grads[1:]:
                np.testing.assert_array_equal(g, grads[0])
            runs.append((loss, grads[0]))
        cal.close()
        ens.close()
    print("loss", runs[0][0], "grad", np.array2string(runs[0][1], precision=4), flush=True)
    assert np.all(np.isfinite(runs[0][1])) and np.any(runs[0][1] != 0)
    for loss, g in runs[1:]:
        assert loss == runs[0][0]
        np.testing.assert_array_equal(g, runs[0][1])


def test_autograd_wrapper_matches_backward(ctx):
    """torch.autograd over the forward / backward pair returns the library's
    gradient, which points uphill; the wrapper drives a torch optimizer."""
    import torch

    from gpusteer_amd.autograd import calibration_loss

    cfg = scenarios.CalibConfig(steps=6, checkpoint_every=3)
    ens = Ensemble(ctx, _specs(2, 160, 60))
    cal = Calibration(ens, cfg)
    th0 = scenarios.calib_params()
    theta = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    loss = calibration_loss(cal, theta)
    loss.backward()
    cal.forward(th0)
    ref = cal.backward()
    np.testing.assert_array_equal(theta.grad.numpy(), ref)
    assert float(loss.detach()) == cal.forward(th0)
    # A small step against the gradient lowers the loss (first order).
    g = ref / np.linalg.norm(ref)
    assert cal.forward(th0 - 1e-5 * g) < cal.forward(th0)
    # The wrapper drives a torch optimizer.
    opt = torch.optim.Adam([theta], lr=0.01)
    for _ in range(3):
        opt.zero_grad()
        calibration_loss(cal, theta).backward()
        opt.step()
    assert np.isfinite(float(calibration_loss(cal, theta).detach()))
    cal.close()
    ens.close()


# ---------------------------------------------------------------------------
# The calibration workload itself: 4 of the 64 crowds of 16,384 agents (gap
# 15, 50 obstacles), checkpoints every 10 steps.
#
# The crowd is chaotic at this density: the oracle's tangents (RMS position
# derivative) grow about 10x every 4-5 steps from step 20 on (1e0 at step 20,
# 1e4 at 40, 1e7 at 60, 1e13 at 100), so central differences of +-1e-3
# resolve the derivative only over the first ~10 steps (measured: the FP64
# differences equal the FP64 tangent within 7e-5 at 10 steps, 0.2 off at 20,
# 2e3 off at 40), and the FP64 trajectory itself separates from the fp32 one
# the GPU follows.  So the gradient is checked two ways: against FP64
# central differences at h = 1e-3 over 10 steps, and over the full 100 steps
# against the oracle's forward-mode tangent evaluated along the GPU's own
# fp32 trajectory (the final fp32 state is bit-exact with the oracle's).

def _c4_specs():
    return scenarios.calib_scenarios(n_scenarios=4, n_agents=16384)


def _rel_ok(got, want, rel):
    scale = np.abs(want)
    return np.abs(got - want) <= rel * scale + 1e-6 * np.max(scale)


@pytest.mark.timeout(600)
def test_calibration_workload_10_steps_vs_fp64_differences(ctx, orc):
    specs = _c4_specs()
    cfg = scenarios.CalibConfig(steps=10, checkpoint_every=10)
    th = scenarios.calib_params()
    ens = Ensemble(ctx, specs)
    cal = Calibration(ens, cfg)
    loss = cal.forward(th)
    grad = cal.backward()
    cal.close()
    ens.close()
    base, fd, t64 = 0.0, np.zeros(14), np.zeros(14)
    for i, (_, sc) in enumerate(specs):
        print(f"scenario {i}: oracle fp32 record, FP64 differences h = 1e-3, FP64 tangent", flush=True)
        c = orc.C4(sc, cfg)
        b, g = c.fd_grad(th, 1e-3, record_precision=32)
        _, t = c.tangent_grad(th, 64)
        c.close()
        base += b / len(specs)
        fd += g / len(specs)
        t64 += t / len(specs)
    print("grad gpu   ", np.array2string(grad, precision=6))
    print("fd (h=1e-3)", np.array2string(fd, precision=6))
    print("tangent f64", np.array2string(t64, precision=6))
    assert abs(loss - base) <= 1e-5 * abs(base)
    # The tangent code is the FP64 differences' derivative (measured 3e-5).
    assert _rel_ok(t64, fd, 1e-4).all()
    ok = _rel_ok(grad, fd, 1e-3)
    assert ok.all(), f"components {np.nonzero(~ok)[0]}: gpu {grad[~ok]} fd {fd[~ok]}"


@pytest.mark.timeout(900)
def test_calibration_workload_100_steps_vs_tangent(ctx, orc):
    specs = _c4_specs()
    cfg = scenarios.CalibConfig(steps=100, checkpoint_every=10)
    th = scenarios.calib_params()
    ens = Ensemble(ctx, specs)
    cal = Calibration(ens, cfg)
    loss = cal.forward(th)
    grad = cal.backward()
    # A second backward of the same forward: the same gradient up to the
    # order of the float atomics that gather neighbour adjoints (not bitwise
    # reproducible; measured spread below 1e-5 relative).
    grad2 = cal.backward()
    finals = [[cal.final_state(i, g) for g in range(2)] for i in range(len(specs))]
    cal.close()
    ens.close()
    base, t32 = 0.0, np.zeros(14)
    for i, (_, sc) in enumerate(specs):
        print(f"scenario {i}: oracle fp32 record (100 steps), tangent along the fp32 trajectory", flush=True)
        c = orc.C4(sc, cfg)
        _, fs32 = c.run(th, 1, want_state=True, precision=32)
        n0, n1 = sc.groups[0].n, sc.groups[1].n
        for g, (off, n) in enumerate(((0, n0), (13 * n0, n1))):
            for j, f in enumerate(orc.C4_STATE):
                np.testing.assert_array_equal(finals[i][g].arrays[f],
                                              fs32[off + j * n: off + (j + 1) * n].astype(np.float32),
                                              err_msg=f"scenario {i} group {g} field {f}")
        b, t = c.tangent_grad(th, 32)
        c.close()
        base += b / len(specs)
        t32 += t / len(specs)
    print("grad gpu    ", np.array2string(grad, precision=6))
    print("grad gpu (2)", np.array2string(grad2, precision=6))
    print("tangent f32 ", np.array2string(t32, precision=6))
    assert abs(loss - base) <= 1e-5 * abs(base)
    ok = _rel_ok(grad, t32, 1e-3)
    assert ok.all(), f"components {np.nonzero(~ok)[0]}: gpu {grad[~ok]} tangent {t32[~ok]}"
    assert _rel_ok(grad2, grad, 1e-4).all()
