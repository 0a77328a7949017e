"""GPU adjoint of the step: against the FP64 CPU adjoint on a 4,096-particle
scene, against FP32 central differences (eps = 1e-3), checkpoint-interval
invariance, and the torch.autograd.Function path."""
import numpy as np
import pytest

import wcsph
from oracle import oracle as O
from tests.helpers import baseline_params
from wcsph import scenes

pytestmark = pytest.mark.gpu

STEPS, DT, ALPHA = 20, 2e-4, 0.01


def _scene(seed=5):
    desc = scenes.DamBreak(dims=(16, 16, 16), box_max=(0.64, 0.5, 0.5), seed=seed)
    fl, bd, box = scenes.dam_break(desc)
    lo = fl[:, :3].min(axis=0)
    pos = fl.copy()
    pos[:, :3] = (lo + (fl[:, :3] - lo) * np.float32(0.86)).astype(np.float32)
    rng = np.random.default_rng(seed)
    v0 = np.zeros_like(pos)
    v0[:, :3] = rng.normal(0, 0.1, (len(pos), 3)).astype(np.float32)
    return pos, bd, box, v0, rng


def _targets(prm, pos, bd, box, v0, rng):
    with wcsph.FluidSimulator(prm, pos, bd, box) as sim:
        sim.set_viscosity(0.02)
        xt = sim.grad_forward(v0 + rng.normal(0, 0.05, v0.shape).astype(np.float32) * [1, 1, 1, 0], DT, STEPS, STEPS)
    return xt.astype(np.float64)


def _gpu_grads(prm, pos, bd, box, v0, targets, every=STEPS, alpha=ALPHA):
    with wcsph.FluidSimulator(prm, pos, bd, box) as sim:
        sim.set_viscosity(alpha)
        xt = sim.grad_forward(v0, DT, STEPS, every)
        n = len(pos)
        g = np.zeros((n, 4), np.float32)
        g[:, :3] = (2.0 * (xt[:, :3] - targets[:, :3]) / n).astype(np.float32)
        gv, ga = sim.grad_backward(g)
        loss = np.sum((xt[:, :3].astype(np.float64) - targets[:, :3]) ** 2) / n
    return loss, gv, ga


def test_gpu_adjoint_matches_fp64_cpu_adjoint():
    prm = baseline_params()
    pos, bd, box, v0, rng = _scene()
    targets = _targets(prm, pos, bd, box, v0, rng)
    loss, gv, ga = _gpu_grads(prm, pos, bd, box, v0, targets)
    ref = O.Solver(prm, pos, bd, box, precision=64)
    ref.set_state64(pos.astype(np.float64), v0.astype(np.float64))
    ref.set_viscosity(ALPHA)
    rl = ref.grad_run(DT, STEPS, targets)
    rgv, rga = ref.grad_backward()
    assert loss == pytest.approx(rl, rel=1e-3)
    err = np.linalg.norm(gv[:, :3] - rgv[:, :3]) / np.linalg.norm(rgv[:, :3])
    assert err < 2e-3, err
    assert ga == pytest.approx(rga, rel=5e-3)
    assert rga != 0.0


def test_checkpoint_interval_does_not_change_gradients():
    prm = baseline_params()
    pos, bd, box, v0, rng = _scene(7)
    targets = _targets(prm, pos, bd, box, v0, rng)
    out = [_gpu_grads(prm, pos, bd, box, v0, targets, every=k) for k in (1, 7, STEPS)]
    for o in out[1:]:
        assert o[0] == out[0][0]
        np.testing.assert_array_equal(o[1], out[0][1])
        assert o[2] == out[0][2]


def test_gpu_adjoint_vs_central_differences():
    """FP32 central differences (eps = 1e-3) along random directions of the
    initial velocity field.  (d L / d alpha is checked against the FP64
    adjoint above: at alpha = 0.01 its effect on the loss over +-eps is far
    below the FP32 rounding of the loss, so an FP32 difference cannot see it.)"""
    prm = baseline_params()
    pos, bd, box, v0, rng = _scene(9)
    targets = _targets(prm, pos, bd, box, v0, rng)
    _, gv, _ = _gpu_grads(prm, pos, bd, box, v0, targets)
    eps = 1e-3
    for _ in range(2):
        d = rng.normal(0, 1, v0.shape).astype(np.float32)
        d[:, 3] = 0
        lp, _, _ = _gpu_grads(prm, pos, bd, box, v0 + eps * d, targets)
        lm, _, _ = _gpu_grads(prm, pos, bd, box, v0 - eps * d, targets)
        fd = (lp - lm) / (2 * eps)
        an = float(np.sum(gv.astype(np.float64) * d))
        assert an == pytest.approx(fd, rel=2e-2), (an, fd)


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    from wcsph.torch_adjoint import rollout

    prm = baseline_params()
    pos, bd, box, v0, rng = _scene(11)
    targets = _targets(prm, pos, bd, box, v0, rng)
    dev = torch.device("cuda:0")
    with wcsph.FluidSimulator(prm, pos, bd, box) as sim:
        v = torch.tensor(v0, device=dev, requires_grad=True)
        alpha = torch.tensor(ALPHA, dtype=torch.float64, device=dev, requires_grad=True)
        xt = rollout(v, alpha, sim, DT, STEPS, 5)
        tgt = torch.tensor(targets, dtype=torch.float32, device=dev)
        loss = ((xt[:, :3] - tgt[:, :3]) ** 2).sum(dim=1).mean()
        loss.backward()
        gv_t = v.grad.cpu().numpy()
        ga_t = alpha.grad.item()
    loss_np, gv, ga = _gpu_grads(prm, pos, bd, box, v0, targets, every=5)
    np.testing.assert_allclose(gv_t[:, :3], gv[:, :3], rtol=1e-4, atol=1e-9)
    assert ga_t == pytest.approx(ga, rel=1e-4)


def test_backward_repeats_and_keeps_its_forward_run():
    """The backward pass seeds from the forward run's final order and uses
    its alpha: called twice (checkpoints every 7 of 20 steps, so the handle
    ends the first backward in an intermediate step's order), and again
    after set_viscosity and plain steps on the handle, it returns the
    gradients of a fresh handle's single backward, bit for bit."""
    prm = baseline_params()
    pos, bd, box, v0, rng = _scene(13)
    targets = _targets(prm, pos, bd, box, v0, rng)
    loss0, gv0, ga0 = _gpu_grads(prm, pos, bd, box, v0, targets, every=7)
    with wcsph.FluidSimulator(prm, pos, bd, box) as sim:
        assert sim.grad_generation() == 0
        sim.set_viscosity(ALPHA)
        xt = sim.grad_forward(v0, DT, STEPS, 7)
        gen = sim.grad_generation()
        assert gen > 0
        n = len(pos)
        g = np.zeros((n, 4), np.float32)
        g[:, :3] = (2.0 * (xt[:, :3] - targets[:, :3]) / n).astype(np.float32)
        runs = [sim.grad_backward(g), sim.grad_backward(g)]
        sim.set_viscosity(0.5)
        sim.reset()
        sim.step(DT, 3)
        runs.append(sim.grad_backward(g))
        assert sim.grad_generation() == gen
    for gv, ga in runs:
        np.testing.assert_array_equal(gv, gv0)
        assert ga == ga0


def test_torch_rollout_refuses_a_superseded_forward():
    """Two rollouts on one simulator: the first one's backward raises (the
    simulator holds the second run's checkpoints), the second one's works."""
    torch = pytest.importorskip("torch")
    from wcsph.torch_adjoint import rollout

    prm = baseline_params()
    pos, bd, box, v0, rng = _scene(17)
    dev = torch.device("cuda:0")
    with wcsph.FluidSimulator(prm, pos, bd, box) as sim:
        v = torch.tensor(v0, device=dev, requires_grad=True)
        alpha = torch.tensor(ALPHA, dtype=torch.float64, device=dev, requires_grad=True)
        first = rollout(v, alpha, sim, DT, 6, 3)
        second = rollout(v, alpha, sim, DT, 8, 4)
        with pytest.raises(RuntimeError, match="another forward rollout"):
            first[:, :3].sum().backward()
        second[:, :3].sum().backward()
        assert v.grad is not None and torch.isfinite(v.grad).all()
