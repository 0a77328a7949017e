"""The FP64 CPU adjoint of the step against central finite differences
(FP64, so the differences are accurate to ~1e-9 relative).  This pins the
reverse-mode formulas the GPU adjoint kernels restate.  The tolerance
(1e-5 relative) leaves room for a probe of +-eps that straddles one of the
step's kinks (the Tait clamp at rho = rho0, the viscosity switch at
v_ij . x_ij = 0), where the one-sided derivatives differ."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import baseline_params
from wcsph import scenes


def _setup(seed=3):
    desc = scenes.DamBreak(dims=(7, 7, 7), box_max=(0.3, 0.3, 0.3), seed=seed)
    fl, bd, box = scenes.dam_break(desc)
    rng = np.random.default_rng(seed)
    lo = fl[:, :3].min(axis=0)
    pos = fl.copy()
    pos[:, :3] = (lo + (fl[:, :3] - lo) * np.float32(0.85)).astype(np.float32)  # pressure active
    v0 = np.zeros((len(fl), 4))
    v0[:, :3] = rng.normal(0, 0.3, (len(fl), 3))
    return pos, bd, box, v0, rng


def _loss(prm, pos, bd, box, v, alpha, targets, steps, dt):
    s = O.Solver(prm, pos, bd, box, precision=64, parallel=False)
    s.set_state64(pos.astype(np.float64), v)
    s.set_viscosity(alpha)
    return s.grad_run(dt, steps, targets), s


@pytest.mark.parametrize("seed", [3, 4])
def test_fp64_adjoint_matches_central_differences(seed):
    pos, bd, box, v0, rng = _setup(seed)
    prm = baseline_params()
    steps, dt, alpha = 12, 2e-4, 0.01
    ref = O.Solver(prm, pos, bd, box, precision=64, parallel=False)
    ref.set_state64(pos.astype(np.float64), v0)
    ref.step(dt, steps)
    st = O.by_id(ref.state())
    assert st["p"].max() > 0 and st["vel"][:, :3].std() > 0.1
    targets = st["pos"] + rng.normal(0, 0.003, st["pos"].shape)
    L, s = _loss(prm, pos, bd, box, v0, alpha, targets, steps, dt)
    gv, ga = s.grad_backward()
    assert np.all(gv[:, 3] == 0)
    for _ in range(3):
        d = rng.normal(0, 1, v0.shape)
        d[:, 3] = 0
        eps = 1e-7
        lp, _ = _loss(prm, pos, bd, box, v0 + eps * d, alpha, targets, steps, dt)
        lm, _ = _loss(prm, pos, bd, box, v0 - eps * d, alpha, targets, steps, dt)
        fd = (lp - lm) / (2 * eps)
        assert np.sum(gv * d) == pytest.approx(fd, rel=1e-5)
    eps = 1e-6
    lp, _ = _loss(prm, pos, bd, box, v0, alpha + eps, targets, steps, dt)
    lm, _ = _loss(prm, pos, bd, box, v0, alpha - eps, targets, steps, dt)
    assert ga == pytest.approx((lp - lm) / (2 * eps), rel=1e-5)
    assert ga != 0.0
