"""GPU checks of the SPH adjoint (mfl_adjoint_*): central finite differences of the FP32 GPU rollout
on the 4,096-particle variant of the differentiable case, an independent FP64 CPU adjoint (torch
autograd of a float64 restatement of the step), checkpoint-interval invariance, and the program's
state after backward."""
import numpy as np
import pytest
import torch

from libcl_amd import scenes
from sph_autograd import SPHRollout, Simulation, cpu_rollout_fp64

pytestmark = pytest.mark.gpu


def _loss(xt, target):
    return ((xt.double() - target) ** 2).sum(1).mean()


def _setup(n_side=(16, 16, 16), steps=200, every=20, clear=65536):
    sc, target = scenes.config_grad(n_side)
    sim = Simulation(sc, steps, every, clear=clear)
    return sc, torch.from_numpy(target).cuda(), sim


def test_adjoint_matches_central_differences():
    """4,096 particles, 200 steps, checkpoints every 20; gradients of L = mean |x_T - target|^2 with
    respect to the initial velocities (directional derivative), viscosity, stiffness and rest density
    agree with central differences of the FP32 rollout (FP64 loss) to 1e-3 relative."""
    sc, target, sim = _setup()
    v0 = torch.zeros(sc.n, 3, dtype=torch.float32, device="cuda", requires_grad=True)
    mu = torch.tensor(6.0, dtype=torch.float64, requires_grad=True)
    k = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    rho0 = torch.tensor(1000.0, dtype=torch.float64, requires_grad=True)
    loss = _loss(SPHRollout.apply(v0, mu, k, rho0, sim), target)
    loss.backward()
    print("loss", loss.item(), "grads", mu.grad.item(), k.grad.item(), rho0.grad.item(), flush=True)

    def L(v, m_, k_, r_):
        pos = sim.run(v, m_, k_, r_)
        return _loss(torch.from_numpy(pos[:, :3]).cuda(), target).item()

    g = v0.grad.cpu().numpy().astype(np.float64)
    zero = np.zeros((sc.n, 3), np.float32)

    def central(fn, base, eps):
        """Richardson-extrapolated central difference (truncation O(eps^4))."""
        d1 = (fn(base + eps) - fn(base - eps)) / (2 * eps)
        d2 = (fn(base + eps / 2) - fn(base - eps / 2)) / eps
        return (4 * d2 - d1) / 3

    # along the gradient: the FD signal is |g|, far above the FP32 rounding noise of the rollout
    ghat = (g / np.linalg.norm(g)).astype(np.float32)
    fd_g = central(lambda t: L(t * ghat, 6.0, 1.0, 1000.0), 0.0, 0.5)
    print("v0 along gradient", np.linalg.norm(g), fd_g, flush=True)
    assert abs(np.linalg.norm(g) - fd_g) <= 1e-3 * abs(fd_g)
    # along a random direction the signal is ~|g|/sqrt(3N), two orders smaller: the same check holds
    # to the rounding-noise floor of two FP32 rollouts (about 1e-9 in L)
    rng = np.random.default_rng(3)
    u = rng.normal(size=(sc.n, 3)).astype(np.float32)
    u /= np.linalg.norm(u)
    fd_u = central(lambda t: L(t * u, 6.0, 1.0, 1000.0), 0.0, 1.0)
    ad_u = float((g * u).sum())
    print("v0 random direction", ad_u, fd_u, flush=True)
    assert abs(ad_u - fd_u) <= 2e-2 * abs(fd_u)
    for name, base, eps, grad in (("viscosity", 6.0, 3.0, mu.grad.item()), ("stiffness", 1.0, 0.5, k.grad.item()),
                                  ("density", 1000.0, 50.0, rho0.grad.item())):
        idx = ["viscosity", "stiffness", "density"].index(name)

        def fn(x, idx=idx):
            args = [6.0, 1.0, 1000.0]
            args[idx] = x
            return L(zero, *args)

        fd = central(fn, base, eps)
        print(name, grad, fd, flush=True)
        assert abs(grad - fd) <= 1e-3 * abs(fd), (name, grad, fd)


def test_adjoint_matches_fp64_cpu_autograd():
    """512 particles near a wall corner, 12 steps, full bucket clear: the hand-written adjoint of the
    FP32 GPU rollout agrees with torch autograd of the FP64 CPU restatement."""
    sc, target = scenes.config_grad((8, 8, 8))
    steps = 12
    sim = Simulation(sc, steps, 5, clear=262144)
    rng = np.random.default_rng(5)
    v0_np = (rng.normal(size=(sc.n, 3)) * 0.05).astype(np.float32)
    v0 = torch.from_numpy(v0_np).cuda().requires_grad_(True)
    mu = torch.tensor(6.0, dtype=torch.float64, requires_grad=True)
    k = torch.tensor(20.0, dtype=torch.float64, requires_grad=True)
    rho0 = torch.tensor(1000.0, dtype=torch.float64, requires_grad=True)
    tgt = torch.from_numpy(target)
    _loss(SPHRollout.apply(v0, mu, k, rho0, sim), tgt.cuda()).backward()

    v0c = torch.from_numpy(v0_np.astype(np.float64)).requires_grad_(True)
    muc = torch.tensor(6.0, dtype=torch.float64, requires_grad=True)
    kc = torch.tensor(20.0, dtype=torch.float64, requires_grad=True)
    rc = torch.tensor(1000.0, dtype=torch.float64, requires_grad=True)
    prm = sim.params(6.0, 20.0, 1000.0)
    xt = cpu_rollout_fp64(sc.pos, v0c, sc.force, prm, sc.box_min, sc.box_max, steps, muc, kc, rc)
    _loss(xt, tgt).backward()
    gv, gvc = v0.grad.cpu().numpy().astype(np.float64), v0c.grad.numpy()
    rel = np.linalg.norm(gv - gvc) / np.linalg.norm(gvc)
    print("v0 grad rel", rel, flush=True)
    assert rel < 1e-3
    for a, b in ((mu.grad, muc.grad), (k.grad, kc.grad), (rho0.grad, rc.grad)):
        print("param", a.item(), b.item(), flush=True)
        assert abs(a.item() - b.item()) <= 1e-3 * abs(b.item()) + 1e-12


def test_checkpoint_interval_does_not_change_gradients():
    sc, target = scenes.config_grad((16, 16, 8))
    grads = []
    for every in (1, 7, 40):
        sim = Simulation(sc, 40, every)
        v0 = torch.zeros(sc.n, 3, dtype=torch.float32, device="cuda", requires_grad=True)
        mu = torch.tensor(6.0, dtype=torch.float64, requires_grad=True)
        xt = SPHRollout.apply(v0, mu, torch.tensor(1.0, dtype=torch.float64), torch.tensor(1000.0, dtype=torch.float64),
                              sim)
        _loss(xt, torch.from_numpy(target).cuda()).backward()
        grads.append((v0.grad.cpu().numpy(), mu.grad.item(), sim.prog.state()["pos"], xt.detach().cpu().numpy()))
    for g, m, pos, xt in grads[1:]:
        np.testing.assert_allclose(g, grads[0][0], rtol=1e-4, atol=1e-6 * np.abs(grads[0][0]).max())
        assert abs(m - grads[0][1]) <= 1e-5 * abs(grads[0][1])
        # forward values do not depend on the checkpoint interval, and backward leaves the final state
        np.testing.assert_array_equal(xt, grads[0][3])
        np.testing.assert_array_equal(pos[:, :3], xt)
