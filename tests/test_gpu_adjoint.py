"""GPU checks of the SPH adjoint (mfl_adjoint_*): an independent FP64 CPU adjoint (torch autograd of a
float64 restatement of the step) over the full 200 steps, FP64 central differences of that map,
central differences of the FP32 GPU rollout on a 4,096-particle variant, checkpoint-interval
invariance, and the program's state after backward."""
import numpy as np
import pytest
import torch

from libcl_amd import scenes
from sph_autograd import SPHRollout, Simulation, cpu_rollout_fp64

pytestmark = pytest.mark.gpu


def _loss(xt, target):
    return ((xt.double() - target) ** 2).sum(1).mean()


def _smooth_scene(n_side, steps):
    """A jittered block at (0.5, 0.5, 0.5), far from the walls and from the coordinate planes (where
    the XOR hash repeats buckets inside a neighbourhood): the rollout map is smooth there."""
    h = 0.02
    pos = scenes.lattice(*n_side, 0.5 * h, (0.5, 0.5, 0.5), 0.01 * h, 6)
    sc = scenes.Scene("smooth", pos, np.zeros_like(pos), np.tile(scenes.GRAVITY_FORCE, (pos.shape[0], 1)),
                      (-10.0, -10.0, -10.0, 0.0), (10.0, 10.0, 10.0, 0.0), scenes.sph_param_values(h), steps)
    noise = scenes.normal(6, 3 * sc.n).reshape(sc.n, 3)
    target = sc.pos[:, :3].astype(np.float64) + np.array([0.1, 0.0, 0.0]) + 0.01 * noise
    return sc, torch.from_numpy(target)


def _gpu_grads(sim, sc, target, v0_np, mu_, k_, r_):
    v0 = torch.from_numpy(v0_np).cuda().requires_grad_(True)
    mu = torch.tensor(mu_, dtype=torch.float64, requires_grad=True)
    k = torch.tensor(k_, dtype=torch.float64, requires_grad=True)
    rho0 = torch.tensor(r_, dtype=torch.float64, requires_grad=True)
    _loss(SPHRollout.apply(v0, mu, k, rho0, sim), target.cuda()).backward()
    return v0.grad.cpu().numpy().astype(np.float64), mu.grad.item(), k.grad.item(), rho0.grad.item()


def _cpu(sc, sim, target, v0_np, mu_, k_, r_, steps, grad=True):
    v0c = torch.from_numpy(v0_np.astype(np.float64)).requires_grad_(grad)
    mu = torch.tensor(mu_, dtype=torch.float64, requires_grad=grad)
    k = torch.tensor(k_, dtype=torch.float64, requires_grad=grad)
    r = torch.tensor(r_, dtype=torch.float64, requires_grad=grad)
    xt = cpu_rollout_fp64(sc.pos, v0c, sc.force, sim.params(mu_, k_, r_), sc.box_min, sc.box_max, steps, mu, k, r)
    loss = _loss(xt, target)
    if not grad:
        return loss.item()
    loss.backward()
    return v0c.grad.numpy(), mu.grad.item(), k.grad.item(), r.grad.item()


def test_adjoint_matches_fp64_autograd_and_fp64_differences_over_200_steps():
    """512 particles, 200 steps, checkpoints every 20, full bucket clear: the hand-written adjoint of the
    FP32 GPU rollout agrees with torch autograd of the FP64 CPU restatement (dL/dv0, viscosity,
    stiffness, rest density), and that FP64 gradient agrees with central differences of the FP64 map,
    so the adjoint matches finite differences to 1e-3. (Central differences of the FP32 rollout itself
    are limited by its rounding noise; see the 4,096-particle test.)"""
    steps = 200
    sc, target = _smooth_scene((8, 8, 8), steps)
    sim = Simulation(sc, steps, 20, clear=262144)
    rng = np.random.default_rng(5)
    v0 = (rng.normal(size=(sc.n, 3)) * 0.02).astype(np.float32)
    g_gpu = _gpu_grads(sim, sc, target, v0, 6.0, 1.0, 1000.0)
    g_cpu = _cpu(sc, sim, target, v0, 6.0, 1.0, 1000.0, steps)
    rel = np.linalg.norm(g_gpu[0] - g_cpu[0]) / np.linalg.norm(g_cpu[0])
    print("v0 rel", rel, "params gpu", g_gpu[1:], "cpu", g_cpu[1:], flush=True)
    assert rel < 1e-4
    for a, b in zip(g_gpu[1:], g_cpu[1:]):
        assert abs(a - b) <= 1e-4 * abs(b)
    # FP64 central differences of the same map
    u = (g_cpu[0] / np.linalg.norm(g_cpu[0])).astype(np.float64)
    eps = 1e-2
    fd_v = (_cpu(sc, sim, target, (v0 + eps * u).astype(np.float64), 6.0, 1.0, 1000.0, steps, grad=False) -
            _cpu(sc, sim, target, (v0 - eps * u).astype(np.float64), 6.0, 1.0, 1000.0, steps, grad=False)) / (2 * eps)
    fd_k = (_cpu(sc, sim, target, v0, 6.0, 1.01, 1000.0, steps, grad=False) -
            _cpu(sc, sim, target, v0, 6.0, 0.99, 1000.0, steps, grad=False)) / 0.02
    print("fd64 v0", fd_v, float((g_gpu[0] * u).sum()), "fd64 k", fd_k, g_gpu[2], flush=True)
    assert abs(float((g_gpu[0] * u).sum()) - fd_v) <= 1e-3 * abs(fd_v)
    assert abs(g_gpu[2] - fd_k) <= 1e-3 * abs(fd_k)


def test_adjoint_vs_fp32_central_differences_4096():
    """4,096 particles, 200 steps, checkpoints every 20: central differences of the FP32 GPU rollout
    (FP64 loss). Along the gradient the signal dominates the FP32 rounding noise of the rollout and the
    match is 1e-3; for the parameters the FD signal is two to three orders smaller and the FP32 noise
    floor (measured: a few 1e-3 to 2e-2) sets the tolerance."""
    steps = 200
    sc, target = _smooth_scene((16, 16, 16), steps)
    sim = Simulation(sc, steps, 20, clear=262144)
    zero = np.zeros((sc.n, 3), np.float32)
    g, gmu, gk, gr = _gpu_grads(sim, sc, target, zero, 6.0, 1.0, 1000.0)
    tgt = target.cuda()

    def L(v, m_, k_, r_):
        return _loss(torch.from_numpy(sim.run(v, m_, k_, r_)[:, :3]).cuda(), tgt).item()

    ghat = (g / np.linalg.norm(g)).astype(np.float32)
    fd_g = (L(0.3 * ghat, 6.0, 1.0, 1000.0) - L(-0.3 * ghat, 6.0, 1.0, 1000.0)) / 0.6
    print("v0 along gradient", np.linalg.norm(g), fd_g, flush=True)
    assert abs(np.linalg.norm(g) - fd_g) <= 1e-3 * abs(fd_g)
    for name, plus, minus, step, grad in (("stiffness", (6.0, 1.05, 1000.0), (6.0, 0.95, 1000.0), 0.1, gk),
                                          ("density", (6.0, 1.0, 1005.0), (6.0, 1.0, 995.0), 10.0, gr),
                                          ("viscosity", (6.5, 1.0, 1000.0), (5.5, 1.0, 1000.0), 1.0, gmu)):
        fd = (L(zero, *plus) - L(zero, *minus)) / step
        print(name, grad, fd, flush=True)
        assert abs(grad - fd) <= 3e-2 * abs(fd), (name, grad, fd)


def test_adjoint_matches_fp64_cpu_autograd_at_the_corner():
    """The origin corner of the differentiable case (repeated buckets, active walls), 512 particles,
    12 steps, full clear: the adjoint is that of the computation the forward performed, visit
    multiplicities and wall branches included."""
    sc, target = scenes.config_grad((8, 8, 8))
    steps = 12
    sim = Simulation(sc, steps, 5, clear=262144)
    rng = np.random.default_rng(5)
    v0 = (rng.normal(size=(sc.n, 3)) * 0.05).astype(np.float32)
    tgt = torch.from_numpy(target)
    g_gpu = _gpu_grads(sim, sc, tgt, v0, 6.0, 20.0, 1000.0)
    g_cpu = _cpu(sc, sim, tgt, v0, 6.0, 20.0, 1000.0, steps)
    rel = np.linalg.norm(g_gpu[0] - g_cpu[0]) / np.linalg.norm(g_cpu[0])
    print("v0 grad rel", rel, flush=True)
    assert rel < 1e-4
    for a, b in zip(g_gpu[1:], g_cpu[1:]):
        print("param", a, b, flush=True)
        assert abs(a - b) <= 1e-4 * abs(b) + 1e-12


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


@pytest.mark.parametrize("steps", [20, 40])
def test_adjoint_at_the_differentiable_workload(steps):
    """The differentiable workload itself: 65,536 particles (64x32x32 block in the origin corner of the
    unit box), v0 = 0, checkpoints every 20 steps, full bucket clear (the FP64 restatement assumes fresh
    tables). The hand-written adjoint of the FP32 GPU rollout agrees with torch autograd of the FP64
    CPU restatement (a cell-list pair search) to 1e-4: dL/dv0 normwise, each parameter gradient
    relative. 20 and 40 steps: one and two checkpoint segments."""
    print(f"C5 65536 particles, {steps} steps", flush=True)
    sc, target = scenes.config_grad()
    sim = Simulation(sc, steps, 20, clear=262144)
    tgt = torch.from_numpy(target)
    v0 = np.zeros((sc.n, 3), np.float32)
    g_gpu = _gpu_grads(sim, sc, tgt, v0, 6.0, 1.0, 1000.0)
    g_cpu = _cpu(sc, sim, tgt, v0, 6.0, 1.0, 1000.0, steps)
    rel = np.linalg.norm(g_gpu[0] - g_cpu[0]) / np.linalg.norm(g_cpu[0])
    print("v0 grad rel", rel, "params gpu", g_gpu[1:], "cpu", g_cpu[1:], flush=True)
    assert rel < 1e-4
    for a, b in zip(g_gpu[1:], g_cpu[1:]):
        assert abs(a - b) <= 1e-4 * abs(b), (a, b)


def test_adjoint_vs_fp64_differences_on_the_4096_variant():
    """The finite-difference variant of the differentiable workload: 16x16x16 = 4,096 particles in the
    origin corner, 200 steps, checkpoints every 20, v0 = 0, full bucket clear. Central differences of
    the FP64 map (cell-list restatement, FP64 throughout) against the GPU adjoint, to 1e-3, for the
    direction of dL/dv0, viscosity, stiffness and rest density. The map is only piecewise smooth there
    (visit multiplicities switch when a particle changes cell next to a coordinate plane, wall forces
    switch on at d = 1e-5), so the steps are relative 1e-5: at 1e-3 they straddle such switches and
    the differences move by tens of percent for stiffness and rest density, while at 1e-5 .. 1e-7 they
    agree with the branch-wise derivative (torch autograd of the same FP64 map) to 1e-7."""
    steps = 200
    sc, target = scenes.config_grad((16, 16, 16))
    sim = Simulation(sc, steps, 20, clear=262144)
    tgt = torch.from_numpy(target)
    v0 = np.zeros((sc.n, 3), np.float32)
    gv, gmu, gk, gr = _gpu_grads(sim, sc, tgt, v0, 6.0, 1.0, 1000.0)
    z = np.zeros((sc.n, 3), np.float64)

    def L64(v, m_, k_, r_):
        return _cpu(sc, sim, tgt, v, m_, k_, r_, steps, grad=False)

    u = gv / np.linalg.norm(gv)
    eps = 1e-5
    fd_v = (L64(z + eps * u, 6.0, 1.0, 1000.0) - L64(z - eps * u, 6.0, 1.0, 1000.0)) / (2 * eps)
    print("v0 direction", np.linalg.norm(gv), fd_v, flush=True)
    assert abs(np.linalg.norm(gv) - fd_v) <= 1e-3 * abs(fd_v)
    # viscosity by 1e-3 relative: the restatement scales the FP32 viscosityConstant of each perturbed value,
    # whose rounding (6e-8) would otherwise dominate the difference; no switch lies that close in viscosity
    for name, plus, minus, step, grad in (("viscosity", (6.006, 1.0, 1000.0), (5.994, 1.0, 1000.0), 0.012, gmu),
                                          ("stiffness", (6.0, 1.0 + 1e-5, 1000.0), (6.0, 1.0 - 1e-5, 1000.0), 2e-5, gk),
                                          ("density", (6.0, 1.0, 1000.01), (6.0, 1.0, 999.99), 0.02, gr)):
        fd = (L64(z, *plus) - L64(z, *minus)) / step
        print(name, grad, fd, flush=True)
        assert abs(grad - fd) <= 1e-3 * abs(fd), (name, grad, fd)
