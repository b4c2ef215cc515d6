"""Hand-written adjoint: velocity sensitivity kernel, LinearSolve gradients
and the coupled drift-diffusion model's gradients.

References used here: finite differences of the forward operator, a dense
FP64 torch assembly of the same Scharfetter-Gummel operator differentiated
by torch autograd (16^3 copy), and central finite differences of the whole
coupled run (relative step 1e-6).
"""
import numpy as np
import pytest
import torch

from mps3d_amd import FieldSolver, ghosted
from mps3d_amd import workloads as W
from mps3d_amd.autograd import Operator, linear_solve
from mps3d_amd.coupled import GlowModel, central_difference, seeded_source, source_pulse

from dense_sg import dense_sg

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0))).cuda()


@pytest.mark.parametrize("flux", ["sg", "upwind"])
def test_velocity_sensitivity_matches_finite_differences(gpu, flux):
    """-d(lam^T A(v) x)/dv from mps_grad_velocity vs a central difference of
    lam^T A(v) x evaluated through find_AX, along a random direction."""
    c = W.random_cdr((10, 8, 6), seed=17)
    rng = np.random.default_rng(3)
    lam = rng.standard_normal(c.n)
    x = rng.standard_normal(c.n)
    dirs = [rng.standard_normal(c.n) for _ in range(3)]
    vel0 = [c.velx + 0.7 * np.sign(c.velx), c.vely, c.velz]  # keep upwind faces away from c = 0

    def pairing(t):
        with FieldSolver(c.n, c.length, flux=flux) as s:
            s.set_bc(c.bc_codes, c.bc_vals)
            s.set_coeffs(c.dcoeff, *[v + t * d for v, d in zip(vel0, dirs)], c.reac, c.source, True, c.dt)
            ax, _, _ = s.find_AX(ghosted(x))
        return float((lam * ax).sum())

    with FieldSolver(c.n, c.length, flux=flux) as s:
        s.set_bc(c.bc_codes, c.bc_vals)
        s.set_coeffs(c.dcoeff, *vel0, c.reac, c.source, True, c.dt)
        lg, xg = dev(ghosted(lam)), dev(ghosted(x))
        out = [torch.empty((c.n[2], c.n[1], c.n[0]), dtype=torch.float64, device="cuda") for _ in range(3)]
        import ctypes as C
        from mps3d_amd import _lib as M
        M.check(s.lib.mps_grad_velocity(s._h, C.c_void_p(lg.data_ptr()), C.c_void_p(xg.data_ptr()),
                                        *[C.c_void_p(o.data_ptr()) for o in out]))
    g = [o.cpu().numpy().transpose(2, 1, 0) for o in out]
    adj = -sum(float((gi * di).sum()) for gi, di in zip(g, dirs))  # d(pairing)/dt
    eps = 1e-6
    fd = (pairing(eps) - pairing(-eps)) / (2 * eps)
    assert abs(adj - fd) <= 1e-6 * abs(fd)


def test_linear_solve_gradients_match_dense_fp64_autograd(gpu):
    """16^3 copy: gradients of J = w . A(v)^-1 b with respect to b and the
    three velocity fields, hand-written adjoint vs torch autograd through a
    dense FP64 solve of the same operator."""
    n = (16, 16, 16)
    L = 0.01
    h = L / 16
    dt = 1e-9
    codes = ("ZGRD", "ZGRD", "ZGRD", "ZGRD", "DIRC", "DIRC")
    g = torch.Generator().manual_seed(11)
    shp = (16, 16, 16)
    D = (50.0 + 50.0 * torch.rand(shp, generator=g, dtype=torch.float64)).cuda()
    vel = [(2e5 * torch.randn(shp, generator=g, dtype=torch.float64)).cuda() for _ in range(3)]
    b = (1e20 * torch.rand(shp, generator=g, dtype=torch.float64)).cuda()
    w = torch.randn(shp, generator=g, dtype=torch.float64).cuda()

    op = Operator(n, (L, L, L), D, codes, flux="sg", timederiv=True, dt=dt, tol=1e-13)
    bb = b.clone().requires_grad_(True)
    vv = [v.clone().requires_grad_(True) for v in vel]
    x = linear_solve(op, bb, vv)
    J = (w * x).sum()
    J.backward()
    ours = [bb.grad] + [v.grad for v in vv]
    op.close()

    bd = b.clone().requires_grad_(True)
    vd = [v.clone().requires_grad_(True) for v in vel]
    A = dense_sg(n, h, D, vd, codes, 1.0 / dt)
    xd = torch.linalg.solve(A, bd.reshape(-1)).reshape(shp)
    assert torch.linalg.norm(x.detach() - xd.detach()) / torch.linalg.norm(xd) < 1e-11
    Jd = (w * xd).sum()
    Jd.backward()
    ref = [bd.grad] + [v.grad for v in vd]
    for o, r, name in zip(ours, ref, ("b", "vx", "vy", "vz")):
        rel = float(torch.linalg.norm(o - r) / torch.linalg.norm(r))
        assert rel < 1e-8, (name, rel)


def test_linear_solve_with_one_velocity_component(gpu):
    """A velocity given as (None, None, vz): backward returns a gradient for
    vz only (the missing components have no gradient slot), equal to torch
    autograd through the dense FP64 solve with vx = vy = 0."""
    n = (8, 8, 8)
    L = 0.01
    h = L / 8
    dt = 1e-9
    codes = ("ZGRD", "ZGRD", "ZGRD", "ZGRD", "DIRC", "DIRC")
    g = torch.Generator().manual_seed(5)
    shp = (8, 8, 8)
    D = (50.0 + 50.0 * torch.rand(shp, generator=g, dtype=torch.float64)).cuda()
    vz = (2e5 * torch.randn(shp, generator=g, dtype=torch.float64)).cuda()
    b = (1e20 * torch.rand(shp, generator=g, dtype=torch.float64)).cuda()
    w = torch.randn(shp, generator=g, dtype=torch.float64).cuda()
    op = Operator(n, (L, L, L), D, codes, flux="sg", timederiv=True, dt=dt, tol=1e-13)
    bb = b.clone().requires_grad_(True)
    zz = vz.clone().requires_grad_(True)
    x = linear_solve(op, bb, (None, None, zz))
    (w * x).sum().backward()
    op.close()
    bd = b.clone().requires_grad_(True)
    zd = vz.clone().requires_grad_(True)
    zero = torch.zeros_like(vz)
    A = dense_sg(n, h, D, [zero, zero, zd], codes, 1.0 / dt)
    xd = torch.linalg.solve(A, bd.reshape(-1)).reshape(shp)
    (w * xd).sum().backward()
    for o, r, name in ((bb.grad, bd.grad, "b"), (zz.grad, zd.grad, "vz")):
        rel = float(torch.linalg.norm(o - r) / torch.linalg.norm(r))
        assert rel < 1e-8, (name, rel)


def glow_gradients(model, nsteps, ckpt, V0=1000.0, s0=1.0, seed=20261202):
    nz, ny, nx = model.n[2], model.n[1], model.n[0]
    q0 = seeded_source((nz, ny, nx), seed=seed, device="cuda")
    V = torch.tensor(V0, dtype=torch.float64, device="cuda", requires_grad=True)
    s = torch.tensor(s0, dtype=torch.float64, device="cuda", requires_grad=True)
    q = q0.clone().requires_grad_(True)
    J = model.run(V, s, q, nsteps, checkpoint_every=ckpt)
    J.backward()
    return float(J.detach()), float(V.grad), float(s.grad), q.grad.detach(), q0


def fd_checks(model, nsteps, grads, V0=1000.0, s0=1.0, rel=1e-6):
    """Central differences (relative step 1e-6) for V_app and s_mu, and for
    the source field along a localised pulse p (peak 1e22 m^-3 s^-1, about
    5e-4 of the peak density over 50 steps): q = q0 +- p. J is linear in q
    up to the weak space-charge feedback, so the unit step costs no accuracy,
    while a relative step on the seeded N(0, 1e13) field itself would move J
    by ~1e-13 relatively, below what an FP64 difference resolves. The full
    dJ/dq field is checked against FP64 autograd on the 16^3 copy."""
    J, gV, gs, gq, q0 = grads
    t64 = lambda v: torch.tensor(v, dtype=torch.float64, device="cuda")  # noqa: E731

    def run(V, s, q):
        with torch.no_grad():
            return float(model.run(t64(V), t64(s), q, nsteps))

    p = source_pulse(model, amplitude=1e22)
    fdV = central_difference(lambda v: run(v, s0, q0), V0, rel)
    fds = central_difference(lambda sv: run(V0, sv, q0), s0, rel)
    fdq = (run(V0, s0, q0 + p) - run(V0, s0, q0 - p)) / 2.0
    adq = float((gq * p).sum())
    return (gV, fdV), (gs, fds), (adq, fdq)


class DenseGlow(GlowModel):
    """The same coupled model with every linear solve replaced by a dense
    FP64 torch solve of the same operator, so torch autograd differentiates
    the whole run (the 16^3 reference copy)."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        nx, ny, nz = self.n
        shp = (nz, ny, nx)
        z = torch.zeros(shp, dtype=torch.float64, device="cuda")
        one = torch.ones(shp, dtype=torch.float64, device="cuda")
        self.codes = ("ZGRD", "ZGRD", "ZGRD", "ZGRD", "DIRC", "DIRC")
        self.Ap = dense_sg(self.n, self.h, one, (z, z, z), self.codes, 0.0, mean="arith")
        self.De = self.D_e * one
        self.shp = shp

    def solve_poisson(self, rhs):
        return torch.linalg.solve(self.Ap, rhs.reshape(-1)).reshape(self.shp)

    def solve_electrons(self, rhs, vel, x0):
        A = dense_sg(self.n, self.h, self.De, vel, self.codes, 1.0 / self.dt)
        return torch.linalg.solve(A, rhs.reshape(-1)).reshape(self.shp)


def test_coupled_gradients_match_dense_fp64_autograd_16cubed(gpu):
    """Full gradient fields (V_app, s_mu and dJ/dq for every cell) of the
    coupled run: hand-written adjoint with checkpointing vs FP64 torch
    autograd through dense solves of the same operators."""
    nsteps = 8
    ours = glow_gradients(GlowModel(n=16, dt=1e-11), nsteps, ckpt=3)
    ref = glow_gradients(DenseGlow(n=16, dt=1e-11), nsteps, ckpt=0)
    assert abs(ours[0] - ref[0]) <= 1e-10 * abs(ref[0])
    for name, a, b in (("V_app", ours[1], ref[1]), ("s_mu", ours[2], ref[2])):
        assert abs(a - b) <= 1e-7 * abs(b), (name, a, b)
    rel = float(torch.linalg.norm(ours[3] - ref[3]) / torch.linalg.norm(ref[3]))
    assert rel <= 1e-7, rel


def test_coupled_gradients_match_central_differences_small(gpu):
    model = GlowModel(n=16, dt=1e-11, cycles=(30, 20))
    grads = glow_gradients(model, nsteps=6, ckpt=2)
    for name, (adj, fd) in zip(("V_app", "s_mu", "q"), fd_checks(model, 6, grads)):
        assert abs(adj - fd) <= 1e-6 * abs(fd), (name, adj, fd)
    model.close()


@pytest.mark.timeout(900)
def test_config4_adjoint_96cubed_50_steps(gpu):
    """96^3, 50 implicit steps, checkpoints every 10 steps: adjoint gradients
    for the applied voltage, the electron-mobility scale and the source field
    match central differences (relative step 1e-6) to 1e-6, and the adjoint
    pass costs at most 3 forward passes."""
    import time
    model = GlowModel(n=96, dt=1e-11, p=(3, 3, 3), cycles=(30, 20))  # 96 = 3 * 2^5: a 3^3 coarsest grid
    q0 = seeded_source((96, 96, 96), device="cuda")
    V = torch.tensor(1000.0, dtype=torch.float64, device="cuda")
    s = torch.tensor(1.0, dtype=torch.float64, device="cuda")
    with torch.no_grad():
        model.run(V, s, q0, 2)  # warm
        torch.cuda.synchronize()
        t = time.perf_counter()
        model.run(V, s, q0, 50)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t
    t = time.perf_counter()
    grads = glow_gradients(model, nsteps=50, ckpt=10)
    torch.cuda.synchronize()
    t_both = time.perf_counter() - t
    ratio = (t_both - t_fwd) / t_fwd  # adjoint pass (recompute + reverse) over one forward
    print("forward %.3f s, forward+adjoint %.3f s, adjoint/forward %.2f" % (t_fwd, t_both, ratio))
    for name, (adj, fd) in zip(("V_app", "s_mu", "q"), fd_checks(model, 50, grads)):
        print(name, adj, fd, abs(adj - fd) / abs(fd))
        assert abs(adj - fd) <= 1e-6 * abs(fd), (name, adj, fd)
    assert ratio <= 3.0
    model.close()
