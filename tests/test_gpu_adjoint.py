"""Adjoint of the RK2 projection step (GPU, through the C ABI).

1. Every transposed piece satisfies <A x, y> = <x, A^T y> for random x, y,
   with A the forward kernel itself: boundary fills, forcing, correctVel,
   calcRHS, the PPE solve, and the Jacobian of the momentum right-hand side
   (its JVP is also checked against central differences of the RK1 kernel).
2. The window gradient dJ/db (J = mean wall shear over the window, b a body
   force) matches central finite differences of the GPU forward along seeded
   directions (bar 1e-6 relative), and of the oracle's FP64 forward on the
   CPU; it does not depend on the checkpoint spacing.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DOT_TOL = 1e-12


def _grid(nx, ny, nz, sr=3.0):
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ
    return Grid(nx, ny, nz, LX, LY, LZ, sr)


def _channel(g, nu=1e-4, dt=2e-4, u_crf=0.0, is_forced=(True, False, False),
             vel_force=(1.0, 0.0, 0.0)):
    from powerllel_channel_amd import Channel
    return Channel(g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.dzf, nu, dt, u_crf=u_crf,
                   is_forced=is_forced, vel_force=vel_force)


def _rand(torch, cuda, g, n, seed):
    gen = torch.Generator(device="cpu").manual_seed(seed)
    return [torch.randn(g.shape, generator=gen, dtype=torch.float64).to(cuda) for _ in range(n)]


def _interior(t):
    return t[1:-1, 1:-1, 1:-1]


def _dot(xs, ys, interior=False):
    s = 0.0
    for x, y in zip(xs, ys):
        if interior:
            x, y = _interior(x), _interior(y)
        s += float((x * y).sum())
    return s


def _close(a, b, tol=DOT_TOL):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


SHAPES = [(24, 16, 12), (32, 24, 16)]


@pytest.mark.parametrize("shape", SHAPES)
def test_bc_transposes(cuda, shape):
    import torch
    g = _grid(*shape)
    ch = _channel(g)
    print(f"[adjoint bc] {shape}", flush=True)
    x = _rand(torch, cuda, g, 3, 1)
    y = _rand(torch, cuda, g, 3, 2)
    Px = [t.clone() for t in x]
    ch.bc_vel(*Px)
    Pty = [t.clone() for t in y]
    ch.bc_vel_adj(*Pty)
    a, b = _dot(Px, y), _dot(x, Pty)
    print(f"  bc_vel  <Px,y> {a:.15e}  <x,P'y> {b:.15e}")
    assert _close(a, b)
    p = _rand(torch, cuda, g, 1, 3)[0]
    q = _rand(torch, cuda, g, 1, 4)[0]
    Pp = p.clone()
    ch.bc_p(Pp)
    Ptq = q.clone()
    ch.bc_p_adj(Ptq)
    a, b = _dot([Pp], [q]), _dot([p], [Ptq])
    print(f"  bc_p    <Px,y> {a:.15e}  <x,P'y> {b:.15e}")
    assert _close(a, b)


@pytest.mark.parametrize("forced", [(True, False, False), (True, True, True)])
def test_force_transpose(cuda, forced):
    import torch
    g = _grid(24, 16, 12)
    ch = _channel(g, is_forced=forced, vel_force=(0.0, 0.0, 0.0))  # linear part only
    x = _rand(torch, cuda, g, 3, 5)
    y = _rand(torch, cuda, g, 3, 6)
    Fx = [t.clone() for t in x]
    ch.force_vel(*Fx)
    Fty = [t.clone() for t in y]
    ch.force_vel_adj(*Fty)
    a, b = _dot(Fx, y), _dot(x, Fty)
    print(f"[adjoint force] {forced}: {a:.15e} {b:.15e}")
    assert _close(a, b)


def test_correct_and_rhs_transposes(cuda):
    import torch
    g = _grid(32, 24, 16)
    ch = _channel(g)
    p = _rand(torch, cuda, g, 1, 7)[0]
    lam = _rand(torch, cuda, g, 3, 8)
    z = [torch.zeros(g.shape, dtype=torch.float64, device=cuda) for _ in range(3)]
    ch.correct_vel(p, *z)  # z = -G p on the interior
    lp = torch.empty_like(p)
    ch.correct_vel_adj(*lam, lp)
    a, b = _dot(z, lam, interior=True), _dot([p], [lp])
    print(f"[adjoint correctVel] {a:.15e} {b:.15e}")
    assert _close(a, b)
    U = _rand(torch, cuda, g, 3, 9)
    lr = _rand(torch, cuda, g, 1, 10)[0]
    r = torch.zeros_like(p)
    ch.calc_rhs(*U, r)
    acc = [torch.zeros_like(p) for _ in range(3)]
    ch.calc_rhs_adj(lr, *acc)
    a, b = _dot([r], [lr], interior=True), _dot(U, acc)
    print(f"[adjoint calcRHS] {a:.15e} {b:.15e}")
    assert _close(a, b)


@pytest.mark.parametrize("shape,sr", [((32, 24, 16), 3.0), ((64, 48, 32), 4.0)])
def test_ppe_transpose(cuda, shape, sr):
    import torch
    from powerllel_channel_amd import PoissonSolver
    g = _grid(*shape, sr=sr)
    s = PoissonSolver(g.nx, g.ny, g.nz, g.dx, g.dy, g.dzf)
    f = _rand(torch, cuda, g, 1, 11)[0]
    h = _rand(torch, cuda, g, 1, 12)[0]
    Sf = f.clone()
    s.solve(Sf)
    Sth = h.clone()
    s.solve_adjoint(Sth)
    a, b = _dot([Sf], [h], interior=True), _dot([f], [Sth], interior=True)
    print(f"[adjoint PPE] {shape} sr={sr}: {a:.15e} {b:.15e} rel {abs(a-b)/abs(a):.1e}")
    # the solves themselves carry their (conditioning-limited) roundoff
    assert _close(a, b, 1e-10)


def test_rk_jacobian(cuda):
    import torch
    from powerllel_channel_amd.workloads import poiseuille_fields
    g = _grid(32, 24, 16)
    ch = _channel(g, u_crf=0.3)
    q = [torch.from_numpy(a).to(cuda) for a in poiseuille_fields(g, 2)]
    ch.bc_vel(*q)
    dq = _rand(torch, cuda, g, 3, 13)
    l = _rand(torch, cuda, g, 3, 14)
    Jdq = [torch.zeros(g.shape, dtype=torch.float64, device=cuda) for _ in range(3)]
    ch.rk_jvp(q, dq, Jdq)
    Jtl = [torch.zeros(g.shape, dtype=torch.float64, device=cuda) for _ in range(3)]
    ch.rk_vjp(q, l, Jtl, alpha=1.0, a0=0.0, beta=0.0)
    a, b = _dot(Jdq, l, interior=True), _dot(dq, Jtl)
    print(f"[adjoint RK] <J dq, l> {a:.15e}  <dq, J'l> {b:.15e}")
    assert _close(a, b)
    # the JVP is the derivative of the RK1 kernel itself: (u1(q+e dq) - u1(q-e dq)) / (2 e dt)
    eps = 1e-4
    outs = []
    for sgn in (1.0, -1.0):
        qs = [a_ + sgn * eps * d for a_, d in zip(q, dq)]
        o = [torch.zeros(g.shape, dtype=torch.float64, device=cuda) for _ in range(3)]
        ch.rk1(*qs, *o)
        outs.append(o)
    for c in range(3):
        fd = _interior(outs[0][c] - outs[1][c]) / (2 * eps * 2e-4) - _interior(dq[c]) / 2e-4
        rel = float((fd - _interior(Jdq[c])).abs().max() / _interior(Jdq[c]).abs().max())
        print(f"  component {c}: JVP vs central difference {rel:.1e}")
        assert rel < 1e-6


def _window_setup(torch, cuda, g, seed=3, amp=1e-3):
    from powerllel_channel_amd.workloads import poiseuille_fields
    u, v, w = poiseuille_fields(g, seed)
    U0 = [torch.from_numpy(a).to(cuda) for a in (u, v, w)]
    gen = torch.Generator(device="cpu").manual_seed(seed + 100)
    B = []
    for _ in range(3):
        b = torch.zeros(g.shape, dtype=torch.float64)
        b[1:-1, 1:-1, 1:-1] = amp * torch.randn((g.nz, g.ny, g.nx), generator=gen,
                                               dtype=torch.float64)
        B.append(b.to(cuda))
    return U0, B


def _directions(torch, cuda, g, n, seed):
    gen = torch.Generator(device="cpu").manual_seed(seed)
    out = []
    for _ in range(n):
        d = []
        for _ in range(3):
            t = torch.zeros(g.shape, dtype=torch.float64)
            t[1:-1, 1:-1, 1:-1] = torch.randn((g.nz, g.ny, g.nx), generator=gen,
                                              dtype=torch.float64)
            d.append(t.to(cuda))
        out.append(d)
    return out


def _fd_check(torch, cuda, g, ch, s, nsteps, every, ndir, eps, tol, report_eps=None):
    from powerllel_channel_amd.adjoint import WallShearWindow
    U0, B = _window_setup(torch, cuda, g)
    ch.bc_vel(*U0)
    win = WallShearWindow(ch, s, nsteps, every)
    J = win.forward(*U0, *B)
    grad = [torch.empty(g.shape, dtype=torch.float64, device=cuda) for _ in range(3)]
    win.backward(*grad)
    worst = 0.0
    for d in _directions(torch, cuda, g, ndir, 77):
        Jp = win.forward(*U0, *[b + eps * x for b, x in zip(B, d)])
        Jm = win.forward(*U0, *[b - eps * x for b, x in zip(B, d)])
        fd = (Jp - Jm) / (2 * eps)
        ad = _dot(grad, d, interior=True)
        rel = abs(fd - ad) / abs(fd)
        worst = max(worst, rel)
        print(f"  J {J:.12e}  FD {fd:.12e}  adjoint {ad:.12e}  rel {rel:.1e}")
        if report_eps:
            Jp = win.forward(*U0, *[b + report_eps * x for b, x in zip(B, d)])
            Jm = win.forward(*U0, *[b - report_eps * x for b, x in zip(B, d)])
            fd2 = (Jp - Jm) / (2 * report_eps)
            print(f"    eps {report_eps:.0e}: FD {fd2:.12e}  rel {abs(fd2 - ad) / abs(fd2):.1e}")
    win.close()
    assert worst <= tol, worst
    return grad


def test_window_gradient_vs_fd_small(cuda):
    import torch
    from powerllel_channel_amd import PoissonSolver
    g = _grid(32, 24, 16, sr=3.0)
    ch = _channel(g)
    s = PoissonSolver(g.nx, g.ny, g.nz, g.dx, g.dy, g.dzf)
    print("[adjoint window] 32x24x16, 6 steps, checkpoints every 2", flush=True)
    g2 = _fd_check(torch, cuda, g, ch, s, nsteps=6, every=2, ndir=8, eps=1e-2, tol=1e-6)
    # checkpoint spacing changes only the recomputation, not the gradient
    from powerllel_channel_amd.adjoint import WallShearWindow
    U0, B = _window_setup(torch, cuda, g)
    ch.bc_vel(*U0)
    for every in (1, 6):
        win = WallShearWindow(ch, s, 6, every)
        win.forward(*U0, *B)
        gr = [torch.empty(g.shape, dtype=torch.float64, device=cuda) for _ in range(3)]
        win.backward(*gr)
        for a, b in zip(gr, g2):
            assert torch.equal(a, b), f"ckpt_every={every}"
        win.close()


def test_window_gradient_vs_oracle_fd(orc, cuda):
    """Directional derivatives of the GPU adjoint against central differences
    of the oracle's FP64 forward (the reference's step on the CPU, plus the
    body force), and the forward objective against the oracle's."""
    import torch
    from powerllel_channel_amd import PoissonSolver
    from powerllel_channel_amd.adjoint import WallShearWindow
    g = _grid(16, 12, 8, sr=3.0)
    nu, dt, nsteps = 1e-4, 2e-4, 4
    ch = _channel(g, nu=nu, dt=dt)
    s = PoissonSolver(g.nx, g.ny, g.nz, g.dx, g.dy, g.dzf)
    U0, B = _window_setup(torch, cuda, g)
    ch.bc_vel(*U0)
    win = WallShearWindow(ch, s, nsteps, 2)
    J = win.forward(*U0, *B)
    grad = [torch.empty(g.shape, dtype=torch.float64, device=cuda) for _ in range(3)]
    win.backward(*grad)
    G = orc.OracleGrid(g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.dzf, nu, dt)
    ppe = orc.OraclePPE(g.nx, g.ny, g.nz, g.dx, g.dy, g.dzf, pin_gauge=True)

    def J_oracle(Bh):
        u, v, w = [t.cpu().numpy().copy() for t in U0]
        z = [np.zeros(g.shape) for _ in range(4)]
        acc = 0.0
        for _ in range(nsteps):
            orc.step_body(G, ppe, u, v, w, *z, *Bh)
            acc += orc.wall_shear(G, u)
        return acc / nsteps

    Bh = [b.cpu().numpy().copy() for b in B]
    Jo = J_oracle(Bh)
    print(f"[adjoint vs oracle] J gpu {J:.15e} oracle {Jo:.15e}")
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    eps = 1e-2
    for d in _directions(torch, cuda, g, 4, 91):
        dh = [x.cpu().numpy() for x in d]
        fd = (J_oracle([b + eps * x for b, x in zip(Bh, dh)]) -
              J_oracle([b - eps * x for b, x in zip(Bh, dh)])) / (2 * eps)
        ad = _dot(grad, d, interior=True)
        rel = abs(fd - ad) / abs(fd)
        print(f"  oracle FD {fd:.12e}  adjoint {ad:.12e}  rel {rel:.1e}")
        assert rel <= 1e-6
    win.close()


def test_autograd_function(cuda):
    import torch
    from powerllel_channel_amd import PoissonSolver
    from powerllel_channel_amd.adjoint import WallShearWindow, wall_shear_objective
    g = _grid(24, 16, 12)
    ch = _channel(g)
    s = PoissonSolver(g.nx, g.ny, g.nz, g.dx, g.dy, g.dzf)
    U0, B = _window_setup(torch, cuda, g)
    ch.bc_vel(*U0)
    win = WallShearWindow(ch, s, 3, 2)
    Bt = [b.clone().requires_grad_(True) for b in B]
    J = wall_shear_objective(win, *U0, *Bt)
    (2.0 * J).backward()
    ref = [torch.empty(g.shape, dtype=torch.float64, device=cuda) for _ in range(3)]
    win.forward(*U0, *B)
    win.backward(*ref)
    for b, r in zip(Bt, ref):
        assert torch.equal(b.grad, 2.0 * r)


@pytest.mark.slow
def test_window_gradient_c4(cuda):
    """256x192x128, nu = 1e-4, dt = 2e-4, 20 steps, checkpoints every 5, body
    force N(0, 1e-3^2) per cell, 16 seeded directions."""
    import torch
    from powerllel_channel_amd import PoissonSolver
    g = _grid(256, 192, 128, sr=3.0)
    ch = _channel(g)
    s = PoissonSolver(g.nx, g.ny, g.nz, g.dx, g.dy, g.dzf)
    print("[adjoint window] 256x192x128, 20 steps, checkpoints every 5", flush=True)
    _fd_check(torch, cuda, g, ch, s, nsteps=20, every=5, ndir=16, eps=1e-2, tol=1e-6,
              report_eps=1e-6)
