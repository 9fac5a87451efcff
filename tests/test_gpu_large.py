"""The large one-GPU workload: 1536x1024x384 (604M cells), Lx = 4 pi,
Ly = 2 pi, tanh gamma = 2.2, nu = 1/40000, dt = 1e-4, seeded perturbed
Poiseuille start, 10 RK2 steps; FP64 PPE against the FP32 + refinement PPE at
a 1e-11 pressure tolerance; a 5-step adjoint window with checkpoints.

The oracle cannot step 604M cells in test time, so the checks here are the
size-independent ones.  The kernels this grid runs are pinned to the oracle
elsewhere at the same lengths and layouts: the x and y transform lengths 768
and 1024 by the PPE parity case 1536x1024x8, the z length 384 (12 rows per
lane, 32 lanes per column, 16-column tiles) by the cases 64x48x384 and
32x36x384 of tests/test_gpu_ppe.py, and the step kernels at 768x384x256 by
tests/test_gpu_step.py::test_steps_full_size_vs_oracle:
* after every projection step, the bulk velocity equals the forcing target
  (forceVel, src/mod_calcVel.f90:455-509), and max |div u| < 1e-12 (FP64)
  from step 2 on.  Step 1 projects the white-noise start, whose divergence
  is about 12 (noise 0.1 over dx = 0.008): what remains is about 1e-13 of
  it (1.4e-12 measured; the FP64 solve's relative residual), so step 1 is
  held to 1e-11;
* the refined solve's error against the FP64 solve per sweep on a step's
  right-hand side, the tolerance form at 1e-11 on it, and a 10-step run with
  the tolerance form against the FP64 run (velocity, pressure, divergence
  under the reference's run-time limit div_limit = 1e-9,
  examples/template/param.in:33);
* the adjoint gradient's directional derivatives against central differences
  of the forward window.
Device memory at this size: one haloed FP64 field is 4.87 GB; the adjoint
window with checkpoints every 2 of 5 steps holds 32 of them.
"""
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

CASE = "channel_1536x1024x384"
PRESSURE_TOL = 1e-11


def _setup(cuda):
    import torch
    from powerllel_channel_amd import Channel, PoissonSolver
    from powerllel_channel_amd.workloads import CASES, poiseuille_fields
    case = CASES[CASE]
    g = case.grid()
    ch = Channel(case.nx, case.ny, case.nz, g.lx, g.ly, g.lz, g.dzf, case.nu, case.dt)
    s = PoissonSolver(case.nx, case.ny, case.nz, g.dx, g.dy, g.dzf)
    U0 = []
    for a in poiseuille_fields(g, seed=11):
        U0.append(torch.from_numpy(a).to(cuda))
        del a
    ch.bc_vel(*U0)
    return case, g, ch, s, U0


def _wmean(torch, g, q, cuda):
    w = torch.tensor(g.dzf[1:-1], device=cuda).view(-1, 1, 1)
    return float((q.sum(dim=(1, 2), keepdim=True) * w).sum() / (w.sum() * g.nx * g.ny))


def _rel_mr(torch, g, a, b, cuda):
    """relative max-norm distance of the dzf-weighted mean-removed interiors"""
    ai, bi = a[1:-1, 1:-1, 1:-1], b[1:-1, 1:-1, 1:-1]
    d = (ai - _wmean(torch, g, ai, cuda)) - (bi - _wmean(torch, g, bi, cuda))
    return float(d.abs().max()) / float((bi - _wmean(torch, g, bi, cuda)).abs().max())


def _rel(a, b):
    return float((a - b).abs().max()) / float(b.abs().max())


def _run(torch, ch, s, U0, case, label, ppe_timing=True):
    u, v, w = [t.clone() for t in U0]
    u1, v1, w1, p = [torch.zeros_like(u) for _ in range(4)]
    s.set_timing(ppe_timing)
    divs = []
    for n in range(case.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ch.step(s, u, v, w, u1, v1, w1, p)
        e1.record()
        torch.cuda.synchronize()
        div, _ = ch.max_div(u, v, w)
        ub = ch.mean_h(u)
        if ppe_timing:
            ppe_ms = f"(PPE {sum(s.stage_times()):6.2f} ms)"
        else:
            ppe_ms = "(sweeps %d, estimate %.1e)" % s.last_refinement()
        divs.append(div)
        print(f"  [{label}] step {n + 1:2d}: {e0.elapsed_time(e1):7.2f} ms {ppe_ms} "
              f"max|div| {div:.2e}  bulk u {ub:.15f}", flush=True)
        assert abs(ub - 1.0) <= 1e-13
    s.set_timing(False)
    return (u, v, w, p), divs


def test_c5_fp64_vs_refined(cuda):
    import torch
    print(f"[large] {CASE}: FP64 run", flush=True)
    case, g, ch, s, U0 = _setup(cuda)
    (u, v, w, p), div64 = _run(torch, ch, s, U0, case, "fp64")
    assert div64[0] < 1e-11
    assert max(div64[1:]) < 1e-12

    # refined PPE per sweep on a step's right-hand side: predictor of the
    # final state (RK1, RK2 with their boundary fills), then calcRHS
    q = [t.clone() for t in (u, v, w)]
    q1 = [torch.zeros_like(u) for _ in range(3)]
    ch.rk1(*q, *q1)
    ch.bc_vel(*q1)
    ch.rk2(*q, *q1)
    ch.bc_vel(*q)
    rhs = torch.zeros_like(u)
    ch.calc_rhs(*q, rhs)
    del q, q1
    ref = rhs.clone()
    s.solve(ref)
    errs = {}
    for k in range(4):
        pk = rhs.clone()
        s.solve_refined(pk, k)
        torch.cuda.synchronize()
        errs[k] = _rel_mr(torch, g, pk, ref, cuda)
        del pk
    print("  refined error per sweep: " + ", ".join(f"k={k}: {e:.2e}" for k, e in errs.items()),
          flush=True)
    assert min(errs[2], errs[3]) <= PRESSURE_TOL, errs
    pk = rhs.clone()
    n, est = s.solve_refined_tol(pk, PRESSURE_TOL)
    e_tol = _rel_mr(torch, g, pk, ref, cuda)
    print(f"  tolerance form: {n} sweeps, estimate {est:.2e}, error {e_tol:.2e}", flush=True)
    assert e_tol <= PRESSURE_TOL
    del rhs, ref, pk

    # the same 10 steps with the refined PPE at the 1e-11 tolerance
    print(f"[large] {CASE}: refined run, tolerance {PRESSURE_TOL:.0e}", flush=True)
    s.set_refinement_tol(PRESSURE_TOL, 8)
    (ur, vr, wr, pr), div32 = _run(torch, ch, s, U0, case, "refined", ppe_timing=False)
    s.set_refinement(-1)
    ev = [_rel(a, b) for a, b in zip((ur, vr, wr), (u, v, w))]
    ep = _rel_mr(torch, g, pr, p, cuda)
    print(f"  refined vs FP64 after {case.steps} steps: velocity {max(ev):.2e} "
          f"pressure {ep:.2e} max|div| {max(div32):.2e}", flush=True)
    assert max(ev) <= 1e-10
    assert ep <= 1e-9
    assert max(div32) < 1e-9
    s.close()
    ch.close()


def test_c5_adjoint_window(cuda):
    """5 steps, checkpoints every 2, body force N(0, 1e-3^2) per interior cell,
    4 seeded directions; central differences with eps = 1e-2 (the response is
    close to linear in the body force, and a smaller step loses digits of
    J(b + eps d) - J(b - eps d) to roundoff of J itself)."""
    import torch
    from powerllel_channel_amd.adjoint import WallShearWindow
    torch.cuda.empty_cache()
    print(f"[large] {CASE}: 5-step adjoint window", flush=True)
    case, g, ch, s, U0 = _setup(cuda)
    gen = torch.Generator(device=cuda).manual_seed(2024)

    def interior_noise(scale):
        t = torch.zeros(g.shape, dtype=torch.float64, device=cuda)
        t[1:-1, 1:-1, 1:-1] = scale * torch.randn((g.nz, g.ny, g.nx), generator=gen,
                                                  dtype=torch.float64, device=cuda)
        return t

    B = [interior_noise(1e-3) for _ in range(3)]
    win = WallShearWindow(ch, s, 5, 2)
    print(f"  adjoint window device bytes {win.device_bytes() / 1e9:.1f} GB", flush=True)
    J = win.forward(*U0, *B)
    grad = [torch.empty_like(B[0]) for _ in range(3)]
    win.backward(*grad)
    torch.cuda.synchronize()
    eps = 1e-2
    worst = 0.0
    for _ in range(4):
        d = [interior_noise(1.0) for _ in range(3)]
        Jp = win.forward(*U0, *[b + eps * x for b, x in zip(B, d)])
        Jm = win.forward(*U0, *[b - eps * x for b, x in zip(B, d)])
        fd = (Jp - Jm) / (2 * eps)
        ad = sum(float((x[1:-1, 1:-1, 1:-1] * y[1:-1, 1:-1, 1:-1]).sum())
                 for x, y in zip(grad, d))
        rel = abs(fd - ad) / abs(fd)
        worst = max(worst, rel)
        print(f"  J {J:.12e}  FD {fd:.12e}  adjoint {ad:.12e}  rel {rel:.1e}", flush=True)
        del d
    win.close()
    s.close()
    ch.close()
    assert worst <= 1e-6, worst
