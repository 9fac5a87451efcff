"""GPU step kernels vs the oracle.

Stencils, boundary fills, calcRHS and correctVel restate the reference's
arithmetic with the same operation order and no fused multiply-add, so they
must match the oracle bit for bit.  Reductions (Mean_h) sum in a different
order: 1e-14 relative.  Whole steps go through the PPE solve: 1e-12 relative
on velocity after one step, 1e-10 after 100 steps (192x192x128).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(orc, nx, ny, nz, sr, nu=1 / 5600, dt=1e-3, u_crf=0.0, seed=5):
    from powerllel_channel_amd import Channel
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ
    g = Grid(nx, ny, nz, LX, LY, LZ, sr)
    G = orc.OracleGrid(nx, ny, nz, LX, LY, LZ, g.dzf, nu, dt)
    ch = Channel(nx, ny, nz, LX, LY, LZ, g.dzf, nu, dt, u_crf=u_crf)
    rng = np.random.default_rng(seed)
    fields = [rng.standard_normal(g.shape) for _ in range(7)]
    return g, G, ch, fields


def _dev(torch, cuda, *arrs):
    return [torch.from_numpy(a.copy()).to(cuda) for a in arrs]


def _host(*ts):
    return [t.cpu().numpy() for t in ts]


@pytest.mark.parametrize("shape", [(16, 12, 10), (64, 48, 32), (130, 66, 20)])
@pytest.mark.parametrize("u_crf", [0.0, 0.37])
def test_rk1_rk2_bitwise(orc, cuda, shape, u_crf):
    import torch
    nx, ny, nz = shape
    print(f"[rk bitwise] {shape} u_crf={u_crf}", flush=True)
    g, G, ch, (u, v, w, u1, v1, w1, _) = _setup(orc, nx, ny, nz, 3.0, u_crf=u_crf)
    du, dv, dw, du1, dv1, dw1 = _dev(torch, cuda, u, v, w, u1, v1, w1)
    ch.rk1(du, dv, dw, du1, dv1, dw1)
    orc.rk1(G, u, v, w, u1, v1, w1, u_crf)
    torch.cuda.synchronize()
    for a, b in zip(_host(du1, dv1, dw1), (u1, v1, w1)):
        assert np.array_equal(a, b)
    ch.rk2(du, dv, dw, du1, dv1, dw1)
    orc.rk2(G, u, v, w, u1, v1, w1, u_crf)
    torch.cuda.synchronize()
    for a, b in zip(_host(du, dv, dw), (u, v, w)):
        assert np.array_equal(a, b)


def test_rhs_correct_bc_bitwise(orc, cuda):
    import torch
    nx, ny, nz = 40, 24, 18
    print(f"[rhs/correct/bc bitwise] {nx}x{ny}x{nz}", flush=True)
    g, G, ch, (u, v, w, p, _, _, _) = _setup(orc, nx, ny, nz, 4.0, u_crf=0.25)
    du, dv, dw, dp = _dev(torch, cuda, u, v, w, p)
    ch.calc_rhs(du, dv, dw, dp)
    orc.calc_rhs(G, u, v, w, p)
    torch.cuda.synchronize()
    assert np.array_equal(dp.cpu().numpy(), p)
    ch.bc_p(dp)
    orc.bc_p(G, p)
    torch.cuda.synchronize()
    assert np.array_equal(dp.cpu().numpy(), p)
    ch.correct_vel(dp, du, dv, dw)
    orc.correct_vel(G, p, u, v, w)
    torch.cuda.synchronize()
    for a, b in zip(_host(du, dv, dw), (u, v, w)):
        assert np.array_equal(a, b)
    ch.bc_vel(du, dv, dw)
    orc.bc_vel(G, u, v, w, 0.25)
    torch.cuda.synchronize()
    for a, b in zip(_host(du, dv, dw), (u, v, w)):
        assert np.array_equal(a, b)


def test_force_mean_maxdiv(orc, cuda):
    import torch
    nx, ny, nz = 48, 32, 24
    print(f"[force/mean/maxdiv] {nx}x{ny}x{nz}", flush=True)
    g, G, ch, (u, v, w, _, _, _, _) = _setup(orc, nx, ny, nz, 3.0)
    du, dv, dw = _dev(torch, cuda, u, v, w)
    m_g = ch.mean_h(du)
    m_o = orc.mean_h(G, u)
    assert abs(m_g - m_o) <= 1e-14 * max(1.0, abs(m_o))
    val_g, ijk_g = ch.max_div(du, dv, dw)
    val_o, ijk_o = orc.max_div(G, u, v, w)
    assert val_g == val_o and ijk_g == ijk_o
    ch.force_vel(du, dv, dw)
    orc.force_vel(G, u, v, w)
    torch.cuda.synchronize()
    a = du.cpu().numpy()
    assert np.abs(a - u).max() <= 1e-14
    assert abs(ch.mean_h(du) - 1.0) <= 1e-14


def _poiseuille_start(orc, case, seed=1):
    from powerllel_channel_amd.workloads import poiseuille_fields
    g = case.grid()
    G = orc.OracleGrid(case.nx, case.ny, case.nz, g.lx, g.ly, g.lz, g.dzf, case.nu, case.dt)
    u, v, w = poiseuille_fields(g, seed)
    orc.bc_vel(G, u, v, w, 0.0)
    return g, G, u, v, w


def _run_steps(orc, cuda, case, nsteps):
    import torch
    from powerllel_channel_amd import Channel, PoissonSolver
    g, G, u, v, w = _poiseuille_start(orc, case)
    zeros = [np.zeros(g.shape) for _ in range(4)]
    u1, v1, w1, p = zeros
    du, dv, dw, du1, dv1, dw1, dp = _dev(torch, cuda, u, v, w, u1, v1, w1, p)
    ch = Channel(case.nx, case.ny, case.nz, g.lx, g.ly, g.lz, g.dzf, case.nu, case.dt)
    s = PoissonSolver(case.nx, case.ny, case.nz, g.dx, g.dy, g.dzf)
    ppe = orc.OraclePPE(case.nx, case.ny, case.nz, g.dx, g.dy, g.dzf, pin_gauge=True)
    s.set_timing(True)
    worst = 0.0
    for n in range(nsteps):
        # per step: GPU step time, its PPE time and max |div| (the log of the
        # reference's driver, src/main.f90:213-222)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ch.step(s, du, dv, dw, du1, dv1, dw1, dp)
        e1.record()
        orc.step(G, ppe, u, v, w, u1, v1, w1, p)
        torch.cuda.synchronize()
        div_n, _ = ch.max_div(du, dv, dw)
        worst = max(worst, div_n)
        print(f"  step {n + 1:3d}: {e0.elapsed_time(e1):6.3f} ms (PPE {sum(s.stage_times()):6.3f} ms)"
              f"  max|div| {div_n:.2e}", flush=True)
    s.set_timing(False)
    return _host(du, dv, dw), (u, v, w), worst


def test_one_step_64(orc, cuda):
    from powerllel_channel_amd.workloads import CASES
    case = CASES["channel_64x64x64"]
    print(f"[one step] {case.name}", flush=True)
    got, ref, div = _run_steps(orc, cuda, case, 1)
    for a, b in zip(got, ref):
        rel = np.abs(a - b).max() / np.abs(b).max()
        print(f"  vel rel {rel:.2e}")
        assert rel <= 1e-12
    print(f"  max div {div:.2e}")
    assert div < 1e-12


def test_steps_full_size_vs_oracle(orc, cuda):
    """768x384x256 (stretch 4, nu = 1/5600, dt = 1e-4, the step workload of the
    benchmark): three steps through pll_channel_run, whose projection is
    carried into the next step's RK1 on a grid this size, against three
    oracle steps; velocity and pressure within 1e-12, max |div u| < 1e-12."""
    import torch
    from powerllel_channel_amd import Channel, PoissonSolver
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ, mean_removed, poiseuille_fields
    nx, ny, nz, nu, dt, nsteps = 768, 384, 256, 1 / 5600, 1e-4, 3
    print(f"[steps full size] {nx}x{ny}x{nz}, {nsteps} steps via run()", flush=True)
    g = Grid(nx, ny, nz, LX, LY, LZ, 4.0)
    G = orc.OracleGrid(nx, ny, nz, LX, LY, LZ, g.dzf, nu, dt)
    u, v, w = poiseuille_fields(g, 17)
    orc.bc_vel(G, u, v, w, 0.0)
    u1, v1, w1, p = [np.zeros(g.shape) for _ in range(4)]
    d = _dev(torch, cuda, u, v, w, u1, v1, w1, p)
    ch = Channel(nx, ny, nz, LX, LY, LZ, g.dzf, nu, dt)
    s = PoissonSolver(nx, ny, nz, g.dx, g.dy, g.dzf)
    ch.run(s, nsteps, *d)
    ppe = orc.OraclePPE(nx, ny, nz, g.dx, g.dy, g.dzf, pin_gauge=True)
    for _ in range(nsteps):
        orc.step(G, ppe, u, v, w, u1, v1, w1, p)
    torch.cuda.synchronize()
    div, _ = ch.max_div(d[0], d[1], d[2])
    for name, t, ref in zip("uvw", d[:3], (u, v, w)):
        rel = float(np.abs(t.cpu().numpy() - ref).max() / np.abs(ref).max())
        print(f"  {name} rel {rel:.2e}", flush=True)
        assert rel <= 1e-12, (name, rel)
    gp = d[6].cpu().numpy()
    ep = float(np.abs(mean_removed(g, gp) - mean_removed(g, p)).max()
               / np.abs(mean_removed(g, p)).max())
    print(f"  p rel (mean-removed) {ep:.2e}, max|div u| {div:.2e}")
    assert ep <= 1e-12 and div < 1e-12
    s.close()
    ch.close()


@pytest.mark.parametrize("nx,ny,nz", [(160, 70, 48), (110, 26, 16)])
def test_steps_radix57_vs_oracle(orc, cuda, nx, ny, nz, monkeypatch):
    """Grids whose transform lengths have the prime factors 5, 7, 11 and 13
    (x 80 = 2^4 5, y 70 = 2 5 7; x 55 = 5 11, y 26 = 2 13), stepped through
    pll_channel_run with the projection carried into the next step's RK1 (the
    fused right-hand side read inside the x transform), against the oracle.
    After one step the right-hand sides agree bit for bit (the stencils are
    bit-exact), so the pressure is held to the solve's 1e-12.  After two more,
    velocity is held to 1e-12 and pressure to 1e-11: the velocities then
    differ by roundoff (about 2e-15), which the pressure right-hand side
    divides by dx dt."""
    import torch
    from powerllel_channel_amd import Channel, PoissonSolver
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ, mean_removed, poiseuille_fields
    monkeypatch.setenv("PLL_RUN_CHAIN", "1")
    nu, dt = 1 / 5600, 1e-3
    print(f"[steps radix 5/7/11/13] {nx}x{ny}x{nz}", flush=True)
    g = Grid(nx, ny, nz, LX, LY, LZ, 3.0)
    G = orc.OracleGrid(nx, ny, nz, LX, LY, LZ, g.dzf, nu, dt)
    u, v, w = poiseuille_fields(g, 23)
    orc.bc_vel(G, u, v, w, 0.0)
    u1, v1, w1, p = [np.zeros(g.shape) for _ in range(4)]
    d = _dev(torch, cuda, u, v, w, u1, v1, w1, p)
    ch = Channel(nx, ny, nz, LX, LY, LZ, g.dzf, nu, dt)
    s = PoissonSolver(nx, ny, nz, g.dx, g.dy, g.dzf)
    ppe = orc.OraclePPE(nx, ny, nz, g.dx, g.dy, g.dzf, pin_gauge=True)
    for nsteps, ptol in ((1, 1e-12), (2, 1e-11)):
        ch.run(s, nsteps, *d)
        for _ in range(nsteps):
            orc.step(G, ppe, u, v, w, u1, v1, w1, p)
        torch.cuda.synchronize()
        div, _ = ch.max_div(d[0], d[1], d[2])
        for name, t, ref in zip("uvw", d[:3], (u, v, w)):
            rel = float(np.abs(t.cpu().numpy() - ref).max() / np.abs(ref).max())
            print(f"  +{nsteps} step(s): {name} rel {rel:.2e}", flush=True)
            assert rel <= 1e-12, (name, rel)
        gp = d[6].cpu().numpy()
        ep = float(np.abs(mean_removed(g, gp) - mean_removed(g, p)).max()
                   / np.abs(mean_removed(g, p)).max())
        print(f"  +{nsteps} step(s): p rel (mean-removed) {ep:.2e}, max|div u| {div:.2e}")
        assert ep <= ptol and div < 1e-12
    s.close()
    ch.close()


def _step_vs_parts(orc, cuda, nx, ny, nz, is_forced):
    """One fused step and the same step as separate calls, from one start."""
    import torch
    from powerllel_channel_amd import Channel, PoissonSolver
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ, poiseuille_fields
    g = Grid(nx, ny, nz, LX, LY, LZ, 3.0)
    G = orc.OracleGrid(nx, ny, nz, LX, LY, LZ, g.dzf, 1 / 5600, 1e-3)
    u, v, w = poiseuille_fields(g, 4)
    orc.bc_vel(G, u, v, w, 0.0)
    ch = Channel(nx, ny, nz, LX, LY, LZ, g.dzf, 1 / 5600, 1e-3, is_forced=is_forced,
                 vel_force=(1.0, 0.25, 0.0))
    s = PoissonSolver(nx, ny, nz, g.dx, g.dy, g.dzf)
    z = np.zeros(g.shape)
    a = _dev(torch, cuda, u, v, w, z, z, z, z)
    b = _dev(torch, cuda, u, v, w, z, z, z, z)
    ch.step(s, *a)
    du, dv, dw, du1, dv1, dw1, dp = b
    ch.rk1(du, dv, dw, du1, dv1, dw1)
    ch.bc_vel(du1, dv1, dw1)
    ch.rk2(du, dv, dw, du1, dv1, dw1)
    ch.bc_vel(du, dv, dw)
    ch.calc_rhs(du, dv, dw, dp)
    s.solve(dp)
    ch.bc_p(dp)
    ch.correct_vel(dp, du, dv, dw)
    ch.force_vel(du, dv, dw)
    ch.bc_vel(du, dv, dw)
    torch.cuda.synchronize()
    return ch, a, b


@pytest.mark.parametrize("nx,ny,nz", [(64, 48, 32), (768, 384, 32)])
def test_fused_rhs_step_bitwise(orc, cuda, nx, ny, nz):
    """Unforced: the step forms the PPE right-hand side inside the solve's x
    transform; it must equal calcRHS followed by the solve bit for bit."""
    import torch
    print(f"[fused rhs step] {nx}x{ny}x{nz}", flush=True)
    _, a, b = _step_vs_parts(orc, cuda, nx, ny, nz, (False, False, False))
    # u, v, w and p; the RK1 stage (u1, v1, w1) is work space whose boundary
    # fill (ghosts and w's wall row) the step leaves out
    for q in (0, 1, 2, 6):
        assert torch.equal(a[q], b[q]), q


@pytest.mark.parametrize("is_forced", [(True, False, False), (True, True, False),
                                       (False, False, True)])
def test_fused_forcing_step(orc, cuda, is_forced):
    """Forced u/v: Mean_h comes from the RK2 output and the shift is applied in
    correctVel; equal to the separate passes up to summation roundoff, and the
    forced bulk velocity is met.  A forced w takes the separate passes."""
    import torch
    nx, ny, nz = 64, 48, 32
    print(f"[fused forcing step] forced={is_forced}", flush=True)
    ch, a, b = _step_vs_parts(orc, cuda, nx, ny, nz, is_forced)
    for x, y in zip(a[:3], b[:3]):
        rel = float((x - y).abs().max() / y.abs().max())
        print(f"  vel rel {rel:.2e}")
        if is_forced[2]:
            assert torch.equal(x, y)
        else:
            assert rel <= 1e-14
    for d, target in zip(range(3), (1.0, 0.25, 0.0)):
        if is_forced[d]:
            m = ch.mean_h(a[d])
            assert abs(m - target) <= 1e-13, (d, m)


@pytest.mark.slow
def test_100_steps_192(orc, cuda):
    from powerllel_channel_amd.workloads import CASES
    case = CASES["channel_192x192x128"]
    print(f"[100 steps] {case.name}", flush=True)
    got, ref, div = _run_steps(orc, cuda, case, case.steps)
    for a, b in zip(got, ref):
        rel = np.abs(a - b).max() / np.abs(b).max()
        print(f"  vel rel {rel:.2e}")
        assert rel <= 1e-10
    print(f"  max div over the steps {div:.2e}")
    assert div < 1e-12


@pytest.mark.parametrize("shape", [(48, 32, 24), (40, 36, 21), (32, 96, 96), (40, 90, 100)])
def test_maxloc_row_orders(orc, cuda, shape):
    """calcMaxDiv and calcMaxCFL against the oracle, value and location, with the
    rows handed out as y slabs per XCD (ny a multiple of 8) and in plain order
    (ny = 36, 90), and with more than 8192 row partials (the final reduction's
    eight-at-a-time loop); a zero field makes every cell tie, so the location
    must be the lowest cell index whatever order the rows were reduced in."""
    import torch
    nx, ny, nz = shape
    print(f"[maxloc] {shape}", flush=True)
    g, G, ch, f = _setup(orc, nx, ny, nz, 3.0, seed=11)
    u, v, w = f[0], f[1], f[2]
    du, dv, dw = _dev(torch, cuda, u, v, w)
    assert ch.max_div(du, dv, dw) == orc.max_div(G, u, v, w)
    assert ch.max_cfl(du, dv, dw) == orc.max_cfl(G, u, v, w)
    z = np.zeros(g.shape)
    dz = _dev(torch, cuda, z, z, z)
    val, ijk = ch.max_div(*dz)
    assert (val, ijk) == orc.max_div(G, z, z, z)
    assert val == 0.0


def test_monitors_vs_oracle(orc, cuda):
    """calcMaxCFL / calcMaxDiv / checkNaN on the GPU against the oracle, and the
    run-time check of main.f90:213-222 (NaN planted in the interior and in a
    halo cell)."""
    import torch
    g, G, ch, f = _setup(orc, 48, 32, 24, 3.0)
    u, v, w = f[0], f[1], f[2]
    du, dv, dw = _dev(torch, cuda, u, v, w)
    val, ijk = ch.max_cfl(du, dv, dw)
    oval, oijk = orc.max_cfl(G, u, v, w)
    print(f"[monitors] cfl gpu {val!r} {ijk} oracle {oval!r} {oijk}")
    assert val == oval and ijk == oijk
    dval, dijk = ch.max_div(du, dv, dw)
    odval, odijk = orc.max_div(G, u, v, w)
    assert dval == odval and dijk == odijk
    assert not ch.check_nan(du)
    passed, rep = ch.check(du, dv, dw, cfl_limit=1e9, div_limit=1e9)
    assert passed and not rep["nan"]
    passed, rep = ch.check(du, dv, dw, cfl_limit=0.5 * val, div_limit=1e9)
    assert not passed
    for idx in [(5, 7, 9), (0, 0, 0), (g.nz + 1, g.ny + 1, g.nx + 1)]:
        bad = du.clone()
        bad[idx] = float("nan")
        assert ch.check_nan(bad)
        passed, rep = ch.check(bad, dv, dw, cfl_limit=1e9, div_limit=1e9)
        assert not passed and rep["nan"]


def test_monitor_files_vs_oracle(orc, cuda, tmp_path):
    """Skin friction per wall (calcAndWriteSkf), Mean_h of the forced
    components and the probe point on the GPU against the oracle, then the
    monitor files of two outputs against records formatted from the oracle's
    values.  The wall sums reduce in another order than the reference's
    running sum: 1e-13 relative; the probe is a copy: exact."""
    import torch
    from powerllel_channel_amd import Channel
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.monitor import Monitor, _record
    from powerllel_channel_amd.workloads import LX, LY, LZ, poiseuille_fields
    nx, ny, nz, u_crf, u_ref = 64, 48, 40, 0.3, 1.25
    print(f"[monitor files] {nx}x{ny}x{nz} u_crf={u_crf}", flush=True)
    g = Grid(nx, ny, nz, LX, LY, LZ, 3.0)
    G = orc.OracleGrid(nx, ny, nz, LX, LY, LZ, g.dzf, 1 / 5600, 1e-3)
    u, v, w = poiseuille_fields(g, 8)
    u -= u_crf
    orc.bc_vel(G, u, v, w, u_crf)
    p = np.random.default_rng(9).standard_normal(g.shape)
    ch = Channel(nx, ny, nz, LX, LY, LZ, g.dzf, 1 / 5600, 1e-3, u_crf=u_crf,
                 is_forced=(True, True, False), vel_force=(1.0, 0.0, 0.0))
    du, dv, dw, dp = _dev(torch, cuda, u, v, w, p)
    got = ch.skin_friction(du, u_ref)
    ref = orc.skin_friction(G, u, u_ref=u_ref, u_crf=u_crf)
    print(f"  skf gpu {got} oracle {ref}", flush=True)
    for a, b in zip(got, ref):
        for x, y in zip(a, b):
            assert abs(x - y) <= 1e-13 * abs(y)
    ijk = (17, 5, nz)
    assert ch.probe(du, dv, dw, dp, ijk) == orc.probe(u, v, w, p, ijk, u_crf)
    with pytest.raises(Exception):
        ch.probe(du, dv, dw, dp, (nx + 1, 1, 1))
    means = [orc.mean_h(G, u), orc.mean_h(G, v), 0.0]
    assert abs(ch.mean_h(du) - means[0]) <= 1e-14 * abs(means[0])
    with Monitor(ch, out_forcing=True, out_probe_point=True, probe_ijk=ijk,
                 out_skf_z=(True, True), nt_out_moni=10, u_ref=u_ref,
                 directory=str(tmp_path)) as mon:
        for nt in (10, 15, 20):
            mon.output(nt, du, dv, dw, dp)
    want = {
        "monitor_forcing.out": (means[0] + u_crf, means[1], means[2]),
        "monitor_probe.out": orc.probe(u, v, w, p, ijk, u_crf),
        "monitor_skf_z1.out": ref[0],
        "monitor_skf_z2.out": ref[1],
    }
    for name, vals in want.items():
        lines = open(tmp_path / name).read().splitlines()
        recs = [ln for ln in lines if ln[:9].strip().isdigit()]
        assert [r[:9] for r in recs] == ["       10", "       20"], name
        for r in recs:
            fields = [float(r[9 + 15 * q: 24 + 15 * q]) for q in range(len(vals))]
            expect = [float(x) for x in _record(10, vals)[9:].split()]
            for a, b in zip(fields, expect):  # 7 significant digits
                assert abs(a - b) <= 1.01e-7 * abs(b), (name, r)
    ch.close()


@pytest.mark.parametrize("shape,forced,u_crf,nsteps", [
    ((64, 48, 32), (True, False, False), 0.0, 5),
    ((64, 48, 32), (True, True, False), 0.3, 4),
    ((96, 54, 24), (False, False, False), 0.2, 3),
    ((96, 54, 24), (True, False, False), 0.0, 6),
])
def test_run_equals_steps(orc, cuda, shape, forced, u_crf, nsteps, monkeypatch):
    """pll_channel_run (projection formed in the next step's RK1; forced on
    these small grids with PLL_RUN_CHAIN=1) against the same number of
    pll_channel_step calls: bit for bit, every field, odd and even step
    counts, with and without forcing, with a convecting frame."""
    import torch
    monkeypatch.setenv("PLL_RUN_CHAIN", "1")
    from powerllel_channel_amd import Channel, PoissonSolver
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ, poiseuille_fields
    nx, ny, nz = shape
    print(f"[run vs steps] {shape} forced={forced} u_crf={u_crf} n={nsteps}", flush=True)
    g = Grid(nx, ny, nz, LX, LY, LZ, 3.0)
    ch = Channel(nx, ny, nz, LX, LY, LZ, g.dzf, 1 / 5600, 1e-3, u_crf=u_crf,
                 is_forced=forced, vel_force=(1.0 - u_crf, 0.25, 0.0))
    s = PoissonSolver(nx, ny, nz, g.dx, g.dy, g.dzf)
    u, v, w = poiseuille_fields(g, 21)
    u -= u_crf
    z = np.zeros(g.shape)
    a = _dev(torch, cuda, u, v, w, z, z, z, z)
    ch.bc_vel(*a[:3])
    b = [t.clone() for t in a]
    for _ in range(nsteps):
        ch.step(s, *a)
    ch.run(s, nsteps, *b)
    torch.cuda.synchronize()
    for name, x, y in zip("u v w".split(), a[:3], b[:3]):
        assert torch.equal(x, y), (name, float((x - y).abs().max()))
    assert torch.equal(a[6][1:-1, 1:-1, 1:-1], b[6][1:-1, 1:-1, 1:-1])
    s.close()
    ch.close()


@pytest.mark.parametrize("env", [{"PLL_RK_FULL": "0"}, {"PLL_RK2_FULL": "1"}])
@pytest.mark.parametrize("u_crf", [0.0, 0.37])
def test_rk_tile_forms_bitwise(orc, cuda, env, u_crf, monkeypatch):
    """The step kernels have a general form (bounds-checked tiles) and a
    full-tile form for grids whose x and y extents are multiples of the tile;
    both must give the oracle's stencils bit for bit: RK1 and RK2 on their
    own, then a forced run (Mean_h reduced inside RK2, projection inside
    RK1), against the default selection."""
    import torch
    from powerllel_channel_amd import Channel, PoissonSolver
    from powerllel_channel_amd.workloads import poiseuille_fields
    nx, ny, nz = 128, 48, 24
    print(f"[rk tile forms] {env} u_crf={u_crf}", flush=True)
    g, G, ch, (u, v, w, u1, v1, w1, _) = _setup(orc, nx, ny, nz, 3.0, u_crf=u_crf)
    for k, val in env.items():
        monkeypatch.setenv(k, val)
    du, dv, dw, du1, dv1, dw1 = _dev(torch, cuda, u, v, w, u1, v1, w1)
    ch.rk1(du, dv, dw, du1, dv1, dw1)
    orc.rk1(G, u, v, w, u1, v1, w1, u_crf)
    ch.rk2(du, dv, dw, du1, dv1, dw1)
    orc.rk2(G, u, v, w, u1, v1, w1, u_crf)
    torch.cuda.synchronize()
    for a, b in zip(_host(du1, dv1, dw1, du, dv, dw), (u1, v1, w1, u, v, w)):
        assert np.array_equal(a, b)
    ch.close()
    monkeypatch.setenv("PLL_RUN_CHAIN", "1")
    ch = Channel(nx, ny, nz, g.lx, g.ly, g.lz, g.dzf, 1 / 5600, 1e-3, u_crf=u_crf,
                 is_forced=(True, True, False), vel_force=(1.0 - u_crf, 0.25, 0.0))
    s = PoissonSolver(nx, ny, nz, g.dx, g.dy, g.dzf)
    u, v, w = poiseuille_fields(g, 24)
    u -= u_crf
    z = np.zeros(g.shape)
    a = _dev(torch, cuda, u, v, w, z, z, z, z)
    ch.bc_vel(*a[:3])
    b = [t.clone() for t in a]
    ch.run(s, 4, *a)
    for k in env:
        monkeypatch.delenv(k)
    ch.run(s, 4, *b)
    torch.cuda.synchronize()
    for name, x, y in zip("u v w".split(), a[:3], b[:3]):
        assert torch.equal(x, y), (name, float((x - y).abs().max()))
    s.close()
    ch.close()


@pytest.mark.parametrize("chain", ["1", "0"])
def test_run_calls_vs_steps(orc, cuda, chain, monkeypatch):
    """Four pll_channel_run calls in a row (5, 3, 7 and 5 steps), with
    per-step calls and a Mean_h read-back between them, each checked bit for
    bit against the same steps taken one call at a time; chain = "0": run()
    as a loop of steps (PLL_RUN_CHAIN=0, the small-grid default)."""
    import torch
    from powerllel_channel_amd import Channel, PoissonSolver
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ, poiseuille_fields
    monkeypatch.setenv("PLL_RUN_CHAIN", chain)
    nx, ny, nz = 64, 48, 32
    print(f"[run calls] chain={chain}", flush=True)
    g = Grid(nx, ny, nz, LX, LY, LZ, 3.0)
    ch = Channel(nx, ny, nz, LX, LY, LZ, g.dzf, 1 / 5600, 1e-3)
    s = PoissonSolver(nx, ny, nz, g.dx, g.dy, g.dzf)
    u, v, w = poiseuille_fields(g, 22)
    z = np.zeros(g.shape)
    a = _dev(torch, cuda, u, v, w, z, z, z, z)
    ch.bc_vel(*a[:3])
    b = [t.clone() for t in a]
    for n in (5, 3, 7, 5):
        for _ in range(n):
            ch.step(s, *a)
        ch.run(s, n, *b)
        torch.cuda.synchronize()
        for name, x, y in zip("u v w".split(), a[:3], b[:3]):
            assert torch.equal(x, y), (n, name, float((x - y).abs().max()))
        ch.step(s, *a)  # eager work between the replays
        ch.step(s, *b)
        assert ch.mean_h(a[0]) == ch.mean_h(b[0])
    s.close()
    ch.close()


@pytest.mark.parametrize("u_crf", [0.0, 0.3])
def test_checkpoint_restart_bitwise(orc, cuda, tmp_path, u_crf):
    """Checkpoint after 2 of 4 steps (u + u_crf, v, w to HDF5 files in the
    reference's layout), restart from the files (minus u_crf, then the
    velocity boundary fill, src/main.f90:86-101) and run the last 2 steps:
    the result equals the uninterrupted run bit for bit once the restart's
    two roundings, (u + u_crf) - u_crf, are applied to it as well."""
    import torch
    from powerllel_channel_amd import Channel, PoissonSolver
    from powerllel_channel_amd import checkpoint as ck
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ, poiseuille_fields
    if not os.path.exists(ck.IO_LIB_PATH):
        pytest.skip("libpllio.so not built")
    nx, ny, nz = 64, 48, 32
    print(f"[restart] {nx}x{ny}x{nz} u_crf={u_crf}", flush=True)
    g = Grid(nx, ny, nz, LX, LY, LZ, 3.0)
    ch = Channel(nx, ny, nz, LX, LY, LZ, g.dzf, 1 / 5600, 1e-3, u_crf=u_crf)
    s = PoissonSolver(nx, ny, nz, g.dx, g.dy, g.dzf)
    u, v, w = poiseuille_fields(g, 12)
    u -= u_crf
    a = _dev(torch, cuda, u, v, w, *(np.zeros(g.shape) for _ in range(4)))
    ch.bc_vel(*a[:3])
    for _ in range(2):
        ch.step(s, *a)
    out = ck.FieldOutput(nt_out_save=2, directory=str(tmp_path))
    out.output(2, a[0], a[1], a[2], a[6], u_crf)
    b = [torch.zeros_like(a[0]) for _ in range(7)]
    assert ck.input_data("save_00000002", b[0], b[1], b[2], u_crf, str(tmp_path)) == 2
    ch.bc_vel(*b[:3])
    a[0][1:-1, 1:-1, 1:-1] = (a[0][1:-1, 1:-1, 1:-1] + u_crf) - u_crf
    ch.bc_vel(*a[:3])
    for x, y in zip(a[:3], b[:3]):
        assert torch.equal(x, y)
    for _ in range(2):
        ch.step(s, *a)
        ch.step(s, *b)
    torch.cuda.synchronize()
    for x, y in zip(a[:3], b[:3]):
        assert torch.equal(x, y)
    s.close()
    ch.close()


def test_refined_steps_vs_oracle(orc, cuda):
    """RK2 steps with the PPE in the refined form, to a 1e-11 pressure
    tolerance, against the oracle's FP64 steps (the reference algorithm):
    the precision form's velocity, pressure and divergence errors."""
    import torch
    from powerllel_channel_amd import Channel, PoissonSolver
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ, mean_removed, poiseuille_fields
    nx, ny, nz, nu, dt, nsteps = 64, 48, 32, 1 / 5600, 1e-3, 20
    print(f"[refined steps] {nx}x{ny}x{nz}, {nsteps} steps, tolerance 1e-11", flush=True)
    g = Grid(nx, ny, nz, LX, LY, LZ, 4.0)
    G = orc.OracleGrid(nx, ny, nz, LX, LY, LZ, g.dzf, nu, dt)
    u, v, w = poiseuille_fields(g, 6)
    orc.bc_vel(G, u, v, w, 0.0)
    u1, v1, w1, p = [np.zeros(g.shape) for _ in range(4)]
    du, dv, dw, du1, dv1, dw1, dp = _dev(torch, cuda, u, v, w, u1, v1, w1, p)
    ch = Channel(nx, ny, nz, LX, LY, LZ, g.dzf, nu, dt)
    s = PoissonSolver(nx, ny, nz, g.dx, g.dy, g.dzf)
    s.set_refinement_tol(1e-11, 8)
    ppe = orc.OraclePPE(nx, ny, nz, g.dx, g.dy, g.dzf, pin_gauge=True)
    worst_div = 0.0
    for _ in range(nsteps):
        ch.step(s, du, dv, dw, du1, dv1, dw1, dp)
        orc.step(G, ppe, u, v, w, u1, v1, w1, p)
        torch.cuda.synchronize()
        worst_div = max(worst_div, ch.max_div(du, dv, dw)[0])
    sweeps, est = s.last_refinement()
    ev = max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(_host(du, dv, dw), (u, v, w)))
    gp = dp.cpu().numpy()
    ep = float(np.abs(mean_removed(g, gp) - mean_removed(g, p)).max() / np.abs(mean_removed(g, p)).max())
    print(f"  last solve {sweeps} sweeps (estimate {est:.1e}); velocity {ev:.2e}, "
          f"pressure {ep:.2e}, max|div u| {worst_div:.2e}")
    assert ev <= 1e-10
    assert ep <= 1e-9
    assert worst_div < 1e-9  # the reference's run-time limit, examples/template/param.in:33
