"""The time loop of src/main.f90 through powerllel_channel_amd.driver on the GPU.

The reference's only validation is its demo run (examples/template/param.in
run by build_th2a.sh): a 128^3 channel, stretch ratio 3, re = 2800,
dt = 1e-3, Poiseuille start with noise 0.5, 1000 steps, checked every 100
steps against cfl_limit = 1.0 and div_limit = 1e-9; it passes when the run
ends successfully.  The same run here (the noise is seeded numpy, so the
flow is not the reference's sample), then a restart: checkpoint at step 500,
a second run from the files to step 1000, equal bit for bit to the first.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PARAM = """&PARA p_row = 1, p_col = 1 /
&MESH read_mesh = F, mesh_type = 1, stretch_func = 0, stretch_ratio = 3.0,
      nx = 128, ny = 128, nz = 128, lx = 12.566370614359172, ly = 6.283185307179586, lz = 2.0 /
&TIME nt_end = 1000, dt = 1.0E-3, nt_check = 100, cfl_limit = 1.0, div_limit = 1.0E-9 /
&RESTART is_restart = {restart}, fn_prefix_input_inst = 'save_00000500' /
&PHYSICAL re = 2800, u_ref = 1.0, l_ref = 1.0, initial_field = 'poi', u0 = 1.0,
      init_with_noise = T, noise_intensity = 0.5, smooth_wall_visc = F, u_crf = 0.0 /
&FORCE is_forced = T, F, F, vel_force = 1.0, 0.0, 0.0 /
&STATISTICS nt_init_stat = 100000, sample_interval = 10 /
&OUTPUT nt_out_scrn = 100, nt_out_inst = 100000, nt_out_stat = 100000,
      nt_out_save = 500, overwrite_save = F, auto_cleanup = F, num_retained = 0,
      nt_out_moni = 100 /
&MONITOR out_forcing = T, out_probe_point = T, probe_ijk = 64, 64, 64, out_skf_z = T, T /
"""


def test_template_run_and_restart(cuda, tmp_path):
    import torch
    from powerllel_channel_amd.checkpoint import IO_LIB_PATH
    from powerllel_channel_amd.driver import Simulation
    from powerllel_channel_amd.params import RunParams
    if not os.path.exists(IO_LIB_PATH):
        pytest.skip("libpllio.so not built")
    print("[driver] 128^3 template run, 1000 steps", flush=True)
    lines = []

    def out(s):
        lines.append(s)
        print(s, flush=True)

    sim = Simulation(RunParams.from_text(PARAM.format(restart="F")), str(tmp_path), out=out)
    res = sim.run()
    assert res["passed"] and res["nt"] == 1000
    assert res["div_max"] < 1e-9 and res["cfl_max"] < 1.0
    assert "PowerLLEL.NOTE: Simulation ends successfully!" in lines
    recs = [l for l in lines if l[:9].strip().isdigit()]
    assert [int(l[:9]) for l in recs] == list(range(100, 1001, 100))
    assert abs(sim.ch.mean_h(sim.u) - 1.0) < 1e-13  # forceVel: bulk velocity 1
    final = [t.clone() for t in (sim.u, sim.v, sim.w)]
    sim.close()
    for name in ("monitor_skf_z1.out", "monitor_skf_z2.out", "monitor_forcing.out"):
        rows = open(tmp_path / name).read().splitlines()
        assert len(rows) == 11 and rows[-1].startswith("     1000"), name
        vals = [float(x) for x in rows[-1].split()[1:]]
        assert all(np.isfinite(vals))
    assert float(open(tmp_path / "monitor_forcing.out").read().splitlines()[-1].split()[1]) == 1.0
    assert len(open(tmp_path / "monitor_probe.out").read().splitlines()) == 12
    for v in "uvw":
        assert (tmp_path / f"save_00000500_{v}.h5").exists()
        assert (tmp_path / f"save_00001000_{v}.h5").exists()

    print("[driver] restart from step 500", flush=True)
    rlines = []
    sim2 = Simulation(RunParams.from_text(PARAM.format(restart="T")), str(tmp_path),
                      out=rlines.append)
    assert sim2.nt_start == 501
    res2 = sim2.run()
    assert res2["passed"] and res2["nt"] == 1000
    for a, b in zip(final, (sim2.u, sim2.v, sim2.w)):
        assert torch.equal(a, b)
    sim2.close()
    # the restart appends to the monitor files (header again, then 501..1000)
    rows = open(tmp_path / "monitor_skf_z1.out").read().splitlines()
    assert len(rows) == 11 + 1 + 5
