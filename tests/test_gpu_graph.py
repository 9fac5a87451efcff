"""The solves captured in a CUDA (HIP) graph through PyTorch: every replay,
with eager solves and synchronisations between the replays, gives the eager
result bit for bit.  Covers the persistent z solve, the FP32 tiles, and the
refined form's side stream for the lowest modes (forked from and joined to
the captured stream)."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,nz,sweeps", [(128, 96, 48, -1), (128, 96, 48, 1),
                                             (256, 128, 64, 2), (768, 384, 256, -1),
                                             (768, 384, 256, 1)])
def test_solve_replays_in_graph(cuda, nx, ny, nz, sweeps):
    import torch
    from powerllel_channel_amd import PoissonSolver
    from powerllel_channel_amd.mesh import Grid
    from powerllel_channel_amd.workloads import LX, LY, LZ, spectral_rhs
    print(f"[graph] {nx}x{ny}x{nz} sweeps={sweeps}", flush=True)
    g = Grid(nx, ny, nz, LX, LY, LZ, 3.0)
    inp = torch.from_numpy(spectral_rhs(g, 8)).to(cuda)
    s = PoissonSolver(nx, ny, nz, g.dx, g.dy, g.dzf)
    p = inp.clone()

    def run():
        p.copy_(inp)
        if sweeps < 0:
            s.solve(p)
        else:
            s.solve_refined(p, sweeps)

    run()
    torch.cuda.synchronize()
    ref = p.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        run()
    for i in range(3):
        p.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(p, ref), f"replay {i}"
        run()  # eager work between the replays
        torch.cuda.synchronize()
        assert torch.equal(p, ref), f"eager after replay {i}"
    del graph
    s.close()
