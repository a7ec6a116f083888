"""The graph form of the cell matvec (rbc3d_set_graph) on a batch with two SH grid
types (discs at nlat 13, spheres at nlat 17): recorded once at a geometry, then
replayed; every replay equals the eager matvec bit for bit, with the program's own GPU
work and synchronisations between replays, in both precision forms and after the
cells are set again (a new geometry records a new graph)."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mixed():
    import rbc3d_amd
    from rbc3d_amd import configs
    w = configs.make_workload("C5", seed=1, ncell=16, Nb=(32, 32, 32))
    assert sorted(set(int(n) for n in w.cells.nlat)) == [13, 17]
    g = rbc3d_amd.from_workload(w)
    yield w, g
    g.close()


def _eager(g, us, prec):
    import torch
    g.set_graph(False)
    g.set_precision(prec)
    out = []
    for u in us:
        b = torch.empty_like(u)
        g.cell_matmult_dev(u.data_ptr(), b.data_ptr())
        g.sync()
        out.append(b.clone())
    return out


def _replays(g, us, ref, label):
    import torch
    dev = us[0].device
    b = torch.empty_like(us[0])
    for rep in range(3):
        for k, u in enumerate(us):
            g.cell_matmult_dev(u.data_ptr(), b.data_ptr())
            g.sync()
            assert g.graph_ready(), (label, rep, k)
            assert torch.equal(b, ref[k]), (label, rep, k)
            # the program's own work between replays
            x = torch.arange(4096, dtype=torch.float64, device=dev)
            assert float((x * x).sum()) == float(sum(i * i for i in range(4096)))
            torch.cuda.synchronize()


def test_graph_replays_equal_eager_matvec(mixed):
    import torch
    w, g = mixed
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev).manual_seed(11)
    us = [torch.randn(g.cell_dof(), dtype=torch.float64, device=dev, generator=gen)
          for _ in range(3)]
    g.set_geometry_reuse(True)
    try:
        for prec in (g.PREC_FP64, g.PREC_SPLIT):
            ref = _eager(g, us, prec)
            g.set_graph(True)
            assert not g.graph_ready()
            _replays(g, us, ref, f"precision {prec}")
        # a new geometry (the same cells set again): the recorded graph no longer applies
        g.set_precision(g.PREC_FP64)
        g.set_cells(w.cells)
        g.set_cell_density(1, w.cells.f)
        g.set_cell_density(2, w.cells.g)
        assert not g.graph_ready()
        ref = _eager(g, us, g.PREC_FP64)
        g.set_graph(True)
        _replays(g, us, ref, "new geometry")
    finally:
        g.set_graph(False)
        g.set_precision(g.PREC_FP64)


def test_graph_off_without_geometry_reuse(mixed):
    """Without geometry reuse every matvec recomputes the geometry, so nothing is
    recorded: the matvec runs eagerly."""
    import torch
    w, g = mixed
    dev = torch.device("cuda:0")
    u = torch.randn(g.cell_dof(), dtype=torch.float64, device=dev,
                    generator=torch.Generator(device=dev).manual_seed(3))
    ref = _eager(g, [u], g.PREC_FP64)[0]
    g.set_geometry_reuse(False)
    g.set_graph(True)
    try:
        b = torch.empty_like(u)
        for _ in range(2):
            g.cell_matmult_dev(u.data_ptr(), b.data_ptr())
            g.sync()
            assert not g.graph_ready()
            assert torch.equal(b, ref)
    finally:
        g.set_graph(False)
        g.set_geometry_reuse(True)
