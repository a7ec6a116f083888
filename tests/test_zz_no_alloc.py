"""rbc3d_hashtable_build_noalloc (the entry point that allocates nothing) against
rbc3d_hashtable_build, bit for bit, through the C ABI: two calls in a row on one GPU,
the second on a smaller input, so the second reads scratch left by the first."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _points(rng, n, Nc, L):
    x = rng.uniform(-1.5 * L, 2.5 * L, size=(n, 3))
    # points exactly on bin boundaries and at the box edges
    x[:5] = np.array([[0.0, 0.0, 0.0], [L, L, L], [-L, 2 * L, 0.5 * L], [L / Nc[0], 0, 0],
                      [-L / Nc[0], -L / Nc[1], -L / Nc[2]]])
    return np.ascontiguousarray(x.reshape(-1, order="F"))


def test_noalloc_hashtable_matches_allocating_entry():
    import torch

    import rbc3d_amd as R
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    rng = np.random.default_rng(11)
    # (points, Nc, box): the second call is smaller in points and in bins
    for n, Nc, L in [(300_000, (7, 6, 5), 15.0), (2_000, (4, 3, 3), 8.0)]:
        xf = _points(rng, n, Nc, L)
        iLbNc = np.array([Nc[0] / L, Nc[1] / L, Nc[2] / L])
        hoc_ref, next_ref = R.hashtable_build(Nc, iLbNc, xf)
        nh = (Nc[0] + 2) * (Nc[1] + 2) * (Nc[2] + 2)
        d_x = torch.from_numpy(xf).to(dev)
        # outputs start as garbage the entry point must overwrite
        d_hoc = torch.full((nh,), 12345, dtype=torch.int32, device=dev)
        d_next = torch.full((n,), 12345, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        R.hashtable_build_noalloc(Nc, iLbNc, n, d_x.data_ptr(), n, d_hoc.data_ptr(),
                                  d_next.data_ptr(), stream)
        torch.cuda.synchronize()
        assert np.array_equal(d_hoc.cpu().numpy(), hoc_ref)
        assert np.array_equal(d_next.cpu().numpy(), next_ref)


def test_noalloc_hashtable_rejects_sizes_beyond_scratch():
    """Sizes above the scratch maximum are refused on the host before anything is queued."""
    import torch

    import rbc3d_amd as R
    dev = torch.device("cuda", 0)
    n = (1 << 20) + 1
    d_x = torch.zeros(3 * n, dtype=torch.float64, device=dev)
    d_next = torch.zeros(n, dtype=torch.int32, device=dev)
    d_hoc = torch.zeros(67 * 66 * 66, dtype=torch.int32, device=dev)
    with pytest.raises(R.Rbc3dError, match="scratch maximum") as e:
        R.hashtable_build_noalloc((4, 4, 4), (0.5, 0.5, 0.5), n, d_x.data_ptr(), n,
                                  d_hoc.data_ptr(), d_next.data_ptr(), 0)
    assert e.value.code == 1
    with pytest.raises(R.Rbc3dError, match="scratch maximum") as e:
        R.hashtable_build_noalloc((65, 64, 64), (8.0, 8.0, 8.0), 10, d_x.data_ptr(), 10,
                                  d_hoc.data_ptr(), d_next.data_ptr(), 0)
    assert e.value.code == 1
    torch.cuda.synchronize()
