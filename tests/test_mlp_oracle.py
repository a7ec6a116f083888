"""CPU pins of the network oracle (oracle/mlp_oracle.py): the FP64 backward against
central finite differences of the FP64 loss, BF16 rounding against known values, the
parameter packing, and an Adam step against its closed form."""
import numpy as np

from oracle import mlp_oracle as M


def _problem(seed=0, n=37):
    rng = np.random.default_rng(seed)
    p = M.init_params(rng).astype(np.float64)
    p += 0.05 * rng.standard_normal(p.size)  # nonzero biases and last layer
    X = rng.standard_normal((n, 9)) * np.array([1, 2, 3, 0.5, 0.5, 0.5, 1, 1, 1]) + 0.3
    T = rng.standard_normal((n, 3)) * 0.2
    mean = X.mean(0)
    std = X.std(0)
    scale = np.array([0.1, 0.2, 0.3])
    return p, mean, std, scale, X, T


def test_fp64_gradient_matches_finite_differences():
    p, mean, std, scale, X, T = _problem()
    _, g = M.loss_grad(p, mean, std, scale, X, T, bf=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(M.NPARAMS, 60, replace=False),
                          np.arange(M.NPARAMS - 3, M.NPARAMS)])  # include b4
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, mean, std, scale, X, T, bf=False)[0] -
              M.loss_grad(pm, mean, std, scale, X, T, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max()), (i, fd, g[i])


def test_directional_derivative():
    """A plausible mistake (a transposed weight, a dropped GELU' factor) fails this:
    the gradient's projection on a random direction equals the directional derivative."""
    p, mean, std, scale, X, T = _problem(3)
    _, g = M.loss_grad(p, mean, std, scale, X, T, bf=False)
    d = np.random.default_rng(4).standard_normal(p.size)
    h = 1e-6
    fd = (M.loss_grad(p + h * d, mean, std, scale, X, T, bf=False)[0] -
          M.loss_grad(p - h * d, mean, std, scale, X, T, bf=False)[0]) / (2 * h)
    assert abs(fd - g @ d) <= 1e-6 * abs(fd)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.140625, 65504.0, 1e-3])
    r = M.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: nearest even keeps 1
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7
    assert r[3] == -3.140625
    assert abs(r[5] - 1e-3) <= 1e-3 * 2 ** -8


def test_pack_unpack_and_count():
    assert M.NPARAMS == 64 * 9 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 3 * 64 + 3
    p = np.arange(M.NPARAMS, dtype=np.float64)
    assert np.array_equal(M.pack(M.unpack(p)), p)
    assert M.unpack(p)["W1"][1, 0] == 9  # row-major W[out][in]


def test_adam_first_step():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    p1, m1, m2 = M.adam(p, np.zeros(2), np.zeros(2), g, 1, lr=0.01)
    # first bias-corrected step moves each parameter by lr * sign(g) (up to eps)
    assert np.allclose(p1, p - 0.01 * np.sign(g), atol=1e-9)


def test_bf16_gradient_close_to_fp64():
    """Rounding at the BF16 points changes the gradient by a few BF16 ulps, not more."""
    p, mean, std, scale, X, T = _problem(5, n=200)
    _, g64 = M.loss_grad(p, mean, std, scale, X, T, bf=False)
    _, gbf = M.loss_grad(p, mean, std, scale, X, T, bf=True)
    assert np.linalg.norm(gbf - g64) <= 0.05 * np.linalg.norm(g64)
