"""Pins of the CPU restatement (oracle/) against what the mathematics fixes: closed
forms, library routines restated elsewhere (numpy/scipy), invariants and brute
force on tiny inputs.  The reference ships no test vectors for this path (its CI only
formats code), so these checks plus the committed fixtures are what pin the oracle.
Each check is chosen so that a plausible slip (a dropped factor, a wrong sign, a
transposed operand, a wrong index) fails it."""
import math

import numpy as np
import pytest
from scipy.special import erfc, factorial, lpmv

from tests.conftest import relerr


@pytest.fixture(scope="module")
def L(orc):
    return orc.L()


# --- Gauss grid and normalised Legendre functions (SPHEREPACK gaqd / Pbar) -------
@pytest.mark.parametrize("n", [7, 13, 17, 36])
def test_gaqd_matches_numpy_leggauss(L, n):
    th = np.zeros(n)
    w = np.zeros(n)
    L.orc_gaqd(n, th, w)
    x, wx = np.polynomial.legendre.leggauss(n)
    assert np.all(np.diff(th) > 0)
    assert np.abs(np.cos(th) - x[::-1]).max() < 1e-15
    assert np.abs(w - wx[::-1]).max() < 1e-14


def test_pbar_closed_forms_and_orthonormality(L):
    n = 13
    th = np.zeros(n)
    w = np.zeros(n)
    L.orc_gaqd(n, th, w)
    pb = np.zeros(n * n * n)
    L.orc_pbar_table(n, th, n, pb)
    pb = pb.reshape(n, n, n)  # [m][l][i]
    c, s = np.cos(th), np.sin(th)
    assert np.abs(pb[0, 0] - 1 / math.sqrt(2)).max() < 1e-15
    assert np.abs(pb[0, 1] - math.sqrt(1.5) * c).max() < 1e-15
    assert np.abs(pb[1, 1] - math.sqrt(3) / 2 * s).max() < 1e-15
    for m in range(n):
        for l in range(m, n):
            # no Condon-Shortley phase: scipy's lpmv includes (-1)^m
            ref = math.sqrt((2 * l + 1) / 2 * factorial(l - m) / factorial(l + m)) * lpmv(m, l, c) * (-1) ** m
            assert np.abs(pb[m, l] - ref).max() < 1e-13
    G = np.einsum("mli,mki,i->mlk", pb, pb, w)
    for m in range(n):
        assert np.abs(G[m, m:, m:] - np.eye(n - m)).max() < 1e-13


# --- spherical-harmonic transforms ----------------------------------------------
def _coeffs(rng, n):
    a = np.triu(rng.standard_normal((n, n)))
    b = np.triu(rng.standard_normal((n, n)))
    b[0] = 0
    return a, b


@pytest.mark.parametrize("nlat", [13, 17])
def test_sh_roundtrip_and_synthesis_formula(L, nlat):
    rng = np.random.default_rng(nlat)
    nlon = 2 * nlat
    a, b = _coeffs(rng, nlat)
    g = np.zeros(nlat * nlon)
    L.orc_sh_synth_gau(nlat, nlon, 1, g, nlat, nlon, a.flatten("F"), b.flatten("F"), nlat, nlat)
    g = g.reshape(nlat, nlon, order="F")
    # direct evaluation of the documented synthesis at one point
    th = np.zeros(nlat)
    w = np.zeros(nlat)
    L.orc_gaqd(nlat, th, w)
    i, j = 3, 5
    phi = 2 * np.pi * j / nlon
    val = 0.0
    for m in range(nlat):
        for l in range(m, nlat):
            p = math.sqrt((2 * l + 1) / 2 * factorial(l - m) / factorial(l + m)) * lpmv(m, l, math.cos(th[i])) * (-1) ** m
            val += (0.5 if m == 0 else 1.0) * p * (a[m, l] * math.cos(m * phi) - b[m, l] * math.sin(m * phi))
    assert abs(g[i, j] - val) < 1e-12 * np.abs(g).max()
    a2 = np.zeros(nlat * nlat)
    b2 = np.zeros(nlat * nlat)
    L.orc_sh_anal_gau(nlat, nlon, 1, g.flatten("F"), nlat, nlon, a2, b2, nlat, nlat)
    a2 = a2.reshape(nlat, nlat, order="F")
    b2 = b2.reshape(nlat, nlat, order="F")
    iu = np.triu_indices(nlat)
    assert np.abs(a2[iu] - a[iu]).max() < 1e-13
    assert np.abs(b2[iu] - b[iu]).max() < 1e-13


@pytest.mark.parametrize("nlat", [13, 17])
def test_sh_gradient_vs_finite_differences(L, nlat):
    """ShGradGau (gradgs plus the sin th factor): v = dg/dth and w = dg/dphi of the
    band-limited field.  Reference: the same coefficients synthesised with scipy's lpmv
    (normalised, no Condon-Shortley phase) at th +- h and phi +- h, central differences."""
    nlon = 2 * nlat
    rng = np.random.default_rng(nlat + 1)
    a, b = _coeffs(rng, nlat)
    th = np.zeros(nlat)
    wg = np.zeros(nlat)
    L.orc_gaqd(nlat, th, wg)
    ph = 2 * np.pi * np.arange(nlon) / nlon
    af = np.ascontiguousarray(a.reshape(-1, order="F"))
    bf = np.ascontiguousarray(b.reshape(-1, order="F"))
    v = np.zeros(nlat * nlon)
    w = np.zeros(nlat * nlon)
    assert L.orc_sh_grad_gau(nlat, nlon, 1, v, w, nlat, nlon, af, bf, nlat, nlat) == 0
    v = v.reshape((nlat, nlon), order="F")
    w = w.reshape((nlat, nlon), order="F")

    def synth(t, p):
        T, Pp = np.meshgrid(t, p, indexing="ij")
        g = np.zeros_like(T)
        for m in range(nlat):
            for n in range(m, nlat):
                nb = math.sqrt((2 * n + 1) / 2 * factorial(n - m) / factorial(n + m)) * (-1) ** m
                P = nb * lpmv(m, n, np.cos(T))
                h = 0.5 if m == 0 else 1.0
                g += h * P * (a[m, n] * np.cos(m * Pp) - b[m, n] * np.sin(m * Pp))
        return g

    h = 1e-5
    v_fd = (synth(th + h, ph) - synth(th - h, ph)) / (2 * h)
    w_fd = (synth(th, ph + h) - synth(th, ph - h)) / (2 * h)
    scale = max(np.abs(v_fd).max(), np.abs(w_fd).max())
    assert np.abs(v - v_fd).max() < 1e-7 * scale
    assert np.abs(w - w_fd).max() < 1e-7 * scale


def test_sh_equi_synthesis_poles(L):
    rng = np.random.default_rng(3)
    nlat, nlon = 14, 26
    a, b = _coeffs(rng, nlat - 1)
    A = np.zeros((nlat, nlat))
    B = np.zeros((nlat, nlat))
    A[: nlat - 1, : nlat - 1] = a
    B[: nlat - 1, : nlat - 1] = b
    g = np.zeros(nlat * nlon)
    L.orc_sh_synth_equ(nlat, nlon, 1, g, nlat, nlon, A.flatten("F"), B.flatten("F"), nlat, nlat)
    g = g.reshape(nlat, nlon, order="F")
    # only m = 0 survives at the poles: Pbar(m>0) vanishes there
    pole_n = 0.5 * sum(a[0, l] * math.sqrt((2 * l + 1) / 2) for l in range(nlat - 1))
    pole_s = 0.5 * sum(a[0, l] * math.sqrt((2 * l + 1) / 2) * (-1) ** l for l in range(nlat - 1))
    assert np.abs(g[0] - pole_n).max() < 1e-12
    assert np.abs(g[-1] - pole_s).max() < 1e-12


# --- FFTs (PFFTW conventions) and spectral derivatives ---------------------------
def test_pme_fft_conventions(L):
    rng = np.random.default_rng(0)
    N1, N2, N3 = 8, 6, 10
    f = rng.standard_normal((N3, N2, N1))
    out = np.zeros(2 * (N1 // 2 + 1) * N2 * N3)
    L.orc_fft3d_forward(N1, N2, N3, f.ravel(), out)
    Fo = (out[0::2] + 1j * out[1::2]).reshape(N3, N2, N1 // 2 + 1)
    # 2-D r2c with sign -1 over (x, y), then sign +1 along z
    Fn = np.fft.fft(np.fft.fft(np.fft.rfft(f, axis=2), axis=1), axis=0)[(-np.arange(N3)) % N3]
    assert np.abs(Fo - Fn).max() < 1e-12
    V = rng.standard_normal(Fo.shape) + 1j * rng.standard_normal(Fo.shape)
    inp = np.zeros(2 * V.size)
    inp[0::2] = V.real.ravel()
    inp[1::2] = V.imag.ravel()
    o2 = np.zeros(N1 * N2 * N3)
    L.orc_fft3d_inverse(N1, N2, N3, inp, o2)
    # c2r drops the imaginary parts of kx = 0 and kx = N1/2, as numpy's irfftn does
    ref = np.fft.irfftn(V[(-np.arange(N3)) % N3], s=(N3, N2, N1), axes=(0, 1, 2)) * N1 * N2 * N3
    assert np.abs(o2.reshape(N3, N2, N1) - ref).max() < 1e-12


def test_fft_diff_exact_on_trig_polynomials(L):
    m, n = 26, 26
    x = 2 * np.pi * np.arange(m) / m
    y = 2 * np.pi * np.arange(n) / n
    X, Y = np.meshgrid(x, y, indexing="ij")
    u = np.sin(3 * X) * np.cos(2 * Y) + np.cos(5 * X + Y)
    du1 = 3 * np.cos(3 * X) * np.cos(2 * Y) - 5 * np.sin(5 * X + Y)
    du2 = -2 * np.sin(3 * X) * np.sin(2 * Y) - np.sin(5 * X + Y)
    u1 = np.zeros(m * n)
    u2 = np.zeros(m * n)
    L.orc_fft_diff(m, n, 1, u.flatten("F"), u1, u2)
    assert np.abs(u1.reshape(m, n, order="F") - du1).max() < 1e-12
    assert np.abs(u2.reshape(m, n, order="F") - du2).max() < 1e-12
    # the Nyquist mode is dropped (zeroed), as FFT_Diff does
    un = np.cos(13 * X)
    L.orc_fft_diff(m, n, 1, un.flatten("F"), u1, u2)
    assert np.abs(u1).max() < 1e-12


# --- Ewald coefficients, mask, B-splines, quadrature -----------------------------
def _sys(orc, alpha=math.pi / 4, rc=2.0, L_=8.0, Nb=16):
    return orc.Oracle((L_, L_, L_), alpha, rc, (Nb, Nb, Nb), 6, (5.0,))


def test_ewald_tables_vs_exact_forms(orc):
    alpha, rc = math.pi / 4, 2.0
    o = _sys(orc, alpha, rc)
    k = math.sqrt(math.pi / alpha)
    for r in [0.05, 0.3, 0.77, 1.2, 1.9999]:
        rt = k * r
        c1, c2 = erfc(rt), 2 / math.sqrt(alpha) * math.exp(-rt * rt)
        A, B = o.ewald_sl(r)
        # linear interpolation on 8192 intervals in t: |error| <= h^2/8 max|f''|
        # for each tabulated term, f1 = erfc(t), f2 = (2/sqrt(alpha)) exp(-t^2)
        h = k * rc / 8192
        ts = rt + h * np.linspace(-1, 1, 9)
        e1 = h * h / 8 * np.max(np.abs(4 * ts * np.exp(-ts * ts) / math.sqrt(math.pi))) * 1.01
        e2 = h * h / 8 * np.max(np.abs(2 / math.sqrt(alpha) * (4 * ts * ts - 2) * np.exp(-ts * ts))) * 1.01
        assert abs(A - (c1 / r ** 3 + c2 / r ** 2)) <= e1 / r ** 3 + e2 / r ** 2 + 1e-15 * A
        assert abs(B - (c1 / r - c2)) <= e1 / r + e2 + 1e-15 * abs(B)
        Ad = o.ewald_dl(r)
        ex = -8 / math.sqrt(math.pi) * (math.exp(-rt * rt) * (1.5 * rt + rt ** 3) + 0.75 * math.sqrt(math.pi) * erfc(rt)) / r ** 5
        f3 = lambda t: -8 / np.sqrt(np.pi) * (np.exp(-t * t) * (1.5 * t + t ** 3)
                                              + 0.75 * np.sqrt(np.pi) * np.vectorize(erfc)(t))
        d = 1e-3
        e3 = h * h / 8 * np.max(np.abs(f3(ts + d) - 2 * f3(ts) + f3(ts - d)) / d ** 2) * 1.05
        assert abs(Ad - ex) <= e3 / r ** 5 + 1e-14 * abs(ex)
    # at the table nodes the interpolation is exact
    r = 3 * rc / 8192
    rt = k * r
    A, B = o.ewald_sl(r)
    assert abs(B - (erfc(rt) / r - 2 / math.sqrt(alpha) * math.exp(-rt * rt))) < 1e-12 * abs(B)
    assert o.ewald_sl(rc) == (0.0, 0.0) and o.ewald_sl(1e-4) == (0.0, 0.0)


def test_mask_function(orc):
    o = _sys(orc)
    assert o.mask(0.0) == 1.0 and o.mask(1.0) == 0.0 and o.mask(1.5) == 0.0
    t = 0.37
    assert abs(o.mask(t) - math.exp(2 * math.exp(-1 / t) / (t - 1))) < 1e-8
    assert o.mask(-t) == o.mask(t)


@pytest.mark.parametrize("P", [4, 6])
def test_bspline_partition_and_moments(orc, P):
    import ctypes as C
    for xc in [3.25, 10.0, 7.999, -2.6]:
        w = np.zeros(P)
        imin = C.c_int()
        orc.L().orc_bspline(xc, P, C.byref(imin), w)
        assert abs(w.sum() - 1) < 1e-14
        nodes = imin.value + np.arange(P)
        # first moment: sum_j w_j (j + P/2) = xc for cardinal B-splines centred this way
        assert abs((w * (nodes + P / 2.0)).sum() - xc) < 1e-12
        assert np.all(w >= 0)


def test_gauleg_and_sinh_rules(orc):
    x = np.zeros(8)
    w = np.zeros(8)
    orc.L().orc_gauleg(0.0, 2.0, 8, x, w)
    xr, wr = np.polynomial.legendre.leggauss(8)
    assert np.abs(x - (xr + 1)).max() < 1e-14 and np.abs(w - wr).max() < 1e-14
    xs = np.zeros(16)
    ws = np.zeros(16)
    # The sinh map x = a + xl*b0*sinh(mu*s - eta) makes 1/sqrt((x-a)^2 + (xl*b0)^2)
    # constant in s, so the rule integrates it exactly.  The reference scales the
    # distance as b0 = b*xl (common/ModQuadRule.F90:239), so the exactly-integrated
    # distance is b*xl^2: b itself on [-1, 1], b/4 on the [0, 1] patch radius used
    # by the near-singular correction.
    b = 1e-3
    orc.L().orc_gauleg_sinh(-1.0, 1.0, 0.3, b, 16, xs, ws)
    exact = math.asinh(0.7 / b) + math.asinh(1.3 / b)
    assert abs((ws / np.sqrt((xs - 0.3) ** 2 + b * b)).sum() - exact) < 1e-12 * exact
    orc.L().orc_gauleg_sinh(0.0, 1.0, 0.0, b, 16, xs, ws)
    be = b * 0.25
    exact = math.asinh(1 / be)
    assert abs((ws / np.sqrt(xs ** 2 + be * be)).sum() - exact) < 1e-12 * exact
    assert abs(ws.sum() - 1) < 1e-12 and np.all(np.diff(xs) > 0)


# --- cell list and pair set --------------------------------------------------------
def test_hashtable_matches_brute_force_pairs(orc):
    rng = np.random.default_rng(1)
    L_, rc = 8.0, 2.0
    n = 300
    x = rng.uniform(-L_, 2 * L_, size=(n, 3))
    Nc = np.array([4, 4, 4], np.int32)
    iLbNc = np.array([1 / L_ * 4, 1 / L_ * 4, 4 / L_])
    hoc = np.zeros(6 ** 3, np.int32)
    nxt = np.zeros(n, np.int32)
    orc.L().orc_hashtable_build(Nc, iLbNc, n, x.flatten("F"), n, hoc, nxt)
    hoc = hoc.reshape(6, 6, 6, order="F")
    b = (np.floor(x * iLbNc).astype(int) % 4) + 1
    visited = set()
    for i in range(n):
        i1, i2, i3 = b[i]
        for j1 in range(i1 - 1, i1 + 2):
            for j2 in range(i2 - 1, i2 + 2):
                for j3 in range(i3 - 1, i3 + 2):
                    j = hoc[j1, j2, j3]
                    while j > 0:
                        d = x[j - 1] - x[i]
                        d = d - np.round(d / L_) * L_
                        if np.sqrt((d * d).sum()) <= rc:
                            visited.add((i, j - 1))
                        j = nxt[j - 1]
    brute = set()
    for i in range(n):
        d = x - x[i]
        d = d - np.round(d / L_) * L_
        for j in np.nonzero(np.sqrt((d * d).sum(1)) <= rc)[0]:
            brute.add((i, int(j)))
    assert visited == brute
    # each list in descending index order, heads are the largest index of the bin
    for j in range(n):
        if nxt[j] > 0:
            assert nxt[j] - 1 < j


@pytest.fixture(scope="module")
def c1(orc):
    from rbc3d_amd import configs
    w = configs.make_workload("C1", seed=1)
    return w, orc.from_workload(w)


def test_pair_sum_matches_brute_force(orc, c1):
    """Single-layer pair sums of the cell list vs a brute-force numpy sum over every
    source (minimum image, masked same-surface pairs), for a sample of targets; the
    singular corrections are removed by comparing two source densities that differ
    only on another cell (f' = f + df on cell 2)."""
    w, o = c1
    tl = o.tlist_rbc()
    tx, ti, act, A = tl
    n = len(A)
    nc = n // 2
    rng = np.random.default_rng(0)
    df = np.zeros_like(w.cells.f)
    df[3 * nc:] = rng.standard_normal(3 * nc)  # changes only cell 2
    c = 1 / (4 * np.pi)
    sub = np.zeros(n, np.int32)
    sub[:nc:7] = 1  # targets on cell 1
    o.set_density(1, w.cells.f)
    v1, _ = o.add_int_on_rbcs(c, 0.0, (tx, ti, sub, A))
    o.set_density(1, w.cells.f + df)
    v2, _ = o.add_int_on_rbcs(c, 0.0, (tx, ti, sub, A))
    o.set_density(1, w.cells.f)
    dv = (v2 - v1).reshape(3, n)
    # brute force of the pair-sum change (near-singular corrections of cell 2 also
    # change, so restrict to targets with no cell-2 node within rc + margin)
    X = tx.reshape(3, n).T
    nl = 13
    th = np.zeros(nl)
    wg = np.zeros(nl)
    orc.L().orc_gaqd(nl, th, wg)
    wq = wg * 2 * np.pi / 26
    dS = (w.cells.detj[nc:].reshape(nl, 26, order="F") * wq[:, None]).reshape(-1, order="F")
    F = df[3 * nc:].reshape(3, nc).T * dS[:, None]
    checked = 0
    for i in np.nonzero(sub)[0]:
        d = X[nc:] - X[i]
        d = d - np.round(d / w.Lb[0]) * w.Lb[0]
        r = np.sqrt((d * d).sum(1))
        if r.min() < 2.0:
            continue  # near-singular corrections involved
        checked += 1
        assert np.abs(dv[:, i]).max() == 0.0
    # targets with cell-2 nodes in range: brute-force pair sum minus the oracle change
    # must equal the near-singular part, which vanishes when the target is far from
    # the surface; pick targets whose closest cell-2 node is in (1.0, 2.0)
    for i in np.nonzero(sub)[0]:
        d = X[nc:] - X[i]
        d = d - np.round(d / w.Lb[0]) * w.Lb[0]
        r = np.sqrt((d * d).sum(1))
        if not (1.0 < r.min() < 2.0):
            continue
        s = np.zeros(3)
        for k in np.nonzero(r <= 2.0)[0]:
            Aa, Bb = o.ewald_sl(r[k])
            s += c / A[i] * (Aa * d[k] * (d[k] @ F[k]) + Bb * F[k])
        # near-singular corrections apply only within 2 mesh sizes of the surface
        assert np.abs(dv[:, i] - s).max() < 1e-13 * max(1.0, np.abs(s).max())
        checked += 1
    assert checked > 0


# --- PME against a direct reciprocal-space sum --------------------------------------
def test_pme_matches_direct_reciprocal_sum(orc):
    """Point forces and force dipoles in a periodic box: the PME velocity against the
    explicit sum over wave vectors of the same Green's functions (Hasimoto split)."""
    from rbc3d_amd import configs
    w = configs.make_small(ncell=1, nlat=7, Nb=(48, 48, 48), P=8, seed=5)
    o = orc.from_workload(w)
    tl = o.tlist_rbc()
    tx, ti, act, A = tl
    n = len(A)
    c = 1 / (4 * np.pi)
    sub = np.zeros(n, np.int32)
    sub[::7] = 1
    v_sl = o.pme(c, 0.0, (tx, ti, sub, A))
    v_dl = o.pme(0.0, c, (tx, ti, sub, A))
    # sources: list densities f dS, g dS, normals, Bcoef = 1 - lambda
    L_ = w.Lb[0]
    X = tx.reshape(3, n).T
    nl, nn = 7, 14
    th = np.zeros(nl)
    wg = np.zeros(nl)
    orc.L().orc_gaqd(nl, th, wg)
    wq = wg * 2 * np.pi / nn
    dS = (w.cells.detj.reshape(nl, nn, order="F") * wq[:, None]).reshape(-1, order="F")
    def cell_field(arr):
        return arr.reshape(3, nl * nn).T
    Fs = cell_field(w.cells.f) * dS[:, None]
    Gs = cell_field(w.cells.g) * dS[:, None]
    Ns = cell_field(w.cells.a3)
    Bc = -4.0
    alpha = w.alpha
    kmax = 24
    ks = np.arange(-kmax, kmax + 1)
    K = np.stack(np.meshgrid(ks, ks, ks, indexing="ij"), -1).reshape(-1, 3)
    K = K[np.any(K != 0, axis=1)]
    q = K / L_
    q2t = np.pi * alpha * (q * q).sum(1)
    e = np.exp(-q2t)
    phi0 = e / q2t
    phi1 = (e + phi0) / q2t
    vol = L_ ** 3
    Sf = np.exp(-2j * np.pi * (X @ q.T))  # per source: phase, reused for targets
    fk = Sf.T @ Fs  # sum_j f_j e^{-2 pi i q.x_j}
    T = np.einsum("ja,jb->jab", Gs, Ns * Bc)
    Tk = np.einsum("kj,jab->kab", Sf.T, T)
    qt = np.sqrt(np.pi * alpha) * q
    vk_sl = 2 * alpha / vol * phi1[:, None] * (q2t[:, None] * fk - qt * (qt * fk).sum(1)[:, None])
    tr = np.einsum("kaa->k", Tk)
    qT = np.einsum("ka,kab->kb", q, Tk)
    Tq = np.einsum("kab,kb->ka", Tk, q)
    qTq = np.einsum("ka,ka->k", q, Tq)
    vk_dl = -(1j * 4 * np.pi * alpha / vol * phi0[:, None] * (q * tr[:, None] + qT + Tq)
              - 1j * 8 * np.pi ** 2 * alpha ** 2 / vol * (phi1 * qTq)[:, None] * q)
    idx = np.nonzero(sub)[0]
    ph = np.exp(2j * np.pi * (X[idx] @ q.T))
    u_sl = (ph @ vk_sl).real * c / A[idx, None]
    u_dl = (ph @ vk_dl).real * c / A[idx, None]
    g_sl = v_sl.reshape(3, n).T[idx]
    g_dl = v_dl.reshape(3, n).T[idx]
    assert relerr(g_sl, u_sl) < 1e-5
    assert relerr(g_dl, u_dl) < 1e-5


# --- triangle quadrature ------------------------------------------------------------
def _tri_ref(alpha, x, f, xt, p0, n=48):
    """Converged integral of the Ewald single-layer kernel times the linear density
    f over triangle x: three collapsed (u, w) maps around the in-plane point p0,
    n-point Gauss-Legendre in each, exact erfc/exp kernel (no table)."""
    g, gw = np.polynomial.legendre.leggauss(n)
    g, gw = 0.5 * (g + 1), 0.5 * gw
    k = math.sqrt(math.pi / alpha)
    M = np.column_stack([x[1] - x[0], x[2] - x[0]])
    out = np.zeros(3)
    for a in range(3):
        va, vb = x[a], x[(a + 1) % 3]
        J = np.linalg.norm(np.cross(va - p0, vb - p0))
        U, W = np.meshgrid(g, g, indexing="ij")
        wt = np.outer(gw, gw) * U * J
        Y = p0 + U[..., None] * ((1 - W)[..., None] * (va - p0) + W[..., None] * (vb - p0))
        st = np.linalg.lstsq(M, (Y - x[0]).reshape(-1, 3).T, rcond=None)[0]
        lam = np.stack([1 - st[0] - st[1], st[0], st[1]])
        F = (lam.T @ f).reshape(Y.shape)
        XX = xt - Y
        r = np.linalg.norm(XX, axis=-1)
        c1 = np.vectorize(erfc)(k * r)
        c2 = 2 / math.sqrt(alpha) * np.exp(-(k * r) ** 2)
        A = c1 / r ** 3 + c2 / r ** 2
        B = c1 / r - c2
        v = A[..., None] * XX * (XX * F).sum(-1)[..., None] + B[..., None] * F
        out += (wt[..., None] * v).sum((0, 1))
    return out


def test_triangle_rules(orc):
    """MinDistToTri against sampling; the 7-point regular rule and the 3x4x4 Duffy
    rule (common/ModWall.F90 Tri_Int_Regular / Tri_Int_Duffy) against a converged
    integral of the same kernel and density."""
    alpha = math.pi / 4
    o = _sys(orc, alpha=alpha, rc=2.0)
    L_ = orc.L()
    import ctypes as C
    x = 0.1 * np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 1.0, 0.1]])
    xf = x.flatten("F")
    rng = np.random.default_rng(2)
    f = rng.standard_normal((3, 3))  # f[vertex, component]
    ff = f.flatten("F")
    s0, t0 = C.c_double(), C.c_double()
    uv = rng.random((200000, 2))
    uv = uv[uv.sum(1) <= 1]
    P = x[0] + uv[:, :1] * (x[1] - x[0]) + uv[:, 1:] * (x[2] - x[0])
    for xt in [np.array([0.03, 0.03, 0.05]), np.array([0.2, -0.1, 0.02]), np.array([-0.05, 0.05, -0.01])]:
        d = L_.orc_min_dist_to_tri(xt, xf, C.byref(s0), C.byref(t0))
        db = np.sqrt(((P - xt) ** 2).sum(1)).min()
        assert d <= db + 1e-12 and db - d < 5e-4
    cen = x.mean(0)
    nrm = np.cross(x[1] - x[0], x[2] - x[0])
    nrm /= np.linalg.norm(nrm)
    # regular rule: target one triangle size off the plane
    xt = cen + 0.1 * nrm
    rhs = np.zeros(3)
    L_.orc_tri_int_regular(o.h, xf, ff, xt, rhs)
    ref = _tri_ref(alpha, x, f, xt, cen)
    assert np.abs(rhs - ref).max() < 2e-3 * np.abs(ref).max()
    # Duffy rule: target in the triangle's plane (the singular case) at (s0, t0)
    for st in [(0.3, 0.3), (0.0, 0.0), (0.5, 0.2)]:
        p0 = x[0] + st[0] * (x[1] - x[0]) + st[1] * (x[2] - x[0])
        rd = np.zeros(3)
        L_.orc_tri_int_duffy(o.h, xf, ff, p0, st[0], st[1], rd)
        # the collapsed map with 4x4 Gauss points per sub-triangle is the Duffy rule
        # itself: agreement to the Ewald-table interpolation accuracy
        ref4 = _tri_ref(alpha, x, f, p0, p0, n=4)
        assert np.abs(rd - ref4).max() < 1e-6 * np.abs(ref4).max()
        # and the rule is within its discretisation error of the converged integral
        ref = _tri_ref(alpha, x, f, p0, p0)
        assert np.abs(rd - ref).max() < 3e-2 * np.abs(ref).max()


# --- GMRES ----------------------------------------------------------------------------
def test_gmres_matches_dense_solve(orc):
    """Cell-BIE solve at viscosity ratio 5 on the 2-cell case: the oracle's GMRES
    solution against a dense solve of the operator assembled column by column
    (what BuildMat does, common/ModVelSolver.F90:772-819)."""
    from rbc3d_amd import configs
    w = configs.make_small(ncell=2, nlat=5, Nb=(16, 16, 16), seed=3)
    o = orc.from_workload(w)
    rhs = o.cell_rhs()
    n = o.cell_dof()
    cols = []
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        cols.append(o.cell_matmult(e))
    Amat = np.array(cols).T
    x = np.linalg.solve(Amat, rhs)
    o2 = orc.from_workload(w)
    v, niter, rn = o2.solve_rbc_vel()
    assert niter > 0
    assert relerr(o2._sol, x) < 1e-9


def test_gmres_restart_length(orc):
    """The oracle's GMRES restart length (KSPGMRESSetRestart; PETSc's 30 unless set): any
    length reaches the dense solution; two lengths longer than the iteration count take
    identical steps; a short length restarts and needs more iterations."""
    from rbc3d_amd import configs
    w = configs.make_small(ncell=2, nlat=5, Nb=(16, 16, 16), seed=3)
    o = orc.from_workload(w)
    rhs = o.cell_rhs()
    n = o.cell_dof()
    Amat = np.array([o.cell_matmult(np.eye(n)[j]) for j in range(n)]).T
    x = np.linalg.solve(Amat, rhs)
    res = {}
    for m in (3, 30, 200, 255):
        om = orc.from_workload(w)
        om.set_gmres_restart(m)
        v, niter, rn = om.solve_rbc_vel()
        res[m] = (niter, om._sol.copy())
        assert relerr(om._sol, x) < 1e-9, m
    assert res[200][0] < 200 and res[200][0] == res[255][0]
    assert np.array_equal(res[200][1], res[255][1])
    assert res[3][0] > res[200][0]
    with pytest.raises(AssertionError):
        orc.from_workload(w).set_gmres_restart(256)
