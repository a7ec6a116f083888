"""Pins of the oracle's per-cell geometry and membrane mechanics (oracle/mech_oracle.c,
the restatement of common/ModRbc.F90:404-969) by what the mathematics fixes:

- a sphere's exact geometry (normals, detJ = R^2, metric, curvature b = -a/R, area,
  volume, centroid, meshSize);
- an ellipsoid's exact volume (the quadrature is exact for it) and normals, and its
  curvature tensor against the closed form b_ij = a3 . d^2x/dth_i dth_j, converging with
  the grid;
- the Skalak force as the variational derivative of the membrane energy: for a
  volume-neutral smooth displacement dx, sum f.dx dA = dE/dh by central differences of
  the energy sum(W dA_ref), W = ES/8 (I1^2 + 2 I1 - 2 I2) + ED/8 I2^2 (the reference's
  tension is half the Cauchy tension of ES/4 (I1^2 + 2 I1 - 2 I2) + ED/4 I2^2);
- zero force in the reference state, zero net force and torque on a deformed shape;
- the projection's idempotence and exactness on band-limited fields.
A plausible slip (a transposed index, a dropped term, a wrong sign of the curvature
term, a missing sin th) fails at least one of these."""
import numpy as np
import pytest

from oracle import oracle_py as O


def grid(nlat):
    th = np.zeros(nlat)
    w = np.zeros(nlat)
    O.L().orc_gaqd(nlat, th, w)
    phi = 2 * np.pi * np.arange(2 * nlat) / (2 * nlat)
    return th, phi, w * (2 * np.pi / (2 * nlat))


def flat(arr):
    """(nlat, nlon, k...) numpy array -> the reference's column-major flat layout"""
    return np.asfortranarray(arr).ravel(order="F")


def unflat(v, nlat, *tail):
    return np.reshape(v, (nlat, 2 * nlat) + tail, order="F")


def shape_x(nlat, fn):
    th, phi, _ = grid(nlat)
    T, P = np.meshgrid(th, phi, indexing="ij")
    return flat(np.stack(fn(T, P), axis=-1)), T, P


def sphere(R, c=(0.0, 0.0, 0.0)):
    return lambda T, P: (c[0] + R * np.sin(T) * np.cos(P), c[1] + R * np.sin(T) * np.sin(P),
                         c[2] + R * np.cos(T))


def ellipsoid(A, B, C):
    return lambda T, P: (A * np.sin(T) * np.cos(P), B * np.sin(T) * np.sin(P), C * np.cos(T))


def test_sphere_geometry():
    nlat, R, c = 17, 1.3, (0.2, -0.4, 1.1)
    x, T, P = shape_x(nlat, sphere(R, c))
    g = O.rbc_geometry(x)
    a3 = unflat(g["a3"], nlat, 3)
    X = unflat(x, nlat, 3)
    assert np.abs(a3 - (X - np.array(c)) / R).max() < 1e-12
    assert np.abs(unflat(g["detj"], nlat) / R**2 - 1).max() < 1e-12
    a = unflat(g["a"], nlat, 2, 2)
    b = unflat(g["b"], nlat, 2, 2)
    ar = unflat(g["a_rcp"], nlat, 2, 2)
    assert np.abs(a[..., 0, 0] - R**2).max() < 1e-12
    assert np.abs(a[..., 1, 1] - (R * np.sin(T))**2).max() < 1e-12
    assert np.abs(a[..., 0, 1]).max() < 1e-12
    assert np.abs(b + a / R).max() < 1e-11
    assert np.abs(np.einsum("...ij,...jk->...ik", a, ar) - np.eye(2)).max() < 1e-12
    assert abs(g["area"] / (4 * np.pi * R**2) - 1) < 1e-13
    assert abs(g["vol"] / (4 / 3 * np.pi * R**3) - 1) < 1e-13
    assert np.abs(g["xc"] - c).max() < 1e-13
    assert abs(g["mesh"] - np.sqrt(g["area"]) / nlat) < 1e-15


def test_ellipsoid_volume_normals_and_curvature():
    A, B, C = 1.2, 0.9, 0.6
    errs = []
    for nlat in (12, 24):
        x, T, P = shape_x(nlat, ellipsoid(A, B, C))
        g = O.rbc_geometry(x)
        X = unflat(x, nlat, 3)
        assert abs(g["vol"] / (4 / 3 * np.pi * A * B * C) - 1) < 1e-13
        n = X / np.array([A * A, B * B, C * C])
        n /= np.linalg.norm(n, axis=-1, keepdims=True)
        assert np.abs(unflat(g["a3"], nlat, 3) - n).max() < 1e-12
        st, ct, sp, cp = np.sin(T), np.cos(T), np.sin(P), np.cos(P)
        xtt = np.stack([-A * st * cp, -B * st * sp, -C * ct], -1)
        xtp = np.stack([-A * ct * sp, B * ct * cp, 0 * T], -1)
        xpp = np.stack([-A * st * cp, -B * st * sp, 0 * T], -1)
        bx = np.zeros(T.shape + (2, 2))
        bx[..., 0, 0] = np.sum(n * xtt, -1)
        bx[..., 0, 1] = bx[..., 1, 0] = np.sum(n * xtp, -1)
        bx[..., 1, 1] = np.sum(n * xpp, -1)
        errs.append(np.abs(unflat(g["b"], nlat, 2, 2) - bx).max() / np.abs(bx).max())
    assert errs[1] < 1e-5 and errs[1] < errs[0] / 100, errs


def deformed(nlat, seed=0, amp=0.08, lmax=4):
    """an ellipsoid carrying a smooth random displacement (degrees <= lmax)"""
    rng = np.random.default_rng(seed)
    th, phi, _ = grid(nlat)
    T, P = np.meshgrid(th, phi, indexing="ij")
    base = np.stack(ellipsoid(1.1, 0.95, 0.7)(T, P), -1)
    d = np.zeros_like(base)
    for l in range(1, lmax + 1):
        for m in range(0, l + 1):
            for k in range(3):
                ylm = np.cos(T) ** (l - m) * np.sin(T) ** m
                ph = rng.uniform(0, 2 * np.pi)
                d[..., k] += rng.standard_normal() / (l + 1) ** 2 * ylm * np.cos(m * P + ph)
    return base + amp * d, T


def skalak_energy(geo, ref, ES, ED, w):
    nlat = geo["nlat"]
    a = unflat(geo["a"], nlat, 2, 2)
    arf = unflat(ref["a_rcp"], nlat, 2, 2)
    V2 = np.einsum("...ij,...jk->...ik", a, arf)
    I1 = V2[..., 0, 0] + V2[..., 1, 1] - 2
    I2 = V2[..., 0, 0] * V2[..., 1, 1] - V2[..., 0, 1] * V2[..., 1, 0] - 1
    W = ES / 8 * (I1**2 + 2 * I1 - 2 * I2) + ED / 8 * I2**2
    return float(np.sum(W * unflat(ref["detj"], nlat) * w[:, None]))


def test_skalak_force_is_the_energy_variation():
    nlat, ES, ED = 20, 1.0, 20.0
    _, _, w = grid(nlat)
    xr, _ = deformed(nlat, seed=1, amp=0.0)
    ref = O.rbc_geometry(flat(xr))
    X, _ = deformed(nlat, seed=3, amp=0.1)
    geo = O.rbc_geometry(flat(X))
    f = unflat(O.shell_res_force(geo, ref, ES, ED, 0.0), nlat, 3)
    # a smooth displacement with no first-order volume change
    dX, _ = deformed(nlat, seed=7, amp=1.0, lmax=3)
    dX = dX - np.stack(ellipsoid(1.1, 0.95, 0.7)(*np.meshgrid(grid(nlat)[0], grid(nlat)[1],
                                                            indexing="ij")), -1)
    dA = unflat(geo["detj"], nlat) * w[:, None]
    a3 = unflat(geo["a3"], nlat, 3)
    dV = np.sum(np.sum(a3 * dX, -1) * dA)
    dX = dX - dV / geo["area"] * a3
    work = np.sum(np.sum(f * dX, -1) * dA)
    h = 1e-5
    Ep = skalak_energy(O.rbc_geometry(flat(X + h * dX)), ref, ES, ED, w)
    Em = skalak_energy(O.rbc_geometry(flat(X - h * dX)), ref, ES, ED, w)
    dE = (Ep - Em) / (2 * h)
    print(f"virtual work {work:.10e}, dE/dh {dE:.10e}")
    # the residual force is the energy's gradient (Compute_Rbc_Vel applies -f)
    assert abs(work - dE) <= 1e-6 * abs(dE)


def test_reference_state_and_inflation():
    nlat = 17
    X, _ = deformed(nlat, seed=2)
    g = O.rbc_geometry(flat(X))
    f = O.shell_res_force(g, g, 1.0, 10.0, 0.1)
    assert np.abs(f).max() < 1e-10
    # a uniformly stretched sphere: isotropic tension, a uniform normal force (Laplace
    # law) that the subtraction of the mean normal force removes
    s0 = O.rbc_geometry(shape_x(nlat, sphere(1.0))[0])
    s1 = O.rbc_geometry(shape_x(nlat, sphere(1.1))[0])
    f = O.shell_res_force(s1, s0, 1.0, 10.0, 0.5)
    assert np.abs(f).max() < 1e-10


def test_net_force_and_torque_vanish():
    nlat = 24
    _, _, w = grid(nlat)
    xr, _ = deformed(nlat, seed=1, amp=0.0)
    ref = O.rbc_geometry(flat(xr))
    X, _ = deformed(nlat, seed=5, amp=0.1)
    geo = O.rbc_geometry(flat(X))
    dA = (unflat(geo["detj"], nlat) * w[:, None])[..., None]
    for ES, ED, EB in ((1.0, 20.0, 0.0), (0.0, 0.0, 0.3), (1.0, 20.0, 0.3)):
        f = unflat(O.shell_res_force(geo, ref, ES, ED, EB), nlat, 3)
        scale = np.sum(np.linalg.norm(f, axis=-1)[..., None] * dA)
        F = np.sum(f * dA, axis=(0, 1))
        print(f"ES {ES} ED {ED} EB {EB}: |net force| / sum|f| dA = {np.abs(F).max() / scale:.2e}")
        assert np.abs(F).max() < 1e-6 * scale
        if EB == 0.0:
            Tq = np.sum(np.cross(X - geo["xc"], f) * dA, axis=(0, 1))
            assert np.abs(Tq).max() < 1e-6 * scale


def test_sph_project():
    nlat, nlat0 = 16, 11
    rng = np.random.default_rng(4)
    u = rng.standard_normal(2 * nlat * nlat * 3)
    p = O.rbc_sph_project(u, nlat, nlat0)
    assert np.abs(O.rbc_sph_project(p, nlat, nlat0) - p).max() < 1e-12
    full = O.rbc_sph_project(u, nlat, nlat)
    assert np.abs(O.rbc_sph_project(full, nlat, nlat) - full).max() < 1e-12
    # a field of degree below nlat0 passes unchanged
    x, _, _ = shape_x(nlat, ellipsoid(1.0, 0.8, 0.6))
    assert np.abs(O.rbc_sph_project(x, nlat, nlat0) - x).max() < 1e-13
    # the dropped part is orthogonal to what is kept
    _, _, w = grid(nlat)
    dA = np.tile(w, 2 * nlat)
    r = (u - p).reshape(3, -1)
    pp = p.reshape(3, -1)
    assert abs(np.sum(r * pp * dA)) < 1e-10 * np.sum(pp * pp * dA)


def shape_fn(seed, amp, lmax=4, axes=(1.1, 0.95, 0.7)):
    """the surface of deformed() as a function of (th, phi) arrays"""
    rng = np.random.default_rng(seed)
    terms = []
    for l in range(1, lmax + 1):
        for m in range(0, l + 1):
            for k in range(3):
                ph = rng.uniform(0, 2 * np.pi)
                terms.append((l, m, k, rng.standard_normal() / (l + 1) ** 2, ph))

    def X(T, P):
        out = np.stack(ellipsoid(*axes)(T, P), -1)
        for l, m, k, c, ph in terms:
            out[..., k] += amp * c * np.cos(T) ** (l - m) * np.sin(T) ** m * np.cos(m * P + ph)
        return out
    return X


def D(fn, k, T, P, h=1e-2):
    """fourth-order central difference of fn(T, P) along th (k = 0) or phi (k = 1)"""
    e = (h, 0.0) if k == 0 else (0.0, h)
    f = lambda s: fn(T + s * e[0], P + s * e[1])
    return (-f(2) + 8 * f(1) - 8 * f(-1) + f(-2)) / (12 * h)


def geom_fn(X):
    """tangents, metric, inverse metric, normal and curvature tensor of the surface X"""
    def g(T, P):
        a1, a2 = D(X, 0, T, P), D(X, 1, T, P)
        A = np.stack([np.stack([np.sum(a1 * a1, -1), np.sum(a1 * a2, -1)], -1),
                      np.stack([np.sum(a2 * a1, -1), np.sum(a2 * a2, -1)], -1)], -2)
        n = np.cross(a1, a2)
        n /= np.linalg.norm(n, axis=-1, keepdims=True)
        xx = [[D(lambda t, p: D(X, i, t, p), k, T, P) for k in range(2)] for i in range(2)]
        B = np.stack([np.stack([np.sum(n * xx[i][k], -1) for k in range(2)], -1)
                      for i in range(2)], -2)
        return dict(a1=a1, a2=a2, a=A, ar=np.linalg.inv(A), a3=n, b=B, xx=xx)
    return g


def analytic_force(X, Xref, ES, ED, EB, T, P, w):
    """Shell_ResForce restated with Christoffel symbols G^p_ik = a^p . d_k a_i and
    derivatives taken by finite differences of the exact surface:
    t^{pq}_{;k} = d_k t^{pq} + G^p_{ik} t^{iq} + G^q_{ik} t^{pi}."""
    gX, gR = geom_fn(X), geom_fn(Xref)
    mm = lambda x, y: np.einsum("...ij,...jk->...ik", x, y)

    def tau(t, p):
        g, r = gX(t, p), gR(t, p)
        V2 = mm(g["a"], r["ar"])
        I1 = V2[..., 0, 0] + V2[..., 1, 1] - 2
        I2 = V2[..., 0, 0] * V2[..., 1, 1] - V2[..., 0, 1] * V2[..., 1, 0] - 1
        JS = np.sqrt(I2 + 1)
        return ((0.5 * ES / JS * (I1 + 1))[..., None, None] * r["ar"]
                + (0.5 * JS * (ED * I2 - ES))[..., None, None] * g["ar"])

    def bnd(t, p):
        g, r = gX(t, p), gR(t, p)
        return -EB * mm(mm(g["ar"], g["b"]) - mm(r["ar"], r["b"]), g["ar"])

    def chris(t, p):
        g = gX(t, p)
        rcp = [g["ar"][..., q, 0, None] * g["a1"] + g["ar"][..., q, 1, None] * g["a2"]
               for q in range(2)]
        G = np.zeros(t.shape + (2, 2, 2))
        for q in range(2):
            for i in range(2):
                for k in range(2):
                    G[..., q, i, k] = np.sum(rcp[q] * g["xx"][i][k], -1)
        return G

    def div_first(tf):
        def v(t, p):
            tt, G = tf(t, p), chris(t, p)
            dt = np.stack([D(tf, k, t, p) for k in range(2)], -1)  # dt[p, q, k]
            cov = dt + np.einsum("...pik,...iq->...pqk", G, tt) + \
                np.einsum("...qik,...pi->...pqk", G, tt)
            return cov[..., 0, :, 0] + cov[..., 1, :, 1]
        return v

    tau_n = div_first(bnd)
    g = gX(T, P)
    tn = tau_n(T, P)
    ft = -div_first(tau)(T, P) + np.einsum("...ij,...j->...i", g["ar"],
                                             np.einsum("...ij,...j->...i", g["b"], tn))
    G = chris(T, P)
    divn = D(lambda t, p: tau_n(t, p)[..., 0], 0, T, P) + \
        D(lambda t, p: tau_n(t, p)[..., 1], 1, T, P) + np.einsum("...iki,...k->...", G, tn)
    fn = -divn - np.sum(tau(T, P) * g["b"], axis=(-2, -1))
    dAr = np.sqrt(np.linalg.det(g["a"])) / np.sin(T) * w[:, None]
    fn = fn - np.sum(fn * dAr) / np.sum(dAr)
    return ft[..., 0, None] * g["a1"] + ft[..., 1, None] * g["a2"] + fn[..., None] * g["a3"]


def test_force_against_the_christoffel_form():
    """Covariant derivatives through the Cartesian embedding (the reference's route) against
    the textbook Christoffel form on the exact surface: pins the index order of the
    derivatives, the divergence over the first index and the curvature coupling terms, for
    the non-symmetric bending moment too."""
    nlat = 24
    th, phi, w = grid(nlat)
    T, P = np.meshgrid(th, phi, indexing="ij")
    Xr = shape_fn(1, 0.05, lmax=3)
    Xc = shape_fn(4, 0.1, lmax=3)
    ref = O.rbc_geometry(flat(Xr(T, P)))
    geo = O.rbc_geometry(flat(Xc(T, P)))
    for ES, ED, EB in ((1.0, 20.0, 0.0), (0.0, 0.0, 0.3)):
        f = unflat(O.shell_res_force(geo, ref, ES, ED, EB), nlat, 3)
        fc = analytic_force(Xc, Xr, ES, ED, EB, T, P, w)
        err = np.abs(f - fc).max() / np.abs(fc).max()
        print(f"ES {ES} ED {ED} EB {EB}: embedding vs Christoffel form {err:.2e}")
        assert err < 1e-3
