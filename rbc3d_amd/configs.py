"""Synthetic workloads for the Ewald-split boundary-integral matvec.

Builds the inputs the reference's Fortran host would hold in ``rbcs(:)`` and
``walls(:)`` (common/ModDataStruct.F90:109-196) for the five benchmark shapes:

* cell surfaces on the Gauss-latitude x uniform-longitude grid of ``RBC_Create``
  (common/ModRbc.F90:44-110, ``dealias_fac = 1`` so nlat = nlat0, nlon = 2 nlat);
* Evans-Fung biconcave discs ``z = +-0.5 sqrt(1-r^2)(0.2072 + 2.0026 r^2 - 1.1228 r^4)``
  (radius 1) and spheres, each given a uniform random SO(3) rotation and a
  uniform random centre with a minimum centre separation (rejection sampling);
* surface geometry (x, unit normal a3, area element detJ = |x_th x x_phi| / sin th)
  evaluated analytically from the parametrisation, the quantities
  ``RBC_ComputeGeometry`` (common/ModRbc.F90:404-509) supplies to the matvec;
* surface densities synthesised from random spherical-harmonic coefficients
  ~ N(0,1) (l+1)^-2 in SPHEREPACK's normalisation;
* a periodic cylindrical tube wall as a Tri3 mesh with a duplicated seam ring
  (the layout ``Wall_Build_V2V`` expects, common/ModWall.F90:66-114).

Arrays use the reference's Fortran column-major layouts: per cell
``x(nlat, nlon, 3)`` flattened in Fortran order and concatenated over cells.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

EF_C = (0.2072, 2.0026, -1.1228)
SURFACE_GAP = 0.1
WALL_GAP = 0.25  # closest approach of a cell node to the tube wall


def gauss_colat(nlat: int):
    """Gauss-Legendre colatitudes (ascending) and weights (sum 2), as SPHEREPACK's gaqd."""
    x, w = np.polynomial.legendre.leggauss(nlat)
    th = np.arccos(x[::-1].clip(-1, 1))
    w = w[::-1].copy()
    # polish in theta (Newton on P_n(cos th)) for full accuracy near the poles
    for _ in range(3):
        p, dp = _leg(nlat, np.cos(th))
        th = th - p / (-np.sin(th) * dp)
    p, dp = _leg(nlat, np.cos(th))
    w = 2.0 / (np.sin(th) ** 2 * dp ** 2)
    return th, w


def _leg(n, x):
    p0 = np.ones_like(x)
    p1 = x.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1)
    return p1, dp


def pbar_table(nlat: int, th: np.ndarray, mmax: int) -> np.ndarray:
    """SPHEREPACK-normalised Pbar(m, n, th): int_0^pi Pbar^2 sin = 1, no CS phase.
    Returns P[m, n, i]."""
    nt = len(th)
    P = np.zeros((mmax, nlat, nt))
    c, s = np.cos(th), np.sin(th)
    pmm = np.full(nt, math.sqrt(0.5))
    for m in range(mmax):
        if m > 0:
            pmm = math.sqrt((2 * m + 1) / (2 * m)) * s * pmm
        if m >= nlat:
            break
        P[m, m] = pmm
        if m + 1 < nlat:
            P[m, m + 1] = math.sqrt(2 * m + 3) * c * pmm
            for n in range(m + 2, nlat):
                an = math.sqrt((4 * n * n - 1) / (n * n - m * m))
                bn = math.sqrt(((2 * n + 1) * ((n - 1) ** 2 - m * m)) / ((2 * n - 3) * (n * n - m * m)))
                P[m, n] = an * c * P[m, n - 1] - bn * P[m, n - 2]
    return P


_PBAR_CACHE = {}


def _pbar_cached(nmax, th, mmax):
    key = (nmax, mmax, th.tobytes())
    P = _PBAR_CACHE.get(key)
    if P is None:
        P = _PBAR_CACHE[key] = pbar_table(nmax, th, mmax)
    return P


def sh_synth(a: np.ndarray, b: np.ndarray, th: np.ndarray, nlon: int) -> np.ndarray:
    """g(i, j) = sum_n .5 a[0,n] Pbar(0,n) + sum_{m>=1,n>=m} Pbar (a cos m phi - b sin m phi)."""
    mmax, nmax = a.shape
    P = _pbar_cached(nmax, th, mmax)
    phi = 2 * np.pi * np.arange(nlon) / nlon
    g = np.zeros((len(th), nlon))
    for m in range(mmax):
        Ga = (P[m] * a[m][:, None]).sum(0) * (0.5 if m == 0 else 1.0)
        Gb = (P[m] * b[m][:, None]).sum(0)
        g += Ga[:, None] * np.cos(m * phi)[None] - Gb[:, None] * np.sin(m * phi)[None]
    return g


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform SO(3) rotation from a uniform unit quaternion (Shoemake's method)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3),
                  math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def shape_param(kind: str, nlat: int, radius: float = 1.0):
    """Body-frame surface x(th, phi) and its th/phi derivatives on the Gauss grid."""
    th, _ = gauss_colat(nlat)
    nlon = 2 * nlat
    phi = 2 * np.pi * np.arange(nlon) / nlon
    T, PH = np.meshgrid(th, phi, indexing="ij")
    s, c = np.sin(T), np.cos(T)
    cp, sp = np.cos(PH), np.sin(PH)
    R = radius
    if kind == "disc":
        c0, c1, c2 = EF_C
        poly = c0 + c1 * s ** 2 + c2 * s ** 4
        dpoly = (2 * c1 * s + 4 * c2 * s ** 3) * c
        z = 0.5 * R * c * poly
        dz = 0.5 * R * (-s * poly + c * dpoly)
        x = np.stack([R * s * cp, R * s * sp, z], -1)
        xt = np.stack([R * c * cp, R * c * sp, dz], -1)
    elif kind == "sphere":
        x = np.stack([R * s * cp, R * s * sp, R * c], -1)
        xt = np.stack([R * c * cp, R * c * sp, -R * s], -1)
    else:
        raise ValueError(kind)
    xp = np.stack([-R * s * sp, R * s * cp, np.zeros_like(s)], -1)
    return th, x, xt, xp


def cell_volume(kind: str, nlat: int, radius: float = 1.0) -> float:
    th, w = gauss_colat(nlat)
    _, x, xt, xp = shape_param(kind, nlat, radius)
    n = np.cross(xt, xp)  # |x_th x x_phi| dth dphi; the Gauss weights carry sin(th) dth
    wt = w[:, None] * (2 * np.pi / (2 * nlat)) / np.sin(th)[:, None]
    return float((wt * (n * x).sum(-1)).sum() / 3.0)


def min_image(d, L):
    return d - np.round(d / L) * L


def place_centres(rng, n, L, minsep, cyl=None, max_tries=2_000_000):
    """Uniform random centres with a minimum periodic separation (rejection).
    ``cyl = (xc, yc, Rmax)`` restricts centres to a cylinder (tube case)."""
    L = np.asarray(L, float)
    cs = np.zeros((0, 3))
    tries = 0
    while len(cs) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could not place {n} cells (placed {len(cs)})")
        p = rng.random(3) * L
        if cyl is not None:
            xc, yc, rmax = cyl
            if (p[0] - xc) ** 2 + (p[1] - yc) ** 2 > rmax ** 2:
                continue
        if len(cs):
            d = min_image(cs - p, L)
            if (d * d).sum(1).min() < minsep * minsep:
                continue
        cs = np.vstack([cs, p])
    return cs


def place_cells_gap(rng, shapes, L, gap, max_tries=2_000_000, cyl=None):
    """Random sequential placement of rigidly rotated cells in a periodic box with a
    minimum surface gap: a candidate (uniform centre, uniform rotation) is rejected
    when any of its surface nodes comes closer than ``gap`` to a node of a placed
    cell.  ``shapes``: list of (body-frame nodes (n,3), bounding radius).
    ``cyl = (xc, yc, Rmax)``: every node must also lie within distance Rmax of the
    z-parallel axis through (xc, yc) (cells inside a tube).
    Placed nodes live in a periodic k-d tree, rebuilt every few insertions; the
    cells placed since the last rebuild are checked directly."""
    from scipy.spatial import cKDTree
    L = np.asarray(L, float)
    cs, Qs, rads = [], [], []
    all_nodes = []
    tree, ntree = None, 0
    tries = 0
    for pts, rad in shapes:
        while True:
            tries += 1
            if tries > max_tries:
                raise RuntimeError(f"could not place {len(shapes)} cells (placed {len(cs)})")
            c = rng.random(3) * L
            if cyl is not None and np.hypot(c[0] - cyl[0], c[1] - cyl[1]) > cyl[2] - 0.5 * rad:
                continue
            Q = random_rotation(rng)
            body = pts @ Q.T + c
            if cyl is not None and np.hypot(body[:, 0] - cyl[0], body[:, 1] - cyl[1]).max() > cyl[2]:
                continue
            mine = np.mod(body, L)
            ok = True
            if tree is not None:
                d, _ = tree.query(mine, k=1, distance_upper_bound=gap)
                ok = not np.isfinite(d).any()
            if ok and len(cs) > ntree:
                dc = min_image(np.asarray(cs[ntree:]) - c, L)
                near = np.nonzero(np.sqrt((dc * dc).sum(1)) < rad + np.asarray(rads[ntree:]) + gap)[0]
                for k in near:
                    dd = min_image(mine[:, None, :] - all_nodes[ntree + k][None, :, :], L)
                    if (dd * dd).sum(-1).min() < gap * gap:
                        ok = False
                        break
            if ok:
                cs.append(c)
                Qs.append(Q)
                rads.append(rad)
                all_nodes.append(mine)
                if len(cs) - ntree >= 16:
                    tree = cKDTree(np.concatenate(all_nodes), boxsize=L)
                    ntree = len(cs)
                break
    return np.asarray(cs), Qs


_COLAT_CACHE = {}


def random_density(rng, nlat, nlat0):
    """One surface field (nlat, nlon, 3) from SH coefficients ~ N(0,1)(l+1)^-2, l < nlat0."""
    if nlat not in _COLAT_CACHE:
        _COLAT_CACHE[nlat] = gauss_colat(nlat)
    th, _ = _COLAT_CACHE[nlat]
    nlon = 2 * nlat
    out = np.zeros((nlat, nlon, 3))
    n = np.arange(nlat0)
    decay = 1.0 / (n + 1.0) ** 2
    for d in range(3):
        a = rng.standard_normal((nlat0, nlat0)) * decay[None, :]
        b = rng.standard_normal((nlat0, nlat0)) * decay[None, :]
        a = np.triu(a)  # a[m, n] with n >= m
        b = np.triu(b)
        b[0] = 0.0
        out[:, :, d] = sh_synth(a, b, th, nlon)
    return out


def alpha_ref_from_xi(xi: float) -> float:
    """The reference's Ewald parameter alpha (erfc(sqrt(pi/alpha) r)) from the
    standard splitting xi (erfc(xi r)): alpha = pi / xi^2."""
    return math.pi / (xi * xi)


def xi_for_tolerance(rc: float, eps: float) -> float:
    """xi with erfc(xi rc) = eps (bisection)."""
    lo, hi = 0.1, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if math.erfc(mid * rc) > eps:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


@dataclass
class Cells:
    nlat0: np.ndarray
    nlat: np.ndarray
    celltype: np.ndarray
    x: np.ndarray      # concatenated (nlat,nlon,3) Fortran order
    a3: np.ndarray
    detj: np.ndarray   # concatenated (nlat,nlon)
    f: np.ndarray
    g: np.ndarray
    centres: np.ndarray = field(default=None)

    @property
    def ncell(self):
        return len(self.nlat)

    @property
    def npoint(self):
        return int((self.nlat * 2 * self.nlat).sum())


def build_cells(rng, kinds, nlats, radii, centres, celltypes, nlat0s=None, rots=None) -> Cells:
    xs, a3s, djs, fs, gs = [], [], [], [], []
    nlat0s = nlats if nlat0s is None else nlat0s
    cache = {}
    for k, (kind, nl, R, c, n0) in enumerate(zip(kinds, nlats, radii, centres, nlat0s)):
        key = (kind, nl, R)
        if key not in cache:
            cache[key] = shape_param(kind, nl, R)
        th, x, xt, xp = cache[key]
        Q = random_rotation(rng) if rots is None else rots[k]
        xr = x @ Q.T + c
        xtr = xt @ Q.T
        xpr = xp @ Q.T
        n = np.cross(xtr, xpr)
        nn = np.linalg.norm(n, axis=-1)
        a3 = n / nn[..., None]
        detj = nn / np.sin(th)[:, None]
        xs.append(xr.reshape(-1, order="F"))
        a3s.append(a3.reshape(-1, order="F"))
        djs.append(detj.reshape(-1, order="F"))
        fs.append(random_density(rng, nl, n0).reshape(-1, order="F"))
        gs.append(random_density(rng, nl, n0).reshape(-1, order="F"))
    return Cells(nlat0=np.asarray(nlat0s, np.int32), nlat=np.asarray(nlats, np.int32),
                 celltype=np.asarray(celltypes, np.int32), x=np.concatenate(xs),
                 a3=np.concatenate(a3s), detj=np.concatenate(djs), f=np.concatenate(fs),
                 g=np.concatenate(gs), centres=np.asarray(centres))


def make_tube(R=8.0, Lz=32.0, nazm=128, nax=160, centre=(8.25, 8.25)):
    """Structured periodic tube as Tri3: nazm x (nax+1) vertices (seam ring at z=Lz
    duplicates z=0), 2*nazm*nax triangles.  Returns x(nvert,3) Fortran-flat, e2v(nele,3)."""
    th = 2 * np.pi * np.arange(nazm) / nazm
    z = Lz * np.arange(nax + 1) / nax
    TH, Z = np.meshgrid(th, z, indexing="ij")  # vertex (a, k) -> a + nazm*k
    X = np.stack([centre[0] + R * np.cos(TH), centre[1] + R * np.sin(TH), Z], -1)
    xv = X.reshape(-1, 3, order="F")
    tris = []
    for k in range(nax):
        for a in range(nazm):
            a1 = (a + 1) % nazm
            v00 = a + nazm * k
            v10 = a1 + nazm * k
            v01 = a + nazm * (k + 1)
            v11 = a1 + nazm * (k + 1)
            tris.append((v00, v01, v10))
            tris.append((v10, v01, v11))
    e2v = np.asarray(tris, np.int32) + 1
    return xv.reshape(-1, order="F"), e2v.reshape(-1, order="F"), xv.shape[0], e2v.shape[0]


@dataclass
class Workload:
    name: str
    Lb: tuple
    alpha: float
    rc: float
    Nb: tuple
    P: int
    viscRat: tuple
    vBkg: tuple
    cells: Cells
    wall: tuple = None  # (x, e2v, nvert, nele)

    def conf_kwargs(self):
        return dict(Lb=self.Lb, alpha=self.alpha, rc=self.rc, Nb=self.Nb, P=self.P,
                    viscRat=self.viscRat, vBkg=self.vBkg)


_DATA = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)), "data")


def _placement(key, make):
    """Centres and rotations of a dense packing; cached as a small data file because
    random sequential placement at 20-30% volume fraction takes about a minute."""
    import os
    path = os.path.join(_DATA, f"placement_{key}.npz")
    if os.path.exists(path):
        z = np.load(path)
        return z["centres"], list(z["rots"])
    cs, rots = make()
    os.makedirs(_DATA, exist_ok=True)
    np.savez(path, centres=cs, rots=np.asarray(rots))
    return cs, rots


def make_small(ncell=1, nlat=7, L=8.0, Nb=(16, 16, 16), P=6, xi=2.0, viscRat=5.0, seed=1):
    """A few coarse cells in a cube (test-sized operators)."""
    rng = np.random.default_rng(seed)
    cs = place_centres(rng, ncell, (L, L, L), 2.2)
    cells = build_cells(rng, ["disc"] * ncell, [nlat] * ncell, [1.0] * ncell, cs, [1] * ncell)
    return Workload(f"small{ncell}x{nlat}", (L, L, L), alpha_ref_from_xi(xi), 2.0, tuple(Nb), P,
                    (viscRat,), (0.0, 0.0, 0.0), cells)


def make_workload(name: str, seed: int = 1, ncell: int = None, Nb: tuple = None) -> Workload:
    """The benchmark shapes.  ``ncell``/``Nb`` shrink a shape for quick tests."""
    rng = np.random.default_rng(seed)
    if name == "C1":
        n = ncell or 2
        L = 8.0
        cs = place_centres(rng, n, (L, L, L), 2.2)
        cells = build_cells(rng, ["disc"] * n, [13] * n, [1.0] * n, cs, [1] * n)
        return Workload("C1", (L, L, L), alpha_ref_from_xi(2.0), 2.0, Nb or (32, 32, 32), 6,
                        (5.0,), (0.0, 0.0, 0.0), cells)
    if name == "C2":
        n = ncell or 512
        vol = cell_volume("disc", 17)
        L = (n * vol / 0.2) ** (1.0 / 3.0)
        # a 2.2 minimum centre separation is infeasible at 20% volume fraction
        # (exclusion spheres would fill 71% of the box): use a surface gap instead
        body = shape_param("disc", 17)[1].reshape(-1, 3)
        cs, rots = _placement(f"C2_n{n}_L{L:.3f}_s{seed}", lambda: place_cells_gap(
            rng, [(body, 1.0)] * n, (L, L, L), SURFACE_GAP))
        cells = build_cells(rng, ["disc"] * n, [17] * n, [1.0] * n, cs, [1] * n, rots=rots)
        xi = xi_for_tolerance(2.0, 1e-8)
        return Workload("C2", (L, L, L), alpha_ref_from_xi(xi), 2.0, Nb or (128, 128, 128), 6,
                        (5.0,), (0.0, 0.0, 0.0), cells)
    if name == "C3":
        # 200 discs fill the R=8, Lz=32 tube to 25% haematocrit when their radius is
        # (0.25 * pi * 64 * 32 / (200 * V_disc(1)))^(1/3) = 1.722
        n = ncell or 200
        Lb = (16.5, 16.5, 32.0)
        rad = (0.25 * math.pi * 64.0 * 32.0 / (200 * cell_volume("disc", 13))) ** (1.0 / 3.0)
        body = shape_param("disc", 13, rad)[1].reshape(-1, 3)
        cs, rots = _placement(f"C3_n{n}_R{rad:.3f}_s{seed}", lambda: place_cells_gap(
            rng, [(body, rad)] * n, Lb, SURFACE_GAP, cyl=(8.25, 8.25, 8.0 - WALL_GAP)))
        cells = build_cells(rng, ["disc"] * n, [13] * n, [rad] * n, cs, [1] * n, rots=rots)
        xi = xi_for_tolerance(2.0, 1e-8)
        wall = make_tube()
        return Workload("C3", Lb, alpha_ref_from_xi(xi), 2.0, Nb or (96, 96, 180), 6,
                        (1.0,), (0.0, 0.0, 1.0), cells, wall)
    if name == "C3s":
        # a small tube for parity tests: radius 4, length 16, 40 x 50 quads
        n = ncell or 8
        Lb = (9.0, 9.0, 16.0)
        cs = place_centres(rng, n, Lb, 2.2, cyl=(4.5, 4.5, 4.0 - 1.3))
        cells = build_cells(rng, ["disc"] * n, [13] * n, [1.0] * n, cs, [1] * n)
        xi = xi_for_tolerance(2.0, 1e-8)
        wall = make_tube(R=4.0, Lz=16.0, nazm=40, nax=50, centre=(4.5, 4.5))
        return Workload("C3s", Lb, alpha_ref_from_xi(xi), 2.0, Nb or (32, 32, 60), 6,
                        (1.0,), (0.0, 0.0, 1.0), cells, wall)
    if name == "C5":
        n1, n2 = (3584, 512) if ncell is None else (ncell - ncell // 8, ncell // 8)
        v1 = cell_volume("disc", 13)
        v2 = 4.0 / 3.0 * math.pi * 1.4 ** 3
        L = ((n1 * v1 + n2 * v2) / 0.3) ** (1.0 / 3.0)
        kinds = ["disc"] * n1 + ["sphere"] * n2
        nl = [13] * n1 + [17] * n2
        rad = [1.0] * n1 + [1.4] * n2
        bd = shape_param("disc", 13)[1].reshape(-1, 3)
        bs = shape_param("sphere", 17, 1.4)[1].reshape(-1, 3)
        # random sequential addition jams below 30% with a 0.1 gap and mixed order; the
        # large spheres go first and the gap is 0.05
        shapes = [(bs, 1.4)] * n2 + [(bd, 1.0)] * n1
        kinds = ["sphere"] * n2 + ["disc"] * n1
        nl = [17] * n2 + [13] * n1
        rad = [1.4] * n2 + [1.0] * n1
        cs, rots = _placement(f"C5_n{n1 + n2}_g05_s{seed}", lambda: place_cells_gap(
            rng, shapes, (L, L, L), 0.05, max_tries=50_000_000))
        cells = build_cells(rng, kinds, nl, rad, cs, [1] * (n1 + n2), rots=rots)
        xi = xi_for_tolerance(2.0, 1e-8)
        return Workload("C5", (L, L, L), alpha_ref_from_xi(xi), 2.0, Nb or (256, 256, 256), 6,
                        (5.0,), (0.0, 0.0, 0.0), cells)
    raise ValueError(name)
