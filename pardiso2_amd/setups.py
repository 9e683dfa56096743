"""Problem setups for the Stokes solve: grids, boundary-equation lines, markers.

The boundary grammar is the reference's text format for boundary equations
(parallel/in3mg.c:234-390; examples in parallel/init_inclusion.t3c:43-75):

    VAR  m10 m11  m20 m21  m30 m31  Const  Koef [dm1 dm2 dm3]  Koef1 ...

meaning, for every node in the index box, ``x = Const + Koef * x(shifted)``.
``x``/``y``/``z`` in an index stand for ``xnumx-1``/``ynumy-1``/``znumz-1``
plus the signed offset that follows, and a shift with magnitude > 1 denotes a
periodic partner (in3mg.c:339-341).  Lines are projected on every multigrid
level by the library (``stk_bc_line``), exactly as in3mg.c:307-324 does.

The generators below build the benchmark workloads: a unit-cube free-slip box
with gravity along -z, markers on a regular 2x2x2-per-cell lattice with seeded
jitter (in3mg.c:173-230 places them the same way), and viscosity/density
fields (sphere sinker, Gaussian random field, random spheres).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence, Tuple

import numpy as np

VARS = {"P": 0, "Vx": 1, "Vy": 2, "Vz": 3, "T": 4}

BcLine = Tuple[int, int, int, int, int, int, int, float, float, int, int, int]


def parse_index(tok: str, nmax: int) -> int:
    """Index token of the boundary grammar: 'x-1' -> nmax-1-1 (in3mg.c:258-263)."""
    if tok[0] in "xyz":
        rest = tok[1:]
        return nmax - 1 + (int(rest) if rest else 0)
    return int(tok)


def parse_bc_lines(text: str, nx: int, ny: int, nz: int) -> List[BcLine]:
    """Parse boundary-equation lines in the reference grammar (in3mg.c:240-290).

    Comment lines start with '/', the block ends at '~'.  P/Vx/Vy/Vz lines
    are the Stokes solve's boundary rows, T (4) lines the heat solve's
    (in3mg.c:243-247); other variables are rejected.
    """
    out: List[BcLine] = []
    for raw in text.splitlines():
        s = raw.replace("^M", "").strip()
        if not s or s.startswith("/"):
            continue
        if s.startswith("~"):
            break
        t = s.split()
        if t[0] not in VARS:
            raise ValueError(f"unsupported boundary variable {t[0]!r}")
        var = VARS[t[0]]
        m10, m11 = parse_index(t[1], nx), parse_index(t[2], nx)
        m20, m21 = parse_index(t[3], ny), parse_index(t[4], ny)
        m30, m31 = parse_index(t[5], nz), parse_index(t[6], nz)
        cnst, koef = float(t[7]), float(t[8])
        dm = (0, 0, 0)
        if koef:
            dm = (int(t[9]), int(t[10]), int(t[11]))
        out.append((var, m10, m11, m20, m21, m30, m31, cnst, koef, dm[0], dm[1], dm[2]))
    return out


def free_slip_lines(nx: int, ny: int, nz: int, gauge=(2, 2, 2), pgauge=0.0) -> List[BcLine]:
    """Free-slip box whose walls lie on grid lines 1 and n-2 of each axis.

    The outermost cell layer is a ghost layer: every unknown in it is a
    boundary equation (the coarse levels keep grid lines 0-2 and the last
    three, parallel/load3mg.c:2082-2098, so the walls sit on the same lines
    on every level).  Normal velocity is zero on the wall lines (and in the
    ghost lines beyond them); a tangential velocity in the ghost layer copies
    its interior neighbour (``Koef=1`` with a one-cell shift, the free-slip
    encoding of the reference grammar), which makes the shear stress vanish
    on the wall.  Ghost-layer pressures copy the adjacent interior cell and
    one interior cell carries the pressure gauge (P 2 2 2 2 2 2 0 0, as
    parallel/init_inclusion.t3c:53).  Later lines overwrite earlier ones
    (in3mg.c:325-331), so normal-velocity lines come after the copies.
    """
    X, Y, Z = nx - 1, ny - 1, nz - 1
    L: List[BcLine] = []
    # vx: tangential copies across the y- and z-ghost layers, then normal walls
    L.append((1, 0, X, 0, 0, 0, Z - 1, 0.0, 1.0, 0, +1, 0))
    L.append((1, 0, X, Y - 1, Y - 1, 0, Z - 1, 0.0, 1.0, 0, -1, 0))
    L.append((1, 0, X, 0, Y - 1, 0, 0, 0.0, 1.0, 0, 0, +1))
    L.append((1, 0, X, 0, Y - 1, Z - 1, Z - 1, 0.0, 1.0, 0, 0, -1))
    L.append((1, 0, 1, 0, Y - 1, 0, Z - 1, 0.0, 0.0, 0, 0, 0))
    L.append((1, X - 1, X, 0, Y - 1, 0, Z - 1, 0.0, 0.0, 0, 0, 0))
    # vy
    L.append((2, 0, 0, 0, Y, 0, Z - 1, 0.0, 1.0, +1, 0, 0))
    L.append((2, X - 1, X - 1, 0, Y, 0, Z - 1, 0.0, 1.0, -1, 0, 0))
    L.append((2, 0, X - 1, 0, Y, 0, 0, 0.0, 1.0, 0, 0, +1))
    L.append((2, 0, X - 1, 0, Y, Z - 1, Z - 1, 0.0, 1.0, 0, 0, -1))
    L.append((2, 0, X - 1, 0, 1, 0, Z - 1, 0.0, 0.0, 0, 0, 0))
    L.append((2, 0, X - 1, Y - 1, Y, 0, Z - 1, 0.0, 0.0, 0, 0, 0))
    # vz
    L.append((3, 0, 0, 0, Y - 1, 0, Z, 0.0, 1.0, +1, 0, 0))
    L.append((3, X - 1, X - 1, 0, Y - 1, 0, Z, 0.0, 1.0, -1, 0, 0))
    L.append((3, 0, X - 1, 0, 0, 0, Z, 0.0, 1.0, 0, +1, 0))
    L.append((3, 0, X - 1, Y - 1, Y - 1, 0, Z, 0.0, 1.0, 0, -1, 0))
    L.append((3, 0, X - 1, 0, Y - 1, 0, 1, 0.0, 0.0, 0, 0, 0))
    L.append((3, 0, X - 1, 0, Y - 1, Z - 1, Z, 0.0, 0.0, 0, 0, 0))
    # ghost-layer pressures copy the interior neighbour
    L.append((0, 1, 1, 1, Y, 1, Z, 0.0, 1.0, +1, 0, 0))
    L.append((0, X, X, 1, Y, 1, Z, 0.0, 1.0, -1, 0, 0))
    L.append((0, 1, X, 1, 1, 1, Z, 0.0, 1.0, 0, +1, 0))
    L.append((0, 1, X, Y, Y, 1, Z, 0.0, 1.0, 0, -1, 0))
    L.append((0, 1, X, 1, Y, 1, 1, 0.0, 1.0, 0, 0, +1))
    L.append((0, 1, X, 1, Y, Z, Z, 0.0, 1.0, 0, 0, -1))
    if gauge is not None:
        g = gauge
        L.append((0, g[0], g[0], g[1], g[1], g[2], g[2], float(pgauge), 0.0, 0, 0, 0))
    return L


def box_grid(n: int, size: float = 1.0) -> np.ndarray:
    """Grid lines for a box of the given size whose walls are lines 1 and n-2:
    uniform spacing h = size/(n-3), g[i] = i*h, so the grid starts at 0 as the
    reference's grids do (in3mg.c:102-108) and the box is [h, h+size]."""
    h = size / float(n - 3)
    return np.arange(n, dtype=np.float64) * h


def box_origin(n: int, size: float = 1.0) -> float:
    return size / float(n - 3)


def uniform_grid(n: int, size: float = 1.0) -> np.ndarray:
    """Grid lines as in in3mg.c:102-108: g[0]=0, g[i]=g[i-1]+size/(n-1)."""
    g = np.zeros(n, dtype=np.float64)
    step = size / float(n - 1)
    for i in range(1, n):
        g[i] = g[i - 1] + step
    return g


def levels_for(n: int, L: int) -> List[int]:
    """Level sizes mgx[l] = (mgx[l-1]-5)//2+5 (parallel/load3mg.c:2068-2070)."""
    out = [n]
    for _ in range(L):
        out.append((out[-1] - 5) // 2 + 5)
    return out


def markers_regular(gx, gy, gz, mnum=(2, 2, 2), jitter=0.0, seed=0):
    """Regular marker lattice, mnum per cell, cell by cell (in3mg.c:173-212).

    ``jitter`` is the half-width of a uniform perturbation in units of the
    marker spacing (the reference adds rand()%(2k+1)-k scaled by x/y/znonstab,
    in3mg.c:221-229); a seeded generator replaces rand().  Markers stay inside
    their cell for jitter < 0.5.
    """
    nx, ny, nz = len(gx), len(gy), len(gz)
    dxs = np.diff(gx) / mnum[0]
    dys = np.diff(gy) / mnum[1]
    dzs = np.diff(gz) / mnum[2]
    # cell-major order: m3 (z), m1 (x), m2 (y); inside: z, x, y
    cz = gz[:-1][:, None] + dzs[:, None] * (np.arange(mnum[2])[None, :] + 0.5)  # (nz-1, mz)
    cx = gx[:-1][:, None] + dxs[:, None] * (np.arange(mnum[0])[None, :] + 0.5)
    cy = gy[:-1][:, None] + dys[:, None] * (np.arange(mnum[1])[None, :] + 0.5)
    Z = np.broadcast_to(cz[:, None, None, :, None, None], (nz - 1, nx - 1, ny - 1, mnum[2], mnum[0], mnum[1]))
    X = np.broadcast_to(cx[None, :, None, None, :, None], Z.shape)
    Yy = np.broadcast_to(cy[None, None, :, None, None, :], Z.shape)
    mx = X.reshape(-1).astype(np.float64)
    my = Yy.reshape(-1).astype(np.float64)
    mz = Z.reshape(-1).astype(np.float64)
    if jitter:
        rng = np.random.default_rng(seed)
        sx = np.broadcast_to(dxs[None, :, None, None, None, None], Z.shape).reshape(-1)
        sy = np.broadcast_to(dys[None, None, :, None, None, None], Z.shape).reshape(-1)
        sz = np.broadcast_to(dzs[:, None, None, None, None, None], Z.shape).reshape(-1)
        mx = mx + rng.uniform(-jitter, jitter, mx.shape) * sx
        my = my + rng.uniform(-jitter, jitter, my.shape) * sy
        mz = mz + rng.uniform(-jitter, jitter, mz.shape) * sz
    return mx.astype(np.float32), my.astype(np.float32), mz.astype(np.float32)


def gaussian_random_field(n: int, corr: float, seed: int) -> np.ndarray:
    """Seeded stationary Gaussian field on an n^3 lattice over the unit cube,
    Gaussian covariance with correlation length ``corr``; zero mean, unit
    variance (spectral synthesis on a periodic 2n box, cropped)."""
    m = 2 * n
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(m, d=1.0 / (m - 1) * 1.0) * 2 * math.pi  # box length ~2
    kx, ky, kz = np.meshgrid(k, k, np.fft.rfftfreq(m, d=1.0 / (m - 1)) * 2 * math.pi, indexing="ij")
    spec = np.exp(-0.25 * (kx ** 2 + ky ** 2 + kz ** 2) * corr ** 2)
    noise = rng.standard_normal((m, m, m))
    f = np.fft.irfftn(np.fft.rfftn(noise) * np.sqrt(spec), s=(m, m, m), axes=(0, 1, 2))[:n, :n, :n]
    f = (f - f.mean()) / (f.std() + 1e-300)
    return f


def sample_trilinear(field: np.ndarray, x, y, z) -> np.ndarray:
    """Trilinear sample of a field on the unit cube lattice at points in [0,1]."""
    n = field.shape[0]
    fx = np.clip(np.asarray(x, np.float64) * (n - 1), 0, n - 1 - 1e-9)
    fy = np.clip(np.asarray(y, np.float64) * (n - 1), 0, n - 1 - 1e-9)
    fz = np.clip(np.asarray(z, np.float64) * (n - 1), 0, n - 1 - 1e-9)
    i, j, k = fx.astype(int), fy.astype(int), fz.astype(int)
    a, b, c = fx - i, fy - j, fz - k
    out = np.zeros_like(fx)
    for di in (0, 1):
        wa = a if di else 1 - a
        for dj in (0, 1):
            wb = b if dj else 1 - b
            for dk in (0, 1):
                wc = c if dk else 1 - c
                out += wa * wb * wc * field[i + di, j + dj, k + dk]
    return out


@dataclasses.dataclass
class Problem:
    """One Stokes instance: grid, hierarchy depth, knobs, boundary lines, markers."""

    n: int
    levels: int  # coarsest level index L (reference: multinum)
    gx: np.ndarray
    gy: np.ndarray
    gz: np.ndarray
    bc: List[BcLine]
    mx: np.ndarray
    my: np.ndarray
    mz: np.ndarray
    mnu: np.ndarray
    mro: np.ndarray
    gravity: Tuple[float, float, float] = (0.0, 0.0, -9.81)
    knobs: dict = dataclasses.field(default_factory=dict)
    name: str = ""
    periodic: Tuple[int, int, int] = (0, 0, 0)

    @property
    def nodes(self) -> int:
        return self.n ** 3


DEFAULT_KNOBS = dict(p0koef=1.0, p1koef=1.0, p2koef=0.0, v0koef=1.0, v1koef=1.0,
                     nubeg=1e18, nuend=1e25, viscmod=0)


def _box(n: int, L: int, mnu, mro, mx, my, mz, name, knobs=None) -> Problem:
    g = box_grid(n)
    k = dict(DEFAULT_KNOBS)
    if knobs:
        k.update(knobs)
    return Problem(n=n, levels=L, gx=g, gy=g.copy(), gz=g.copy(), bc=free_slip_lines(n, n, n, gauge=None),
                   mx=mx, my=my, mz=mz, mnu=mnu.astype(np.float32), mro=mro.astype(np.float32),
                   knobs=k, name=name)


def sinker(n: int = 33, L: int = 3, seed: int = 1, r: float = 0.2, eta_in=1e23, eta_out=1e20,
           rho_in=3200.0, rho_out=3300.0, knobs=None) -> Problem:
    """Spherical inclusion at the box centre in a uniform matrix."""
    g = box_grid(n)
    mx, my, mz = markers_regular(g, g, g, jitter=0.25, seed=seed)
    o = box_origin(n)
    d2 = (mx - o - 0.5) ** 2 + (my - o - 0.5) ** 2 + (mz - o - 0.5) ** 2
    inside = d2 < r * r
    mnu = np.where(inside, eta_in, eta_out)
    mro = np.where(inside, rho_in, rho_out)
    return _box(n, L, mnu, mro, mx, my, mz, f"sinker{n}", knobs)


def random_field(n: int = 129, L: int = 5, seed: int = 2, corr: float = 0.1, lo=18.0, hi=24.0,
                 rho=3300.0, drho=50.0, knobs=None, field_n: int = 65) -> Problem:
    """log10(eta) = seeded Gaussian random field (clipped), density 3300 +/- 50."""
    g = box_grid(n)
    mx, my, mz = markers_regular(g, g, g, jitter=0.25, seed=seed)
    f = gaussian_random_field(field_n, corr, seed)
    mid, half = 0.5 * (lo + hi), 0.5 * (hi - lo)
    o = box_origin(n)
    le = np.clip(mid + half * 0.5 * sample_trilinear(f, mx - o, my - o, mz - o), lo, hi)
    rng = np.random.default_rng(seed + 7)
    mro = rho + rng.uniform(-drho, drho, mx.shape)
    return _box(n, L, 10.0 ** le, mro, mx, my, mz, f"randfield{n}", knobs)


def random_spheres(n: int = 257, L: int = 6, seed: int = 3, count: int = 8, rmin=0.03, rmax=0.1,
                   eta_m=1e21, lc_lo=-3.0, lc_hi=3.0, rho_m=3300.0, rho_s=3200.0,
                   knobs=None) -> Problem:
    """Seeded random spheres, log-uniform viscosity contrast."""
    g = box_grid(n)
    mx, my, mz = markers_regular(g, g, g, jitter=0.25, seed=seed)
    rng = np.random.default_rng(seed)
    mnu = np.full(mx.shape, eta_m)
    mro = np.full(mx.shape, rho_m)
    o = box_origin(n)
    for _ in range(count):
        c = rng.uniform(0.15, 0.85, 3) + o
        r = rng.uniform(rmin, rmax)
        eta = eta_m * 10.0 ** rng.uniform(lc_lo, lc_hi)
        ins = (mx - c[0]) ** 2 + (my - c[1]) ** 2 + (mz - c[2]) ** 2 < r * r
        mnu[ins] = eta
        mro[ins] = rho_s
    return _box(n, L, mnu, mro, mx, my, mz, f"spheres{n}", knobs)


def coarsest_levels(n: int, target: int = 8) -> int:
    """Smallest L whose coarsest level has at most ``target`` nodes per side."""
    L, m = 0, n
    while m > target:
        m = (m - 5) // 2 + 5
        L += 1
    return L


def inclusions(n: int, L: int = None, seed: int = 0, count: int = None, rmin=0.05, rmax=0.15, eta_m=1e21,
               decades=(0.0, 6.0), rho_m=3300.0, rho_s=3200.0, knobs=None) -> Problem:
    """1-4 seeded spherical inclusions, each with viscosity eta_m * 10^(+-u),
    u log-uniform contrast in ``decades`` (stronger or weaker at random)."""
    L = coarsest_levels(n) if L is None else L
    g = box_grid(n)
    mx, my, mz = markers_regular(g, g, g, jitter=0.25, seed=seed)
    rng = np.random.default_rng(1000 + seed)
    count = int(rng.integers(1, 5)) if count is None else count
    mnu = np.full(mx.shape, eta_m)
    mro = np.full(mx.shape, rho_m)
    o = box_origin(n)
    for _ in range(count):
        c = rng.uniform(0.2, 0.8, 3) + o
        r = rng.uniform(rmin, rmax)
        eta = eta_m * 10.0 ** (rng.choice([-1.0, 1.0]) * rng.uniform(*decades))
        ins = (mx - c[0]) ** 2 + (my - c[1]) ** 2 + (mz - c[2]) ** 2 < r * r
        mnu[ins] = eta
        mro[ins] = rho_s
    return _box(n, L, mnu, mro, mx, my, mz, f"inclusions{n}", knobs)


def served_request(n: int, seed: int, decades: float, knobs=None) -> Problem:
    """Random-field viscosity spanning ``decades`` orders of magnitude around
    1e21 (seeded), two-phase density 3300 / 3200 split by the field's sign."""
    L = coarsest_levels(n)
    g = box_grid(n)
    mx, my, mz = markers_regular(g, g, g, jitter=0.25, seed=seed)
    f = gaussian_random_field(33, 0.1, seed)
    o = box_origin(n)
    s = sample_trilinear(f, mx - o, my - o, mz - o)
    le = np.clip(21.0 + 0.25 * decades * s, 21.0 - 0.5 * decades, 21.0 + 0.5 * decades)
    mro = np.where(s > 0, 3200.0, 3300.0)
    return _box(n, L, 10.0 ** le, mro, mx, my, mz, f"served{n}", knobs)


# Boundary block of parallel/init_inclusion.t3c:43-75 (the reference's simple
# shear setup, periodic in x and y): pressure fixed in cell (2,2,2); on the z
# walls (outer lines) vz = 0 and the first/last vx cell centres held at +1/-1,
# vy at 0; the last x line of vx and the last y line of vy copy line 0.
PERIODIC_SHEAR_BC = """
 P     2   2     2   2     2   2     0       0   0 0 0
 Vx    0   x     0   y-1   0   0     1       0   0 0 0
 Vy    0   x-1   0   y     0   0     0       0   0 0 0
 Vz    0   x-1   0   y-1   0   0     0       0   0 0 0
 Vx    0   x     0   y-1   z-1 z-1  -1       0   0 0 0
 Vy    0   x-1   0   y     z-1 z-1   0       0   0 0 0
 Vz    0   x-1   0   y-1   z   z     0       0   0 0 0
 Vy    0   x-1   y   y     1   z-2   0       1   +0 -100 +0
 Vx    x   x     0   y-1   1   z-2   0       1  -100  +0 +0
"""


def periodic_shear(n: int = 33, L: int = None, nz: int = None, seed: int = 5, count: int = 1, r: float = 0.2,
                   eta_m: float = 1.0, eta_in: float = 1e3, zsize: float = 1.16, jitter: float = 0.25,
                   knobs=None) -> Problem:
    """The reference's periodic simple-shear box (parallel/init_inclusion.t3c):
    x and y periodic over [0, 1], z walls at 0 and ``zsize`` driving vx = +1 /
    -1, no gravity, ``count`` seeded spherical inclusions of radius ``r``
    (viscosity ``eta_in`` in an ``eta_m`` matrix; count 0 leaves the box
    isoviscous)."""
    nz = n if nz is None else nz
    L = coarsest_levels(min(n, nz)) if L is None else L
    gx = uniform_grid(n, 1.0)
    gz = uniform_grid(nz, zsize)
    mx, my, mz = markers_regular(gx, gx, gz, jitter=jitter, seed=seed)
    rng = np.random.default_rng(seed)
    mnu = np.full(mx.shape, eta_m)
    mro = np.full(mx.shape, 3000.0)
    for k in range(count):
        c = np.array([0.5, 0.5, 0.5 * zsize]) if k == 0 else rng.uniform([0.2, 0.2, 0.2 * zsize], [0.8, 0.8, 0.8 * zsize])
        ins = (mx - c[0]) ** 2 + (my - c[1]) ** 2 + (mz - c[2]) ** 2 < r * r
        mnu[ins] = eta_in
    k = dict(DEFAULT_KNOBS, nubeg=min(eta_m, eta_in) * 1e-3, nuend=max(eta_m, eta_in) * 1e3)
    if knobs:
        k.update(knobs)
    return Problem(n=n, levels=L, gx=gx, gy=gx.copy(), gz=gz, bc=parse_bc_lines(PERIODIC_SHEAR_BC, n, n, nz),
                   mx=mx, my=my, mz=mz, mnu=mnu.astype(np.float32), mro=mro.astype(np.float32),
                   gravity=(0.0, 0.0, 0.0), knobs=k, name=f"pshear{n}", periodic=(1, 1, 0))


# ---------------------------------------------------------------------------
# Heat equation (parallel/heat3mg.c): node properties, temperature and
# boundary lines of variable T in the reference grammar.
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class HeatProblem:
    n: int
    levels: int
    gx: np.ndarray
    gy: np.ndarray
    gz: np.ndarray
    bc: List[BcLine]
    ro: np.ndarray   # FP32 node arrays of level 0, node order (m3, m1, m2)
    cp: np.ndarray
    kt: np.ndarray
    tk: np.ndarray   # temperature at the start of the step (K)
    tk1: np.ndarray  # heat source term of the step (the reference's tk1, W/m^3)
    timestep: float
    periodic: Tuple[int, int, int] = (0, 0, 0)
    name: str = ""


def heat_bc_text(t_top: float, t_bottom: float, xperiodic: bool = False) -> str:
    """T boundary lines: fixed temperatures on the z faces, zero flux on the
    other faces (copy of the inner neighbour), or x periodic (the last x line
    copies line 0, in3mg.c:339-341)."""
    lines = [f" T 0 x 0 y 0 0 {t_top} 0", f" T 0 x 0 y z z {t_bottom} 0"]
    if xperiodic:
        lines.append(" T x x 0 y 1 z-1 0 1 -100 0 0")
    else:
        lines += [" T 0 0 0 y 1 z-1 0 1 +1 0 0", " T x x 0 y 1 z-1 0 1 -1 0 0"]
    lines += [" T 0 x 0 0 1 z-1 0 1 0 +1 0", " T 0 x y y 1 z-1 0 1 0 -1 0"]
    return "\n".join(lines) + "\n"


def heat_box(n: int = 65, L: int = None, seed: int = 4, timestep: float = 1e4, xperiodic: bool = False,
             kt_lo: float = 1.0, kt_hi: float = 5.0, corr: float = 0.15) -> HeatProblem:
    """Unit box, one implicit temperature step: conductivity a seeded Gaussian
    random field in [kt_lo, kt_hi] W/m/K, rho 3000-3300 kg/m^3, cp 1000-1200
    J/kg/K, temperature (scaled to [0, 1]) 0 on top (z = 0) to 1 at the
    bottom with a seeded perturbation, a source term 1e-6 scaled by a second
    field.  timestep 1e4 s makes diffusion dominate each cell (k dt /
    (rho cp h^2) ~ 40 at 65^3, ~650 at 257^3), so the step needs the
    multigrid.  The temperature is kept of order one: the reference's error
    measure reads the FP32 temperature, whose rounding sets its floor."""
    L = coarsest_levels(n) if L is None else L
    g = uniform_grid(n, 1.0)
    zz, xx, yy = np.meshgrid(g, g, g, indexing="ij")  # node order (m3, m1, m2)
    f1 = gaussian_random_field(33, corr, seed)
    f2 = gaussian_random_field(33, corr, seed + 1)
    u1 = np.clip(0.5 + 0.25 * sample_trilinear(f1, xx.ravel(), yy.ravel(), zz.ravel()), 0.0, 1.0)
    u2 = np.clip(0.5 + 0.25 * sample_trilinear(f2, xx.ravel(), yy.ravel(), zz.ravel()), 0.0, 1.0)
    rng = np.random.default_rng(seed)
    kt = kt_lo + (kt_hi - kt_lo) * u1
    ro = 3000.0 + 300.0 * u2
    cp = rng.uniform(1000.0, 1200.0, kt.shape)
    tk = zz.ravel() + 0.02 * (u1 - 0.5)
    tk1 = 1e-6 * (0.5 + u2)
    bc = parse_bc_lines(heat_bc_text(0.0, 1.0, xperiodic), n, n, n)
    return HeatProblem(n=n, levels=L, gx=g, gy=g.copy(), gz=g.copy(), bc=bc, ro=ro.astype(np.float32),
                       cp=cp.astype(np.float32), kt=kt.astype(np.float32), tk=tk.astype(np.float32),
                       tk1=tk1.astype(np.float32), timestep=timestep,
                       periodic=(1 if xperiodic else 0, 0, 0), name=f"heat{n}")
