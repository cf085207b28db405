"""Problem construction with the reference's EnterData semantics.

Host-side logic (numpy only). It turns support and load rows into the free-DOF
list and load vector exactly as ``Implementation/EnterData.m`` does, generates
the synthetic densities used by the workloads, and applies the multigrid level
rules of ``Implementation/StartStrSLP.m:107-133``.

Numbering (0-based here, 1-based in the reference):
    node(ix, iy, iz) = iz*nx*ny + ix*ny + iy      (Implementation/EnterData.m:130)
    dof             = 3*node + c, c = 0:x 1:y 2:z  (Implementation/EnterData.m:146-150)
    elem(ex, ey, ez) = ez*nelx*nely + ex*nely + ey (Implementation/DegreesOfFreedom.m:40-54)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Problem:
    """The ``str`` struct fields the equilibrium solve reads."""

    nelx: int
    nely: int
    nelz: int
    el: float = 1.0
    eh: float = 1.0
    ew: float = 1.0
    E: float = 1.0
    nu: float = 0.3
    freedofs: np.ndarray = field(default=None)  # 1-based int64, sorted (str.freedofs)
    f: np.ndarray = field(default=None)         # load on free DOFs (str.f)
    supp: np.ndarray = field(default=None)
    loads: np.ndarray = field(default=None)
    fix: np.ndarray = field(default=None)       # fixed-density rows [ex0 exf ey0 eyf ez0 ezf val] (str.fix)
    fixed_dens: np.ndarray = field(default=None)      # 1-based elements with fixed densities (str.fixedDens)
    fixed_dens_val: np.ndarray = field(default=None)  # their values, 0 void / 1 solid (str.fixedDensVal)

    def __post_init__(self):
        if self.fixed_dens is None:
            self.fixed_dens = np.zeros(0, dtype=np.int64)
        if self.fixed_dens_val is None:
            self.fixed_dens_val = np.zeros(0, dtype=np.float64)

    def free_dens_mask(self) -> np.ndarray:
        """0/1 per element: 1 where the density is a design variable
        (str.freeDens, used as str.freeDensG by Implementation/Grad.m)."""
        m = np.ones(self.nelem, dtype=np.uint8)
        m[self.fixed_dens - 1] = 0
        return m

    @property
    def nx(self) -> int:
        return self.nelx + 1

    @property
    def ny(self) -> int:
        return self.nely + 1

    @property
    def nz(self) -> int:
        return self.nelz + 1

    @property
    def nnodes(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def ndof(self) -> int:
        return 3 * self.nnodes

    @property
    def nelem(self) -> int:
        return self.nelx * self.nely * self.nelz

    @property
    def nfree(self) -> int:
        return int(self.freedofs.size)

    def free_mask(self) -> np.ndarray:
        m = np.zeros(self.ndof, dtype=np.uint8)
        m[self.freedofs - 1] = 1
        return m

    def full(self, v_free: np.ndarray) -> np.ndarray:
        """Expand a free-DOF vector to all DOFs (fixed = 0), as u(str.freedofs) = v."""
        u = np.zeros(self.ndof, dtype=np.float64)
        u[self.freedofs - 1] = v_free
        return u


def _layer(coord: float, h: float) -> int:
    """1-based node layer nearest to a coordinate (Implementation/EnterData.m:119-124)."""
    return 1 + int(math.floor(coord / h)) + (1 if math.fmod(coord, h) > h / 2 else 0)


def _line_nodes(c0: float, cf: float, h: float, F):
    """Nodal shares of a unit line load on [c0, cf] (Implementation/EnterData.m:215-243)."""
    n0, nf = _layer(c0, h), _layer(cf, h)
    out = []
    if n0 == nf:
        out.append((0, (cf - c0)))
    else:
        out.append((0, (n0 - 0.5) * h - c0))
        for k in range(1, nf - n0):
            out.append((k, h))
        out.append((nf - n0, cf - (nf - 1.5) * h))
    return n0, [(k, [w * F[0], w * F[1], w * F[2]]) for k, w in out]


def _element_layer(coord: float, h: float, n: int, lower: bool) -> int:
    """1-based element layer of a fixed-density region bound
    (Implementation/EnterData.m:838-843): the nearest layer boundary, clamped
    to [1, n]."""
    v = int(math.floor(coord / h)) + (1 if math.fmod(coord, h) > h / 2 else 0)
    return max(v, 1) if lower else min(v, n)


def fixed_regions(nelx, nely, nelz, el, eh, ew, regions):
    """Rows ``[ex0 exf ey0 eyf ez0 ezf val]`` of str.fix and the fixed
    elements from void (val 0) and solid (val 1) regions given by coordinates
    ``[x0 xf y0 yf z0 zf val]`` (Implementation/EnterData.m:800-903). Returns
    (fix rows, str.fixedDens 1-based, str.fixedDensVal)."""
    fixd = np.zeros(nelx * nely * nelz, dtype=np.int8)
    val = np.zeros(nelx * nely * nelz)
    rows = []
    for r in np.atleast_2d(np.asarray(regions, dtype=np.float64)) if len(regions) else []:
        x0, xf, y0, yf, z0, zf, v = r
        ex0, exf = _element_layer(x0, el, nelx, True), _element_layer(xf, el, nelx, False)
        ey0, eyf = _element_layer(y0, eh, nely, True), _element_layer(yf, eh, nely, False)
        ez0, ezf = _element_layer(z0, ew, nelz, True), _element_layer(zf, ew, nelz, False)
        rows.append([ex0, exf, ey0, eyf, ez0, ezf, v])
        _mark_fixed(fixd, val, nelx, nely, [ex0, exf, ey0, eyf, ez0, ezf, v])
    dens = np.nonzero(fixd)[0]
    return np.asarray(rows, dtype=np.float64).reshape(-1, 7), dens.astype(np.int64) + 1, val[dens]


def _mark_fixed(fixd, val, nelx, nely, row):
    ex0, exf, ey0, eyf, ez0, ezf, v = (int(row[0]), int(row[1]), int(row[2]), int(row[3]), int(row[4]), int(row[5]),
                                       float(row[6]))
    for ez in range(ez0, ezf + 1):
        for ex in range(ex0, exf + 1):
            for ey in range(ey0, eyf + 1):
                e = (ey - 1) + (ex - 1) * nely + (ez - 1) * nelx * nely
                fixd[e] = 1
                if v == 1:
                    val[e] = 1.0


def fixed_from_rows(nelx, nely, nelz, fix):
    """str.fixedDens and str.fixedDensVal from stored str.fix rows (element
    layers, Implementation/EnterData.m:829-903)."""
    fixd = np.zeros(nelx * nely * nelz, dtype=np.int8)
    val = np.zeros(nelx * nely * nelz)
    for r in np.atleast_2d(np.asarray(fix, dtype=np.float64)):
        _mark_fixed(fixd, val, nelx, nely, r)
    dens = np.nonzero(fixd)[0]
    return dens.astype(np.int64) + 1, val[dens]


def from_rows(nelx, nely, nelz, supp, loads, l=None, h=None, w=None, E=1.0, nu=0.3, fix=None) -> Problem:
    """Build freedofs and f from support rows ``[x0 xf y0 yf z0 zf ix iy iz]``
    and load rows ``[x0 xf y0 yf z0 zf Fx Fy Fz]``
    (Implementation/EnterData.m:79-156 supports, :159-348 point and line
    loads, :350-782 loads on areas parallel to a coordinate plane, :790-797
    assembly), and the fixed densities from ``fix`` rows
    ``[ex0 exf ey0 eyf ez0 ezf val]`` (str.fix, :829-903). A load on an area
    spreads the per-length shares of the line rule in both directions: each
    node gets the product of its two one-dimensional shares times F.
    """
    l = float(nelx if l is None else l)
    h = float(nely if h is None else h)
    w = float(nelz if w is None else w)
    el, eh, ew = l / nelx, h / nely, w / nelz
    nx, ny, nz = nelx + 1, nely + 1, nelz + 1
    supp = np.atleast_2d(np.asarray(supp, dtype=np.float64))
    loads = np.atleast_2d(np.asarray(loads, dtype=np.float64))

    fixed = []
    for r in supp:
        x0, xf, y0, yf, z0, zf, ix, iy, iz = r
        nx0, nxf = _layer(x0, el), _layer(xf, el)
        ny0, nyf = _layer(y0, eh), _layer(yf, eh)
        nz0, nzf = _layer(z0, ew), _layer(zf, ew)
        for kz in range(nz0, nzf + 1):
            for kx in range(nx0, nxf + 1):
                for ky in range(ny0, nyf + 1):
                    node = (kz - 1) * nx * ny + (kx - 1) * ny + ky  # 1-based
                    for c, on in enumerate((ix, iy, iz)):
                        if on != 0:
                            fixed.append(3 * node - 2 + c)
    ndof = 3 * nx * ny * nz
    fixed = np.unique(np.asarray(fixed, dtype=np.int64))
    keep = np.ones(ndof + 1, dtype=bool)
    keep[0] = False
    keep[fixed] = False
    freedofs = np.nonzero(keep)[0].astype(np.int64)

    f = np.zeros(ndof, dtype=np.float64)
    for r in loads:
        x0, xf, y0, yf, z0, zf, Fx, Fy, Fz = r
        F = (Fx, Fy, Fz)
        dx, dy, dz = xf > x0, yf > y0, zf > z0
        base = lambda kx, ky, kz: (kz - 1) * nx * ny + (kx - 1) * ny + ky
        shares = []
        if not (dx or dy or dz):
            shares = [(base(_layer(x0, el), _layer(y0, eh), _layer(z0, ew)), list(F))]
        elif dx and not (dy or dz):
            n0, sh = _line_nodes(x0, xf, el, F)
            shares = [(base(n0 + k, _layer(y0, eh), _layer(z0, ew)), v) for k, v in sh]
        elif dy and not (dx or dz):
            n0, sh = _line_nodes(y0, yf, eh, F)
            shares = [(base(_layer(x0, el), n0 + k, _layer(z0, ew)), v) for k, v in sh]
        elif dz and not (dx or dy):
            n0, sh = _line_nodes(z0, zf, ew, F)
            shares = [(base(_layer(x0, el), _layer(y0, eh), n0 + k), v) for k, v in sh]
        elif dx and dy and dz:
            raise ValueError("volumetric loads are rejected (Implementation/EnterData.m:784-787)")
        else:
            one = (1.0, 1.0, 1.0)
            ax = _line_nodes(x0, xf, el, one) if dx else (_layer(x0, el), [(0, [1.0, 1.0, 1.0])])
            ay = _line_nodes(y0, yf, eh, one) if dy else (_layer(y0, eh), [(0, [1.0, 1.0, 1.0])])
            az = _line_nodes(z0, zf, ew, one) if dz else (_layer(z0, ew), [(0, [1.0, 1.0, 1.0])])
            shares = []
            for kz, wz in az[1]:
                for kx, wx in ax[1]:
                    for ky, wy in ay[1]:
                        # the two spanning directions' shares, multiplied first
                        wgt = [wa for wa, on in ((wx[0], dx), (wy[0], dy), (wz[0], dz)) if on]
                        sc = wgt[0] * wgt[1]
                        shares.append((base(ax[0] + kx, ay[0] + ky, az[0] + kz), [sc * F[0], sc * F[1], sc * F[2]]))
        for node, v in shares:
            for c in range(3):
                f[3 * node - 3 + c] += v[c]
    fixed_dens = fixed_val = None
    fixrows = None
    if fix is not None and np.asarray(fix).size:
        fixrows = np.atleast_2d(np.asarray(fix, dtype=np.float64))
        fixed_dens, fixed_val = fixed_from_rows(nelx, nely, nelz, fixrows)
    return Problem(nelx, nely, nelz, el, eh, ew, E, nu, freedofs, f[freedofs - 1].copy(),
                   supp, loads, fixrows, fixed_dens, fixed_val)


def from_mat(path, check=True) -> Problem:
    """Read a reference problem file (the ``str`` struct of a ``.mat`` as the
    reference stores them, e.g. ``Implementation/Tests/cb48x16x16.mat``) with
    ``scipy.io.loadmat``, which executes nothing. When the file keeps its
    support and load rows (str.supp, str.loads) the problem is rebuilt from
    them through :func:`from_rows`, and the fixed densities from str.fix; with
    ``check`` the rebuilt freedofs, f and fixedDens / fixedDensVal must equal
    the stored ones. Files without rows (entered interactively and saved
    without them) give their stored freedofs, f and fixed densities."""
    import scipy.io as sio
    d = sio.loadmat(path, squeeze_me=True, struct_as_record=False)
    if "str" not in d:
        raise ValueError(f"{path}: no 'str' struct")
    st = d["str"]
    names = set(st._fieldnames)
    for k in ("nelx", "nely", "nelz"):
        if k not in names:
            raise ValueError(f"{path}: str has no field {k!r}")
    get = lambda k, dflt=None: getattr(st, k) if k in names else dflt
    nelx, nely, nelz = int(st.nelx), int(st.nely), int(st.nelz)
    fix = get("fix")
    fix = None if fix is None or np.asarray(fix).size == 0 else np.atleast_2d(np.asarray(fix, dtype=np.float64))
    stored_fd = np.asarray(get("fixedDens", []), dtype=np.int64).ravel()
    stored_fv = np.asarray(get("fixedDensVal", []), dtype=np.float64).ravel()
    if "supp" in names and "loads" in names:
        prob = from_rows(nelx, nely, nelz, get("supp"), get("loads"), l=get("l"), h=get("h"), w=get("w"),
                         E=float(get("E", 1.0)), nu=float(get("nu", 0.3)), fix=fix)
        if check and "freedofs" in names and "f" in names:
            fd = np.asarray(st.freedofs, dtype=np.int64).ravel()
            f = np.asarray(st.f, dtype=np.float64).ravel()
            if not (np.array_equal(fd, prob.freedofs) and np.allclose(f, prob.f, rtol=0, atol=1e-12)):
                raise ValueError(f"{path}: stored freedofs/f differ from the ones rebuilt from supp/loads")
    else:
        l, h, w = (float(get(k, n)) for k, n in (("l", nelx), ("h", nely), ("w", nelz)))
        prob = Problem(nelx, nely, nelz, l / nelx, h / nely, w / nelz, float(get("E", 1.0)), float(get("nu", 0.3)),
                       np.asarray(st.freedofs, dtype=np.int64).ravel(), np.asarray(st.f, dtype=np.float64).ravel(),
                       fix=fix)
        if fix is not None:
            prob.fixed_dens, prob.fixed_dens_val = fixed_from_rows(nelx, nely, nelz, fix)
        else:
            prob.fixed_dens, prob.fixed_dens_val = stored_fd, stored_fv
    if check and "fixedDens" in names:
        if not (np.array_equal(prob.fixed_dens, stored_fd) and np.array_equal(prob.fixed_dens_val, stored_fv)):
            raise ValueError(f"{path}: stored fixedDens/fixedDensVal differ from the ones rebuilt from str.fix")
    return prob


def cantilever(nelx, nely, nelz, load="line", tip=None) -> Problem:
    """Cantilever of the benchmark workloads: face x = 0 clamped; either a unit
    downward line load along z at (x = nelx, y = 0), or a unit point load
    Fy = -1 at tip node (x = nelx, y = tip[0], z = tip[1])."""
    supp = [[0, 0, 0, nely, 0, nelz, 1, 1, 1]]
    if load == "line":
        loads = [[nelx, nelx, 0, 0, 0, nelz, 0, -1, 0]]
    else:
        ty, tz = tip
        loads = [[nelx, nelx, ty, ty, tz, tz, 0, -1, 0]]
    return from_rows(nelx, nely, nelz, supp, loads)


def filter_weights(rmin: float, opfilter: int = 2, el=1.0, eh=1.0, ew=1.0):
    """Offsets and weights of the density filter neighbourhood
    (Implementation/ElemNeighborhood.m:10-39,60-64)."""
    frz = int(math.floor(rmin / ew))
    offs, wts = [], []
    alpha = 1.0 / (2.0 * (rmin / 3.0) ** 2)
    beta = 1.0 / (2.0 * math.pi * (rmin / 3.0))
    for adz in range(frz + 1):
        dz2 = (adz * ew) ** 2
        relx = int(math.floor(math.sqrt(rmin * rmin - dz2) / el))
        for adx in range(relx + 1):
            dx2 = (adx * el) ** 2
            rely = int(math.floor(math.sqrt(rmin * rmin - dz2 - dx2) / eh))
            for ady in range(rely + 1):
                d2 = dz2 + dx2 + (ady * eh) ** 2
                if opfilter == 1:
                    wt = beta if d2 == 0 else math.exp(-d2 * alpha) * beta
                else:
                    wt = 1.0 if d2 == 0 else (rmin - math.sqrt(d2)) / rmin
                for sz in ((1, -1) if adz else (1,)):
                    for sx in ((1, -1) if adx else (1,)):
                        for sy in ((1, -1) if ady else (1,)):
                            offs.append((sx * adx, sy * ady, sz * adz))
                            wts.append(wt)
    return offs, wts


def apply_filter(x: np.ndarray, nelx, nely, nelz, rmin=1.5, opfilter=2, trans=False):
    """ApplyFilter on a structured grid without building W
    (Implementation/ApplyFilter.m:11-19 with W from ElemNeighborhood.m).
    No fixed elements."""
    offs, wts = filter_weights(rmin, opfilter)
    X = np.asarray(x, dtype=np.float64).reshape(nelz, nelx, nely)
    ones = np.ones_like(X)

    def wsum(a):
        out = np.zeros_like(a)
        for (dx, dy, dz), wt in zip(offs, wts):
            src = a[max(dz, 0):nelz + min(dz, 0), max(dx, 0):nelx + min(dx, 0), max(dy, 0):nely + min(dy, 0)]
            out[max(-dz, 0):nelz + min(-dz, 0), max(-dx, 0):nelx + min(-dx, 0), max(-dy, 0):nely + min(-dy, 0)] += wt * src
        return out

    wrow = wsum(ones)
    if trans:
        return wsum(X / wrow).reshape(-1)
    return (wsum(X) / wrow).reshape(-1)


def random_density(nelx, nely, nelz, seed, lo=1e-3, hi=1.0, rmin=1.5) -> np.ndarray:
    """Element densities drawn from U[lo, hi] with numpy's PCG64 generator
    (default_rng(seed)) in element order, then smoothed by the linear
    (opfilter = 2) density filter of radius rmin."""
    rng = np.random.default_rng(seed)
    raw = rng.uniform(lo, hi, size=nelx * nely * nelz)
    return apply_filter(raw, nelx, nely, nelz, rmin=rmin, opfilter=2)


def choose_ngrids(nelx, nely, nelz, ngrids, nmaxchol=60000, opchol=0):
    """Level-count rules of Implementation/StartStrSLP.m:107-133.

    Returns (ngrids, opchol, use_gmg). ngrids is reduced until every element
    count is divisible by 2^(ngrids-1); then, while the coarsest system is
    larger than nmaxchol, ngrids grows if divisibility allows, else opchol=1.
    """
    def ok(L):
        d = 2 ** (L - 1)
        return nelx % d == 0 and nely % d == 0 and nelz % d == 0

    L0 = ngrids
    while not ok(ngrids):
        ngrids -= 1
    if ngrids < 2:
        return L0, opchol, False
    if opchol == 0:
        while True:
            d = 2 ** (ngrids - 1)
            size = (nelx // d + 1) * (nely // d + 1) * (nelz // d + 1) * 3
            if size <= nmaxchol or opchol != 0:
                break
            if ok(ngrids + 1):
                ngrids += 1
            else:
                opchol = 1
    return ngrids, opchol, True
