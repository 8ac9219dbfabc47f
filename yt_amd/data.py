"""In-memory grid datasets: the inputs the ray-caster's callers hand it.

Restates the parts of yt's data model the ray-caster consumes:
  load_uniform_grid (nprocs split)   yt/loaders.py:204-449,
                                     yt/utilities/decompose.py:20-150
  load_amr_grids                     yt/loaders.py:450-...
  child mask / blocks                selector_object.pxi:372-401,
                                     data_containers.py:1505-1517
  get_vertex_centered_data(no_ghost) grid_patch.py:270-312,
                                     interpolators.pyx:177-226
plus the seeded synthetic fields of the measured workloads.
"""

import numpy as np


# --------------------------------------------------------------------------
# Synthetic fields

def gaussian_random_field(n, seed, slope=-3.0, dtype=np.float64):
    """Gaussian random field on an n^3 grid with P(k) ~ k^slope: white noise
    -> rFFT -> * k^(slope/2) (k=0 mode 0) -> irFFT -> mean 0, sigma 1."""
    try:
        import scipy.fft as sfft

        rfftn = lambda a: sfft.rfftn(a, workers=-1)  # noqa: E731
        irfftn = lambda a, s: sfft.irfftn(a, s=s, workers=-1)  # noqa: E731
    except ImportError:  # pragma: no cover
        rfftn = np.fft.rfftn
        irfftn = lambda a, s: np.fft.irfftn(a, s=s)  # noqa: E731
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, n, n))
    wk = rfftn(w)
    del w
    kx = np.fft.fftfreq(n)[:, None, None]
    ky = np.fft.fftfreq(n)[None, :, None]
    kz = np.fft.rfftfreq(n)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    with np.errstate(divide="ignore"):
        amp = np.where(k2 > 0, k2 ** (slope / 4.0), 0.0)
    wk *= amp
    del amp, k2
    g = irfftn(wk, (n, n, n))
    del wk
    g -= g.mean()
    g /= g.std()
    return g.astype(dtype, copy=False)


def lognormal_density(n, seed=1001, sigma_ln=1.0, dtype=np.float64):
    """rho = exp(sigma_ln * g) for a k^-3 Gaussian random field g."""
    g = gaussian_random_field(n, seed)
    np.multiply(g, sigma_ln, out=g)
    np.exp(g, out=g)
    return g.astype(dtype, copy=False)


def temperature_from_density(rho, seed=1002, sigma_ln=0.3):
    """T = rho^0.6 * exp(0.3 h), h standard-normal white noise."""
    rng = np.random.default_rng(seed)
    h = rng.standard_normal(rho.shape)
    return rho ** 0.6 * np.exp(sigma_ln * h)


# --------------------------------------------------------------------------
# Grids and datasets

class Grid:
    """A grid patch: edges, dimensions, level, cell-centred fields."""

    def __init__(self, gid, left_edge, right_edge, dims, level, fields):
        self.id = gid
        self.LeftEdge = np.asarray(left_edge, dtype="float64")
        self.RightEdge = np.asarray(right_edge, dtype="float64")
        self.ActiveDimensions = np.asarray(dims, dtype="int64")
        self.Level = int(level)
        self.dds = (self.RightEdge - self.LeftEdge) / self.ActiveDimensions
        self.fields = fields
        self.child_mask = np.ones(tuple(self.ActiveDimensions), dtype=bool)
        self.Children = []
        self.Parent = None

    def __getitem__(self, name):
        return self.fields[name]

    def get_vertex_centered_data(self, fields, no_ghost=True):
        """grid_patch.py:270-296, no_ghost path only."""
        if not no_ghost:
            raise NotImplementedError("ghost-zone vertex centring needs neighbour grids")
        return {f: ghost_zone_interpolate_no_ghost(np.asarray(self[f], dtype="=f8"))
                for f in set(fields)}


def ghost_zone_interpolate_no_ghost(a):
    """interpolators.pyx:177-226 with rf=1, input_left=0.5, output_left=0:
    vertex values are 8-cell averages inside, linear extrapolation on faces.
    Same per-element operation order as the reference."""
    n = a.shape
    axes = []
    for d in range(3):
        o = np.arange(n[d] + 1, dtype="float64")
        ropos = (o - 0.5) * 1.0
        ii = np.clip(ropos.astype(np.int64), 0, n[d] - 2)
        xp = ropos - ii
        xm = 1.0 - xp
        axes.append((ii, xp, xm))
    (ii, xp, xm), (ij, yp, ym), (ik, zp, zm) = axes
    xp, xm = xp[:, None, None], xm[:, None, None]
    yp, ym = yp[None, :, None], ym[None, :, None]
    zp, zm = zp[None, None, :], zm[None, None, :]

    def g(di, dj, dk):
        return a[np.ix_(ii + di, ij + dj, ik + dk)]

    out = g(0, 0, 0) * (xm * ym * zm)
    out = out + g(1, 0, 0) * (xp * ym * zm)
    out = out + g(0, 1, 0) * (xm * yp * zm)
    out = out + g(0, 0, 1) * (xm * ym * zp)
    out = out + g(1, 0, 1) * (xp * ym * zp)
    out = out + g(0, 1, 1) * (xm * yp * zp)
    out = out + g(1, 1, 0) * (xp * yp * zm)
    out = out + g(1, 1, 1) * (xp * yp * zp)
    return out


class GridDataset:
    """A grid dataset with a child-masked `blocks` iterator."""

    def __init__(self, grids, domain_left_edge, domain_right_edge, domain_dimensions):
        self.grids = grids
        self.domain_left_edge = np.asarray(domain_left_edge, dtype="float64")
        self.domain_right_edge = np.asarray(domain_right_edge, dtype="float64")
        self.domain_dimensions = np.asarray(domain_dimensions, dtype="int64")
        self.domain_width = self.domain_right_edge - self.domain_left_edge
        self.domain_center = 0.5 * (self.domain_left_edge + self.domain_right_edge)
        self.max_level = max(g.Level for g in grids)
        self._compute_child_masks()

    @property
    def ds(self):
        return self

    def all_data(self):
        return self

    def _compute_child_masks(self):
        by_level = {}
        for g in self.grids:
            by_level.setdefault(g.Level, []).append(g)
        for g in self.grids:
            g.child_mask[...] = True
            g.Children = []
        for lvl, gs in by_level.items():
            kids = by_level.get(lvl + 1, [])
            if not kids:
                continue
            ple = np.array([p.LeftEdge for p in gs])
            pre = np.array([p.RightEdge for p in gs])
            for c in kids:
                lo = np.maximum(ple, c.LeftEdge)
                hi = np.minimum(pre, c.RightEdge)
                for pi in np.nonzero(np.all(hi > lo, axis=1))[0]:
                    p = gs[pi]
                    i0 = np.rint((lo[pi] - p.LeftEdge) / p.dds).astype(np.int64)
                    i1 = np.rint((hi[pi] - p.LeftEdge) / p.dds).astype(np.int64)
                    p.child_mask[i0[0]:i1[0], i0[1]:i1[1], i0[2]:i1[2]] = False
                    p.Children.append(c)
                    c.Parent = p

    def blocks(self):
        """(grid, mask) pairs in index order; grids whose every cell is covered
        by children are skipped (data_containers.py:1513-1516)."""
        for g in self.grids:
            m = g.child_mask
            if not m.any():
                continue
            yield g, m


# decompose.py:4-104 (get_psize) and :106-150 (split_array)

def _primes(x):
    return x and x[:1] + _primes([n for n in x if n % x[0]])


def _decompose_to_primes(max_prime):
    for prime in _primes(list(range(2, max_prime))):
        if prime * prime > max_prime:
            break
        while max_prime % prime == 0:
            yield prime
            max_prime //= prime
    if max_prime > 1:
        yield max_prime


def _evaluate(n_d, pieces, ldom):
    eff_dim = (n_d > 1).sum()
    exp = float(eff_dim - 1) / float(eff_dim)
    ideal_bsize = eff_dim * pieces ** (1.0 / eff_dim) * np.prod(n_d) ** exp
    mask = np.where(n_d > 1)
    nd_arr = np.array(n_d, dtype=np.float64)[mask]
    bsize = int(np.sum(ldom[mask] / nd_arr * np.prod(nd_arr)))
    load_balance = float(np.prod(n_d)) / (float(pieces) * np.prod((n_d - 1) // ldom + 1))
    quality = load_balance / (1 + 0.25 * (bsize / ideal_bsize - 1.0))
    quality *= 1.0 - (0.001 * ldom[0] + 0.0001 * ldom[1]) / pieces
    if np.any(ldom > n_d):
        quality = 0
    return quality


def get_psize(n_d, pieces):
    factors = list(_decompose_to_primes(pieces))
    temp = np.bincount(factors)
    fac = np.array([(p, temp[p], (temp[p] + 1) * (temp[p] + 2) // 2)
                    for p in np.unique(factors)], dtype="int64")
    nfactors = len(fac[:, 2])
    best = 0.0
    p_size = np.ones(3, dtype=np.int64)
    if pieces == 1:
        return p_size
    while np.all(fac[:, 2] > 0):
        ldom = np.ones(3, dtype=np.int64)
        for nfac in range(nfactors):
            i = int(np.sqrt(0.25 + 2 * (fac[nfac, 2] - 1)) - 0.5)
            k = fac[nfac, 2] - (1 + i * (i + 1) // 2)
            i = fac[nfac, 1] - i
            j = fac[nfac, 1] - (i + k)
            ldom *= fac[nfac, 0] ** np.array([i, j, k])
        quality = _evaluate(n_d, pieces, ldom)
        if quality > best:
            best = quality
            p_size = ldom
        for j in range(nfactors):
            if fac[j, 2] > 1:
                fac[j, 2] -= 1
                break
            elif j < nfactors - 1:
                fac[j, 2] = int((fac[j, 1] + 1) * (fac[j, 1] + 2) / 2)
            else:
                fac[:, 2] = 0
    return p_size


def split_array(gle, gre, shape, psize):
    n_d = np.array(shape, dtype=np.int64)
    dds = (gre - gle) / shape
    out = []
    for i in range(psize[0]):
        for j in range(psize[1]):
            for k in range(psize[2]):
                piece = np.array((i, j, k), dtype=np.int64)
                lei = n_d * piece // psize
                rei = n_d * (piece + np.ones(3, dtype=np.int64)) // psize
                out.append((gle + lei * dds, gle + rei * dds, rei - lei,
                            np.s_[lei[0]:rei[0], lei[1]:rei[1], lei[2]:rei[2]]))
    return out


def load_uniform_grid(data, bbox=None, nprocs=1, psize=None):
    """data: {name: (n0, n1, n2) array}; returns a GridDataset of nprocs grids
    (loaders.py:204-449)."""
    if bbox is None:
        bbox = np.array([[0.0, 1.0]] * 3)
    bbox = np.asarray(bbox, dtype="float64")
    shape = next(iter(data.values())).shape
    if nprocs > 1 or psize is not None:
        if psize is None:
            psize = get_psize(np.array(shape), nprocs)
        grids = []
        for gid, (le, re, dims, sl) in enumerate(split_array(bbox[:, 0], bbox[:, 1], shape,
                                                              np.asarray(psize))):
            grids.append(Grid(gid, le, re, dims, 0, {k: v[sl] for k, v in data.items()}))
    else:
        grids = [Grid(0, bbox[:, 0], bbox[:, 1], shape, 0, dict(data))]
    return GridDataset(grids, bbox[:, 0], bbox[:, 1], shape)


def load_amr_grids(grid_data, domain_dimensions, bbox=None):
    """grid_data: list of dicts with left_edge, right_edge, level, dimensions
    and field arrays (loaders.py:450-...)."""
    if bbox is None:
        bbox = np.array([[0.0, 1.0]] * 3)
    bbox = np.asarray(bbox, dtype="float64")
    reserved = {"left_edge", "right_edge", "level", "dimensions"}
    grids = []
    for gid, g in enumerate(grid_data):
        fields = {k: v for k, v in g.items() if k not in reserved}
        grids.append(Grid(gid, g["left_edge"], g["right_edge"], g["dimensions"], g["level"],
                          fields))
    return GridDataset(grids, bbox[:, 0], bbox[:, 1], domain_dimensions)


def _trilinear_prolong(parent_block, r=2):
    """Cell-centred trilinear prolongation by factor r of a (n0,n1,n2) block
    (edge cells use clamped neighbours)."""
    a = parent_block
    for ax in range(3):
        n = a.shape[ax]
        f = (np.arange(n * r) + 0.5) / r - 0.5
        i0 = np.clip(np.floor(f).astype(np.int64), 0, n - 1)
        i1 = np.clip(i0 + 1, 0, n - 1)
        w = np.clip(f - np.floor(f), 0.0, 1.0)
        shp = [1, 1, 1]
        shp[ax] = n * r
        w = w.reshape(shp)
        a = np.take(a, i0, axis=ax) * (1 - w) + np.take(a, i1, axis=ax) * w
    return a


def make_amr_hierarchy(root_n=128, levels=3, seed_density=1001, seed_flag=1003,
                       percentile=90.0, block=16, max_block=32, sigma_level=0.1,
                       blocks_per_level=(330, 280, 230)):
    """The seeded AMR workload: a root_n^3 lognormal root grid (seed
    seed_density) and `levels` levels of refinement by 2.  At each level the
    parent cells above the given percentile are flagged; the (block/2)^3
    parent-cell blocks with the most flagged cells (up to blocks_per_level[l]
    of them, ties broken by a seeded shuffle) are refined into block^3 child
    grids, and adjacent selected blocks along the last axis are merged into
    grids up to max_block cells long.  Child data is the trilinear
    prolongation of the parent times a seeded lognormal perturbation
    (sigma_level).  With the defaults this makes about 600 grids."""
    rng = np.random.default_rng(seed_flag)
    rho = lognormal_density(root_n, seed=seed_density)
    grid_data = [dict(left_edge=np.zeros(3), right_edge=np.ones(3), level=0,
                      dimensions=np.array([root_n] * 3), density=rho)]
    parents = [grid_data[0]]
    pb = block // 2  # parent cells per child block edge
    for lvl in range(1, levels + 1):
        allv = np.concatenate([p["density"].ravel() for p in parents])
        thresh = np.percentile(allv, percentile)
        cands = []  # (count, tiebreak, parent index, bi, bj, bk)
        for pi, p in enumerate(parents):
            dims = np.asarray(p["dimensions"])
            nb = dims // pb
            if np.any(nb == 0):
                continue
            sub = p["density"][:nb[0] * pb, :nb[1] * pb, :nb[2] * pb]
            cnt = (sub > thresh).reshape(nb[0], pb, nb[1], pb, nb[2], pb).sum(axis=(1, 3, 5))
            for (bi, bj, bk), c in np.ndenumerate(cnt):
                if c > 0:
                    cands.append((int(c), rng.random(), pi, bi, bj, bk))
        if not cands:
            break
        cands.sort(key=lambda x: (-x[0], x[1]))
        nsel = blocks_per_level[min(lvl - 1, len(blocks_per_level) - 1)]
        chosen = {}
        for c in cands[:nsel]:
            chosen.setdefault((c[2], c[3], c[4]), []).append(c[5])
        children = []
        for (pi, bi, bj), ks in sorted(chosen.items()):
            p = parents[pi]
            pd = p["density"]
            pdds = (np.asarray(p["right_edge"]) - np.asarray(p["left_edge"])) / np.asarray(p["dimensions"])
            runs = []
            for k in sorted(ks):
                if runs and k == runs[-1][1] and (runs[-1][1] + 1 - runs[-1][0]) * block <= max_block:
                    runs[-1][1] = k + 1
                else:
                    runs.append([k, k + 1])
            for k0, k1 in runs:
                lo = np.array([bi * pb, bj * pb, k0 * pb])
                hi = np.array([(bi + 1) * pb, (bj + 1) * pb, k1 * pb])
                sub = pd[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
                fine = _trilinear_prolong(sub, 2)
                fine = fine * np.exp(sigma_level * rng.standard_normal(fine.shape))
                le = np.asarray(p["left_edge"]) + lo * pdds
                re = np.asarray(p["left_edge"]) + hi * pdds
                children.append(dict(left_edge=le, right_edge=re, level=lvl,
                                     dimensions=np.array(fine.shape), density=fine))
        grid_data.extend(children)
        parents = children
    return grid_data
