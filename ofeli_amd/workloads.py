"""Seeded synthetic inputs of the benchmark configurations.

A counter-based SplitMix64 stream (seed 0x0FE11000 + config id) feeds every
random quantity, so the same arrays are produced on any host.  Meshes come
from the library's structured mesh constructors; the fields below are nodal
or per-element arrays handed to the assembly.
"""
import numpy as np

SEED_BASE = 0x0FE11000
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


class SplitMix64:
    def __init__(self, seed):
        self.state = np.uint64(seed)

    def next_u64(self, n):
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + idx * np.uint64(0x9E3779B97F4A7C15)
            self.state = self.state + np.uint64(n) * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n):
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        u1 = np.maximum(u1, 1e-300)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def gaussian_bumps(rng, xy, nb=8, amp=0.5):
    """f = 1 + amp * sum of nb bumps exp(-|x-c|^2 / (2 s^2)), c in [0.1,0.9]^2, s in [0.02,0.1]."""
    c = rng.uniform(2 * nb, 0.1, 0.9).reshape(nb, 2)
    s = rng.uniform(nb, 0.02, 0.1)
    f = np.ones(xy.shape[0])
    for k in range(nb):
        d2 = (xy[:, 0] - c[k, 0]) ** 2 + (xy[:, 1] - c[k, 1]) ** 2
        f += amp * np.exp(-d2 / (2.0 * s[k] ** 2))
    return f


def fourier_field(rng, pts, modes=64, corr=0.1):
    """Gaussian field of variance 1: g = sqrt(2/M) sum cos(w.x + phi), w ~ N(0, corr^-2 I)."""
    dim = pts.shape[1]
    w = rng.normal(modes * dim).reshape(modes, dim) / corr
    phi = rng.uniform(modes, 0.0, 2.0 * np.pi)
    g = np.zeros(pts.shape[0])
    for m in range(modes):
        g += np.cos(pts @ w[m] + phi[m])
    return np.sqrt(2.0 / modes) * g


def fourier_field_device(rng, pts, modes=64, corr=0.1):
    """fourier_field with the same random modes (drawn from rng on the host)
    evaluated by torch on the device of the tensor pts (float64)."""
    import torch
    dim = pts.shape[1]
    w = torch.as_tensor(rng.normal(modes * dim).reshape(modes, dim) / corr, device=pts.device)
    phi = torch.as_tensor(rng.uniform(modes, 0.0, 2.0 * np.pi), device=pts.device)
    g = torch.zeros(pts.shape[0], dtype=torch.float64, device=pts.device)
    for m in range(modes):
        g += torch.cos(pts @ w[m] + phi[m])
    return np.sqrt(2.0 / modes) * g


def centroids(coord, elem):
    return coord[elem - 1].mean(axis=1)


def config0(n=128):
    """2D Poisson, Laplace2DT3, n x n, f = 1 + 0.5 * 8 bumps, Dirichlet 0."""
    from . import api
    coord, code, elem = api.mesh2d(n, n)
    rng = SplitMix64(SEED_BASE + 0)
    f = gaussian_bumps(rng, coord)
    return dict(coord=coord, code=code, elem=elem, f=f, coef=None)


def config1(n=96):
    """3D variable-k Poisson on Kuhn tets, k_e = exp(g(centroid)), f = 1, Dirichlet 0."""
    from . import api
    coord, code, elem = api.mesh3d_kuhn(n, n, n)
    rng = SplitMix64(SEED_BASE + 1)
    g = fourier_field(rng, centroids(coord, elem))
    return dict(coord=coord, code=code, elem=elem, f=np.ones(coord.shape[0]), coef=np.exp(g))


def config2(nsys=512, nmin=16, nmax=96):
    """Batched steady heat (DC2DT3): nsys unit squares, n x n cells with n
    uniform in [nmin, nmax], per-element lognormal k (sigma 0.75), bump
    source per system, Dirichlet 0.  Returns the concatenated mesh, the
    system row ranges, k and f."""
    from . import api
    rng = SplitMix64(SEED_BASE + 2)
    sizes = nmin + (rng.next_u64(nsys) % np.uint64(nmax - nmin + 1)).astype(np.int64)
    meshes, ks, fs = [], [], []
    for n in sizes:
        coord, code, elem = api.mesh2d(int(n), int(n))
        meshes.append((coord, code, elem))
        ks.append(np.exp(0.75 * rng.normal(elem.shape[0])))
        fs.append(gaussian_bumps(rng, coord))
    coord, code, elem, sys_rows = api.concat_meshes(meshes)
    return dict(coord=coord, code=code, elem=elem, sys_rows=sys_rows, sizes=sizes,
                coef=np.concatenate(ks), f=np.concatenate(fs), meshes=meshes,
                ks=ks, fs=fs)


def config3(n=512):
    """Transient heat (DC2DT3, backward Euler): n x n cells, k_e = exp(g) with
    g the 64-mode Gaussian field of correlation 0.1 at element centroids,
    rho = cp = 1, dt = 1e-3, Dirichlet 0, u^0 = 0; source: a Gaussian bump
    (sigma 0.03, unit integral) moving on a seeded Lissajous path."""
    from . import api
    coord, code, elem = api.mesh2d(n, n)
    rng = SplitMix64(SEED_BASE + 3)
    g = fourier_field(rng, centroids(coord, elem))
    wx, wy = rng.uniform(2, 2 * np.pi * 0.5, 2 * np.pi * 1.5)
    phi = float(rng.uniform(1, 0.0, 2 * np.pi)[0])
    sigma = 0.03
    path = dict(wx=float(wx), wy=float(wy), phi=phi, sigma=sigma, amp=1.0 / (2 * np.pi * sigma ** 2),
                radius=0.3)
    return dict(coord=coord, code=code, elem=elem, coef=np.exp(g), dt=1e-3, path=path)


def source_centre(path, t):
    return (0.5 + path["radius"] * np.sin(path["wx"] * t + path["phi"]),
            0.5 + path["radius"] * np.sin(path["wy"] * t))


def source_field(xy, path, t):
    """Nodal source at time t; xy is a numpy array or a torch tensor."""
    cx, cy = source_centre(path, t)
    d2 = (xy[:, 0] - cx) ** 2 + (xy[:, 1] - cy) ** 2
    e = -d2 / (2.0 * path["sigma"] ** 2)
    return path["amp"] * (e.exp() if hasattr(e, "exp") else np.exp(e))


def config4(nx=2048, ny=256, ncases=256, nbasis=8):
    """Cantilever (Elas2DQ4, plane stress, nu = 0.3) on [0,4] x [0,1], nx x ny
    quadrilaterals, clamped at x = 0 (both DOFs); E_e lognormal with mean 1
    and sigma 0.5 from the 64-mode Gaussian field (correlation 0.1) at the
    element centroids.  Right-edge tractions: 8 basis loads (x and y
    components of Legendre profiles P0..P3 in y, constant per side) and
    ncases seeded N(0,1) combinations of them."""
    from . import api
    coord, code, elem = api.mesh2d(nx, ny, quad=True, nb_dof=2, box=(0.0, 4.0, 0.0, 1.0),
                                   codes=(1, 0, 0, 0))
    code[:, 1] = code[:, 0]  # clamp both DOFs of the left edge
    rng = SplitMix64(SEED_BASE + 4)
    g = fourier_field(rng, centroids(coord, elem))
    E = np.exp(0.5 * g - 0.125)
    # right-edge sides (node ids 1-based, nodes x-outer / y-inner)
    first = nx * (ny + 1) + 1
    sides = np.array([[first + j, first + j + 1] for j in range(ny)], np.int32)
    ymid = 0.5 * (coord[sides[:, 0] - 1, 1] + coord[sides[:, 1] - 1, 1])
    t = 2.0 * ymid - 1.0
    leg = [np.ones_like(t), t, 0.5 * (3 * t ** 2 - 1), 0.5 * (5 * t ** 3 - 3 * t)]
    basis = []
    for k in range(nbasis):
        tr = np.zeros((ny, 2))
        tr[:, 0 if k < 4 else 1] = leg[k % 4]
        basis.append(tr)
    alpha = rng.normal(ncases * nbasis).reshape(ncases, nbasis)
    return dict(coord=coord, code=code, elem=elem, coef=E, sides=sides, basis=basis,
                alpha=alpha, prm=(0.3, 1.0))
