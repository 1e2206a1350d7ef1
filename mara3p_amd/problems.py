"""Seeded synthetic initial conditions for the benchmark configurations.

All randomness comes from a counter-based generator keyed by (seed, global
zone index), so the initial state does not depend on the block decomposition
or on the order of evaluation.  Arrays are returned in the device layout
[nblk, nc, n, n] (2D) or [nblk, nc, n] (1D), conserved variables.
"""
import math

import numpy as np

from .mesh import UniformMesh2D

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def uniform01(seed, index, stream=0):
    """Uniform doubles in (0, 1) keyed by (seed, stream, index)."""
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream))
        z = _splitmix64(np.asarray(index, dtype=np.uint64) ^ key)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed, index, stream=0):
    """Standard normals by Box-Muller from two counter streams."""
    u1 = uniform01(seed, index, 2 * stream)
    u2 = uniform01(seed, index, 2 * stream + 1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def _global_index_2d(mesh):
    """Global zone index (row-major over the whole level) for every zone, [nblk, n, n]."""
    n, nb = mesh.n, mesh.nb
    ij = mesh.block_coordinates()
    i = ij[:, 0, None, None] * n + np.arange(n)[None, :, None]
    j = ij[:, 1, None, None] * n + np.arange(n)[None, None, :]
    return (i * (nb * n) + j).astype(np.uint64)


def cartesian_cell_centres(mesh, domain_size=1.0):
    """mesh::cartesian_2d::geometry_t cell coordinates (src/mesh_geometry.cpp:114-146):
    block extents on [-L/2, L/2], linspace vertices, adjacent means.  [nblk, n, n] each."""
    n = mesh.n
    dl = float(1 << mesh.depth)
    ij = mesh.block_coordinates().astype(np.float64)
    x0 = domain_size * (-0.5 + 1.0 * ij[:, 0] / dl)
    x1 = domain_size * (-0.5 + 1.0 * (ij[:, 0] + 1) / dl)
    y0 = domain_size * (-0.5 + 1.0 * ij[:, 1] / dl)
    y1 = domain_size * (-0.5 + 1.0 * (ij[:, 1] + 1) / dl)
    k = np.arange(n + 1, dtype=np.float64)
    xv = x0[:, None] + (x1 - x0)[:, None] * k[None, :] / float(n)
    yv = y0[:, None] + (y1 - y0)[:, None] * k[None, :] / float(n)
    xc = 0.5 * (xv[:, :-1] + xv[:, 1:])
    yc = 0.5 * (yv[:, :-1] + yv[:, 1:])
    return np.broadcast_to(xc[:, :, None], (mesh.nblk, n, n)), np.broadcast_to(yc[:, None, :], (mesh.nblk, n, n))


def euler_conserved(rho, vx, vy, vz, p, gamma):
    """euler::conserved_density (src/physics_euler.hpp:175-183), vectorised, stacked on axis 1."""
    ls = ((0.0 + vx * vx) + vy * vy) + vz * vz
    return np.stack([rho, rho * vx, rho * vy, rho * vz, rho * ls * 0.5 + p / (gamma - 1.0)], axis=1)


def kelvin_helmholtz(depth=4, block_size=256, seed=1, gamma=5.0 / 3.0):
    """2D Kelvin-Helmholtz on [-1/2, 1/2]^2, periodic: rho = 2 in |y| < 1/4 else 1,
    vx = +-1/2, p = 2.5, vy = 0.01 sin(4 pi x) + U(-1e-3, 1e-3)."""
    mesh = UniformMesh2D(depth, block_size)
    x, y = cartesian_cell_centres(mesh)
    inner = np.abs(y) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5)
    noise = 2e-3 * uniform01(seed, _global_index_2d(mesh)) - 1e-3
    vy = 0.01 * np.sin(4.0 * math.pi * x) + noise
    vz = np.zeros_like(rho)
    p = np.full_like(rho, 2.5)
    return mesh, euler_conserved(rho, vx, vy, vz, p, gamma)


def minidisk_cell_centres(data, mesh):
    """problems/minidisk_scheme.cpp:101-116,153-165: [nblk, n, n] x and y."""
    n = mesh.n
    dl = float(1 << mesh.depth)
    R = data.domain_radius
    ij = mesh.block_coordinates().astype(np.int64)

    def verts(b):
        x0 = R * (-1 + (2 * b) / dl)
        x1 = R * (-1 + (2 * (b + 1)) / dl)
        k = np.arange(n + 1, dtype=np.float64)
        return x0[:, None] + (x1 - x0)[:, None] * k[None, :] / float(n)

    xv, yv = verts(ij[:, 0]), verts(ij[:, 1])
    xc = 0.5 * (xv[:, :-1] + xv[:, 1:])
    yc = 0.5 * (yv[:, :-1] + yv[:, 1:])
    return np.broadcast_to(xc[:, :, None], (mesh.nblk, n, n)), np.broadcast_to(yc[:, None, :], (mesh.nblk, n, n))


def minidisk(data, seed=1, noise=0.01):
    """Locally isothermal binary minidisk: Keplerian initial_primitive
    (problems/minidisk_scheme.cpp:87-99) with surface density
    Sigma = exp(noise * z), z ~ N(0, 1) per zone (log-normal, median 1, no
    mean correction)."""
    mesh = UniformMesh2D(data.depth, data.block_size)
    x, y = minidisk_cell_centres(data, mesh)
    rs = data.softening_length
    r2 = x * x + y * y
    ph = -1.0 / np.sqrt(r2 + rs ** 2.0)
    r0 = np.sqrt(r2)
    vp = np.sqrt(-ph) - data.omega_frame * r0
    vx = vp * (-y / r0)
    vy = vp * (x / r0)
    sigma = np.exp(noise * normal(seed, _global_index_2d(mesh))) if noise else np.ones_like(x)
    u = np.stack([sigma, sigma * vx, sigma * vy], axis=1)
    return mesh, np.ascontiguousarray(u)


def minidisk_dt(data, time):
    """problems/minidisk.cpp:202-206: dt = dx_min / sqrt(G 0.5 / r_soft) * cfl * ramp(t)."""
    vmax = math.sqrt(1.0 * 0.5 / data.softening_length)
    cfl = data.cfl * min(1.0, 0.01 + time / 0.05)
    dx = 2.0 * data.domain_radius / float(data.block_size) / float(1 << data.depth)
    return dx / vmax * cfl


def sod(nblk=16, block_size=64, gamma=1.4):
    """Sod shock tube on [0, 1] (left (1, 0, 1), right (0.125, 0, 0.1)), outflow.
    The reference 1D geometry spans [-1/2, 1/2] (src/mesh_geometry.cpp:72-82);
    the interface is its centre."""
    n = block_size
    N = nblk * n
    xc = (np.arange(N) + 0.5) / N - 0.5
    left = xc < 0.0
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros(N)
    u = euler_conserved(rho, z, z, z, p, gamma)        # [N, 5]
    return np.ascontiguousarray(u.reshape(nblk, n, 5).transpose(0, 2, 1))


def sedov_amr(tree, block_size, energy=1.0, radius=0.01, rho=1.0, p=1e-5, gamma=5.0 / 3.0):
    """2D Sedov blast on a quad-tree of leaves over [-1/2, 1/2]^2: rho = 1,
    v = 0, p = 1e-5, plus total energy `energy` deposited uniformly (per unit
    area) in the zones whose centres lie within `radius` of the origin.
    Returns conserved [nleaf, 5, n, n]."""
    n = block_size
    lev = tree.levels().astype(np.float64)
    ij = tree.coordinates().astype(np.float64)
    k = (np.arange(n) + 0.5)[None, :]
    dl = 1.0 / (n * 2.0 ** lev)
    xc = -0.5 + ij[:, 0, None] / 2.0 ** lev[:, None] + k * dl[:, None]
    yc = -0.5 + ij[:, 1, None] / 2.0 ** lev[:, None] + k * dl[:, None]
    x = np.broadcast_to(xc[:, :, None], (len(lev), n, n))
    y = np.broadcast_to(yc[:, None, :], (len(lev), n, n))
    inside = x * x + y * y < radius * radius
    area = np.broadcast_to((dl * dl)[:, None, None], inside.shape)
    e_density = energy / float(np.sum(area[inside]))
    z = np.zeros(inside.shape)
    pres = np.where(inside, (gamma - 1.0) * e_density, p)
    return np.ascontiguousarray(euler_conserved(np.full(inside.shape, rho), z, z, z, pres, gamma))


def riemann_ensemble(count=512, block_size=64, lmin=512, lmax=8192, seed=1, gamma=1.4):
    """An ensemble of `count` 1D Riemann problems on [0, 1], each split at its
    centre.  Lengths are drawn uniformly from the multiples of `block_size`
    in [lmin, lmax]; left / right rho and p are log-uniform in [0.1, 10];
    v = 0.  Returns (EnsembleMesh1D, conserved [nblk, 5, n])."""
    from .mesh import EnsembleMesh1D
    n = block_size
    k = np.arange(count)
    choices = (lmax - lmin) // n + 1
    lengths = lmin + n * np.minimum((uniform01(seed, k, 0) * choices).astype(np.int64), choices - 1)
    mesh = EnsembleMesh1D(lengths, n)
    logu = lambda s: 10.0 ** (2.0 * uniform01(seed, k, s) - 1.0)
    rl, rr, pl, pr = logu(1), logu(2), logu(3), logu(4)
    u = np.empty((mesh.nblk, 5, n))
    for q in range(count):
        L = int(lengths[q])
        left = (np.arange(L) + 0.5) / L < 0.5
        rho = np.where(left, rl[q], rr[q])
        p = np.where(left, pl[q], pr[q])
        z = np.zeros(L)
        uq = euler_conserved(rho, z, z, z, p, gamma)          # [L, 5]
        u[mesh.problem_slice(q)] = uq.reshape(L // n, n, 5).transpose(0, 2, 1)
    return mesh, u


def orszag_tang_3d(N, seed=1, perturbation=1e-3, gamma=5.0 / 3.0):
    """3D Orszag-Tang-style state on the periodic unit cube with N^3 cells:
    rho = 25/(36 pi), p = 5/(12 pi), v = (-sin 2 pi y, sin 2 pi x, vz),
    B = (-sin 2 pi y, sin 4 pi x, 0) / sqrt(4 pi), vz = perturbation * U(-1, 1)
    seeded per cell.  Face fields come from the vector potential
    A_z = (cos(2 pi y) / (2 pi) + cos(4 pi x) / (4 pi)) / sqrt(4 pi) line-integrated
    on the z edges and differenced (as construct_conserved,
    src/scheme_mhd_v1.cpp:305-342), so div B = 0 to round-off.
    Returns (uc [5][N][N][N], bf1 [N+1][N][N], bf2 [N][N+1][N], bf3 [N][N][N+1])."""
    dl = 1.0 / N
    xv = np.arange(N + 1) * dl
    xc = (np.arange(N) + 0.5) * dl
    s4 = math.sqrt(4.0 * math.pi)

    def Az(x, y):
        return (np.cos(2 * math.pi * y) / (2 * math.pi) + np.cos(4 * math.pi * x) / (4 * math.pi)) / s4
    # x faces at (xv_i, y cell j): B1 = (Az(x_i, y_{j+1}) - Az(x_i, y_j)) / dl
    a = Az(xv[:, None], xv[None, :])                       # [N+1][N+1] at (x vertex, y vertex)
    b1 = (a[:, 1:] - a[:, :-1]) / dl                        # [N+1][N]
    b2 = -(a[1:, :] - a[:-1, :]) / dl                       # [N][N+1]
    bf1 = np.ascontiguousarray(np.broadcast_to(b1[:, :, None], (N + 1, N, N)))
    bf2 = np.ascontiguousarray(np.broadcast_to(b2[:, :, None], (N, N + 1, N)))
    bf3 = np.zeros((N, N, N + 1))
    bx = 0.5 * (bf1[1:] + bf1[:-1])
    by = 0.5 * (bf2[:, 1:] + bf2[:, :-1])
    bz = 0.5 * (bf3[:, :, 1:] + bf3[:, :, :-1])
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    rho = np.full((N, N, N), 25.0 / (36.0 * math.pi))
    p = np.full((N, N, N), 5.0 / (12.0 * math.pi))
    vx = np.broadcast_to(-np.sin(2 * math.pi * Y)[:, :, None], (N, N, N))
    vy = np.broadcast_to(np.sin(2 * math.pi * X)[:, :, None], (N, N, N))
    idx = np.arange(N * N * N, dtype=np.uint64).reshape(N, N, N)
    vz = perturbation * (2.0 * uniform01(seed, idx) - 1.0)
    v2 = vx * vx + vy * vy + vz * vz
    E = 0.5 * rho * v2 + p / (gamma - 1.0) + 0.5 * (bx * bx + by * by + bz * bz)
    uc = np.stack([rho, rho * vx, rho * vy, rho * vz, E])
    return np.ascontiguousarray(uc), bf1, bf2, bf3
