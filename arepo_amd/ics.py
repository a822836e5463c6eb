"""Seeded synthetic particle sets for the gravity workloads.

All generators use numpy's Philox bit generator so that a (config, seed)
pair always yields the same particles.  Units: G = 1; periodic sets live in
the unit box.  Each function returns (pos[N,3] float64, mass[N] float64, eps)
with eps the Plummer-equivalent softening.
"""
from __future__ import annotations

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def _isotropic(r: np.ndarray, g: np.random.Generator) -> np.ndarray:
    mu = g.uniform(-1.0, 1.0, len(r))
    phi = g.uniform(0.0, 2 * np.pi, len(r))
    s = np.sqrt(1 - mu * mu)
    return np.stack([r * s * np.cos(phi), r * s * np.sin(phi), r * mu], axis=1)


def hernquist_radii(n: int, a: float, g: np.random.Generator, rmax_over_a: float = 30.0) -> np.ndarray:
    """Radii from the Hernquist mass profile M(<r) = M r^2/(r+a)^2 truncated at rmax."""
    smax = rmax_over_a / (1.0 + rmax_over_a)
    s = np.sqrt(g.random(n)) * smax
    return a * s / (1.0 - s)


def uniform_box(n: int = 32768, seed: int = 1):
    """Uniform random particles in the periodic unit box, equal masses, eps = (1/N^(1/3))/40."""
    g = rng(seed)
    pos = g.random((n, 3))
    mass = np.full(n, 1.0 / n)
    eps = (1.0 / round(n ** (1.0 / 3.0))) / 40.0
    return pos, mass, eps


def clustered_box(n: int = 128 ** 3, nhalo: int = 256, seed: int = 2, mass_range: float = 1000.0, slope: float = 1.9):
    """Half uniform, half in `nhalo` Hernquist spheres (a ~ U(0.005, 0.02), dN/dM ~ M^-slope on
    [Mmin, mass_range*Mmin], normalised to half the mass), wrapped into the periodic unit box."""
    g = rng(seed)
    n_uni = n // 2
    n_cl = n - n_uni
    # power-law halo masses by inverse CDF of dN/dM ~ M^-slope on [1, mass_range]
    u = g.random(nhalo)
    k = 1.0 - slope
    m = (1.0 + u * (mass_range ** k - 1.0)) ** (1.0 / k)
    counts = np.floor(m / m.sum() * n_cl).astype(np.int64)
    short = n_cl - counts.sum()
    order = np.argsort(-(m / m.sum() * n_cl - counts))
    counts[order[:short]] += 1
    centres = g.random((nhalo, 3))
    scales = g.uniform(0.005, 0.02, nhalo)
    parts = [g.random((n_uni, 3))]
    for c, a, cnt in zip(centres, scales, counts):
        if cnt == 0:
            continue
        r = hernquist_radii(int(cnt), a, g)
        parts.append(c + _isotropic(r, g))
    pos = np.concatenate(parts) % 1.0
    pos[pos >= 1.0] = 0.0
    mass = np.full(n, 1.0 / n)
    eps = (1.0 / round(n ** (1.0 / 3.0))) / 40.0
    return pos, mass, eps


def halo_disk(n: int = 1 << 20, seed: int = 3, eps: float = 0.01):
    """Isolated Hernquist halo (M=1, a=1, r<30a) plus exponential disk (M=0.1, Rd=0.2, sech^2 z0=0.02),
    equal particle masses (total 1.1)."""
    g = rng(seed)
    n_halo = int(round(n / 1.1))
    n_disk = n - n_halo
    r = hernquist_radii(n_halo, 1.0, g)
    halo = _isotropic(r, g)
    R = 0.2 * g.gamma(2.0, 1.0, n_disk)
    ph = g.uniform(0, 2 * np.pi, n_disk)
    z = 0.02 * np.arctanh(np.clip(2 * g.random(n_disk) - 1, -1 + 1e-15, 1 - 1e-15))
    disk = np.stack([R * np.cos(ph), R * np.sin(ph), z], axis=1)
    pos = np.concatenate([halo, disk])
    mass = np.full(n, 1.1 / n)
    return pos, mass, eps


def nfw_like_group(n: int, conc: float, g: np.random.Generator, rvir: float = 1.0):
    """NFW profile truncated at the virial radius: M(<x) ~ ln(1+x) - x/(1+x), x = r/rs, rs = rvir/conc."""
    rs = rvir / conc
    xmax = conc
    mx = np.log1p(xmax) - xmax / (1 + xmax)
    u = g.random(n) * mx
    # invert by bisection on x in [0, xmax]
    lo = np.zeros(n)
    hi = np.full(n, xmax)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        f = np.log1p(mid) - mid / (1 + mid)
        lo = np.where(f < u, mid, lo)
        hi = np.where(f < u, hi, mid)
    return _isotropic(0.5 * (lo + hi) * rs, g)


def half_mass_radius(pos: np.ndarray) -> float:
    r = np.sqrt((pos * pos).sum(1))
    return float(np.median(r))


def _nfw_radii(n: int, conc: float, rvir: float, g: np.random.Generator) -> np.ndarray:
    """NFW radii truncated at rvir, by interpolation in the tabulated inverse of M(<r)."""
    x = np.concatenate([[0.0], np.geomspace(1e-5, conc, 2048)])
    mx = np.log1p(x) - x / (1 + x)
    return np.interp(g.random(n) * mx[-1], mx, x) * (rvir / conc)


def unbinding_batch(nsets: int = 4096, nmin: int = 32, nmax: int = 200000, seed: int = 4, mpart: float = 1e-4):
    """Independent isolated groups for batched potential solves.

    Particle counts are log-uniform on [nmin, nmax]; each group is NFW-like or
    Hernquist-like (alternating), concentration U(4, 15), virial radius growing as
    N^(1/3) at fixed particle mass; softening = half-mass radius / 50.  Returns
    (offsets[nsets+1] int64, pos_rel[N,3] float32 offsets from each group's
    centre, mass[N] float64, eps[nsets], centres[nsets,3]).
    """
    g = rng(seed)
    sizes = np.floor(np.exp(g.uniform(np.log(nmin), np.log(nmax + 1), nsets))).astype(np.int64)
    sizes = np.clip(sizes, nmin, nmax)
    conc = g.uniform(4.0, 15.0, nsets)
    centres = g.uniform(0.0, 100.0, (nsets, 3))
    off = np.zeros(nsets + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    pos = np.empty((int(off[-1]), 3), np.float32)
    eps = np.empty(nsets)
    for s in range(nsets):
        n = int(sizes[s])
        rvir = 0.05 * (n / 1000.0) ** (1.0 / 3.0)
        if s % 2 == 0:
            r = _nfw_radii(n, conc[s], rvir, g)
        else:
            r = hernquist_radii(n, rvir / conc[s], g, rmax_over_a=conc[s])
        p = _isotropic(r, g)
        pos[off[s]:off[s + 1]] = p
        eps[s] = float(np.median(r)) / 50.0
    mass = np.full(int(off[-1]), mpart)
    return off, pos, mass, eps, centres


def _shell_potential(r: np.ndarray, m: float) -> np.ndarray:
    """Potential (G = 1) of equal masses m at radii r from spherical shells: -(M(<r)/r + sum_{r_j > r} m/r_j)."""
    order = np.argsort(r)
    rs = np.maximum(r[order], 1e-12)
    inner = m * np.arange(len(rs)) / rs
    outer = m * (np.cumsum((1.0 / rs)[::-1])[::-1] - 1.0 / rs)
    phi = np.empty_like(r)
    phi[order] = -(inner + outer)
    return phi


def unbinding_candidates(ngroups: int = 4096, nmin: int = 32, nmax: int = 200000, seed: int = 6, mpart: float = 1e-4,
                         f_max: float = 1.3, interlopers: float = 0.05):
    """Candidate groups for SUBFIND-style unbinding: the groups of `unbinding_batch` with velocities.

    Each member moves isotropically at f times its escape speed from the group's
    own spherical mass profile, f ~ U(0, f_max), so about (f_max - 1)/f_max of
    the members start unbound; a fraction `interlopers` of each group is
    replaced by particles spread uniformly over the group's sphere moving at
    2-4 escape speeds; every group has a bulk velocity of about its half-mass
    escape speed.  Returns (offsets[ngroups+1] int64, pos[N,3] float64 offsets
    from each group's centre, vel[N,3] float64, mass[N], eps[ngroups]).
    """
    off, pos32, mass, eps, _ = unbinding_batch(ngroups, nmin, nmax, seed, mpart)
    g = rng(seed + 1000)
    pos = pos32.astype(np.float64)
    vel = np.empty_like(pos)
    for s in range(ngroups):
        a, b = int(off[s]), int(off[s + 1])
        n = b - a
        if n == 0:
            continue
        p = pos[a:b]
        r = np.sqrt((p * p).sum(1))
        ni = int(interlopers * n)
        if ni > 0:
            rmax = float(r.max())
            k = g.choice(n, ni, replace=False)
            p[k] = _isotropic(rmax * g.random(ni) ** (1.0 / 3.0), g)
            r = np.sqrt((p * p).sum(1))
        vesc = np.sqrt(2.0 * -_shell_potential(r, mpart))
        f = g.uniform(0.0, f_max, n)
        if ni > 0:
            f[k] = g.uniform(2.0, 4.0, ni)
        d = g.normal(size=(n, 3))
        d /= np.maximum(np.linalg.norm(d, axis=1), 1e-300)[:, None]
        bulk = g.normal(size=3) * float(np.median(vesc)) / np.sqrt(3.0)
        vel[a:b] = d * (f * vesc)[:, None] + bulk
    return off, pos, vel, mass, eps
