"""Seeded synthetic inputs: jittered lattices and the initial conditions of the benchmark flows.

Positions are a Cartesian lattice x_ij = xmin + ((i + 1/2) + d)*dr with d uniform in [-0.1, 0.1]
per coordinate, drawn from SplitMix64 (Steele, Lea & Flood 2014).  Particle p = j*nx + i takes
draws 2p (x jitter) and 2p+1 (y jitter); u = (top 53 bits) / 2^53.  The generator is vectorised
numpy on uint64 (wrap-around arithmetic), so the same seed gives the same bits on every machine.
"""
import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, start=0):
    """The draws start+1 .. start+count of SplitMix64 seeded with `seed` (uint64 array)."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(bits):
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def jittered_lattice(nx, ny, dr, xmin=(0.0, 0.0), jitter=0.1, seed=0x5EED0000):
    """SoA planes double[2*n] of an nx-by-ny lattice with +-jitter*dr uniform jitter."""
    n = nx * ny
    u = uniform01(splitmix64(seed, 2 * n)).reshape(n, 2)
    d = (2.0 * jitter) * u - jitter
    i = np.tile(np.arange(nx, dtype=np.float64), ny)
    j = np.repeat(np.arange(ny, dtype=np.float64), nx)
    x = np.empty(2 * n)
    x[:n] = xmin[0] + ((i + 0.5) + d[:, 0]) * dr
    x[n:] = xmin[1] + ((j + 0.5) + d[:, 1]) * dr
    return x


# ---- initial conditions: each returns a dict of SoA fields given positions and areas -----------

def taylor_green(x, area, rho0=1.0, mu=1e-3, c0=50.0, gamma=1.4):
    """u = sin(2 pi x) cos(2 pi y), v = -cos(2 pi x) sin(2 pi y), P = (cos 4 pi x + cos 4 pi y)/4,
    stiffened gas with P0 = rho0 c0^2 / gamma (the EOS constants of tests/taylorgreen.jl:17-19)."""
    n = x.size // 2
    X, Y = x[:n], x[n:]
    v = np.concatenate([np.sin(2 * math.pi * X) * np.cos(2 * math.pi * Y),
                        -np.cos(2 * math.pi * X) * np.sin(2 * math.pi * Y)])
    P = 0.25 * (np.cos(4 * math.pi * X) + np.cos(4 * math.pi * Y))
    rho = np.full(n, rho0)
    e = 0.5 * (v[:n] ** 2 + v[n:] ** 2) + P / (rho * (gamma - 1.0))
    return dict(v=v, rho=rho, mass=rho * area, P=P, e=e, mu=np.full(n, mu), P0=rho0 * c0 ** 2 / gamma,
                gamma=gamma)


def gresho(x, area, center=(0.5, 0.5), rho0=1.0):
    """Gresho vortex (examples/gresho.jl:46-62): speed 5r (r < 0.2), 2 - 5r (r < 0.4), 0 outside;
    incompressible (c2 = Inf)."""
    n = x.size // 2
    dx, dy = x[:n] - center[0], x[n:] - center[1]
    r = np.sqrt(dx * dx + dy * dy)
    with np.errstate(divide="ignore"):
        omega = np.where(r < 0.2, 5.0, np.where(r < 0.4, 2.0 / np.maximum(r, 1e-300) - 5.0, 0.0))
        P = np.where(r < 0.2, 12.5 * r ** 2,
                     np.where(r < 0.4, 4.0 + 4 * np.log(np.maximum(5 * r, 1e-300)) - 20.0 * r + 12.5 * r ** 2,
                              -2.0 + 4 * math.log(2)))
    v = np.concatenate([-omega * dy, omega * dx])
    rho = np.full(n, rho0)
    return dict(v=v, rho=rho, mass=rho * area, P=P, e=0.5 * (v[:n] ** 2 + v[n:] ** 2), c2=np.full(n, np.inf))


def rayleigh_taylor(x, area, rho_light=1.0, ratio=3.0, amp=0.05, g=1.0, gamma=1.4, mu=1e-3, P_top=10.0):
    """Heavy fluid (ratio*rho_light) above y = 1 + amp*cos(2 pi x) in [0,1]x[0,2], hydrostatic P."""
    n = x.size // 2
    X, Y = x[:n], x[n:]
    yi = 1.0 + amp * np.cos(2 * math.pi * X)
    heavy = Y > yi
    rho = np.where(heavy, ratio * rho_light, rho_light)
    P = np.where(heavy, P_top + ratio * rho_light * g * (2.0 - Y),
                 P_top + ratio * rho_light * g * (2.0 - yi) + rho_light * g * (yi - Y))
    v = np.zeros(2 * n)
    e = P / (rho * (gamma - 1.0))
    return dict(v=v, rho=rho, mass=rho * area, P=P, e=e, mu=np.full(n, mu), phase=heavy.astype(np.int32))


def cavity(x, area, rho0=1.0, mu=1e-3):
    """Lid-driven cavity (examples/cavity.jl:46-51): fluid at rest, c2 = Inf, lid u = 1 on BDARY_UP."""
    n = x.size // 2
    rho = np.full(n, rho0)
    return dict(v=np.zeros(2 * n), rho=rho, mass=rho * area, mu=np.full(n, mu), c2=np.full(n, np.inf))
