"""Shared set-ups of the tests: the reference test's Taylor-Green case and small seeded flows."""
import math

import numpy as np


def populate_hex(xmin, xmax, dr):
    """populate_hex! positions (src/populate.jl:156-179): rows j, offset (j%2)/2, spacing
    a = (4/3)^(1/4) dr, b = (3/4)^(1/4) dr, kept if inside the closed boundary rectangle."""
    a = (4.0 / 3.0) ** 0.25 * dr
    b = (3.0 / 4.0) ** 0.25 * dr
    i_min = int(math.floor(xmin[0] / a)) - 1
    j_min = int(math.floor(xmin[1] / b))
    i_max = int(math.ceil(xmax[0] / a))
    j_max = int(math.ceil(xmax[1] / b))
    pts = []
    for i in range(i_min, i_max + 1):
        for j in range(j_min, j_max + 1):
            jr = int(math.fmod(j, 2))
            x1 = (i + jr / 2) * a
            x2 = j * b
            if xmin[0] <= x1 <= xmax[0] and xmin[1] <= x2 <= xmax[1]:
                pts.append((x1, x2))
    p = np.array(pts)
    return np.concatenate([p[:, 0], p[:, 1]])


# tests/taylorgreen.jl:9-20
TG_RE = 400.0
TG_N = 80
TG_DR = 1.0 / TG_N
TG_DT = 0.1 * TG_DR
TG_C0 = 50.0
TG_GAMMA = 1.4
TG_P0 = 1.0 * TG_C0 ** 2 / TG_GAMMA
TG_TEND = 0.1


def tg_vmax(t):
    return math.exp(-8.0 * math.pi ** 2 * t / TG_RE)


def tg_v_exact(x, t):
    """v_exact of tests/taylorgreen.jl:34-38 (cos-sin form), x as SoA planes"""
    n = x.size // 2
    X, Y = x[:n], x[n:]
    u0 = np.cos(2 * math.pi * X) * np.sin(2 * math.pi * Y)
    v0 = -np.sin(2 * math.pi * X) * np.cos(2 * math.pi * Y)
    return tg_vmax(t) * np.concatenate([u0, v0])


def tg_P_exact(x, t):
    n = x.size // 2
    X, Y = x[:n], x[n:]
    return 0.5 * (tg_vmax(t) ** 2) * (np.sin(2 * math.pi * X) ** 2 + np.sin(2 * math.pi * Y) ** 2 - 1.0)


def tg_ic(x, area):
    """ic! of tests/taylorgreen.jl:25-32"""
    n = x.size // 2
    v = tg_v_exact(x, 0.0)
    rho = np.ones(n)
    P = tg_P_exact(x, 0.0)
    e = 0.5 * (v[:n] ** 2 + v[n:] ** 2) + P / (rho * (TG_GAMMA - 1.0))
    return dict(v=v, rho=rho, mass=rho * area, P=P, e=e, mu=np.full(n, 1.0 / TG_RE))


def tg_errors(x, v, P, area, mass, e, t):
    """postproc! of tests/taylorgreen.jl:73-99 without the E0 subtraction"""
    p_avg = np.sum(area * P)
    n = x.size // 2
    ve = tg_v_exact(x, t)
    E = np.sum(mass * e)
    v_err = math.sqrt(np.sum(area * ((v[:n] - ve[:n]) ** 2 + (v[n:] - ve[n:]) ** 2)))
    P_err = math.sqrt(np.sum(area * (P - p_avg - tg_P_exact(x, t)) ** 2))
    return E, v_err, P_err


def tg_frame_schedule():
    """The run! loop (src/simulation.jl:59-89) with nframes = 20: yields (k, t, is_frame)."""
    nframes = 20
    k_frame = max(int(round(TG_TEND / (nframes * TG_DT))), 1)
    t = 0.0
    k = 0
    while t < TG_TEND:
        k += 1
        yield k, t, (k % k_frame == 0)
        t += TG_DT
