"""Time marching with frame output: run!, SimulationWorkspace and movingavg (src/simulation.jl).

run() keeps the reference's cadence and file set:
  * one frame every round(t_end / (nframes * dt)) steps;
  * cframe<k>.vtp (export_grid) and pframe<k>.vtp (export_points);
  * the ParaView collections cells.pvd and points.pvd;
  * simdata.csv with the time and the requested global variables.

The columns of simdata.csv are sorted by name, as DataFrame(::Dict) orders them, and numbers are
printed as Julia prints a Float64 (shortest round-trip digits; scientific outside [1e-4, 1e6)).
The frames themselves are written by the library (lv_export_grid / lv_export_points).
"""
import math
import os
from decimal import Decimal

from . import export_grid, export_points


class SimulationWorkspace:
    """SimulationWorkspace (src/simulation.jl:1-10): holds `grid` and whatever else a step needs;
    run() calls step(sim, t) and, before each frame, postproc(sim, t)."""

    grid = None


def donothing(sim, t):
    return None


def julia_float(x):
    """string(x::Float64) in Julia: shortest round-trip digits, plain notation for 1e-4 <= |x| < 1e6,
    otherwise d.ddde<exp> (at least one digit after the point)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign, digits, exp = Decimal(repr(x)).as_tuple()
    ds = "".join(str(d) for d in digits).rstrip("0") or "0"
    e10 = len(digits) + exp - 1  # exponent of the leading digit
    s = "-" if sign else ""
    if -5 < e10 < 6:
        if e10 >= 0:
            ip = ds[:e10 + 1].ljust(e10 + 1, "0")
            fp = ds[e10 + 1:] or "0"
        else:
            ip = "0"
            fp = "0" * (-e10 - 1) + ds
        return s + ip + "." + fp
    return s + ds[0] + "." + (ds[1:] or "0") + "e" + str(e10)


def write_pvd(path, frames):
    """paraview_collection(...) + vtk_save: a Collection of (time, file) datasets, files relative to
    the collection's directory"""
    base = os.path.dirname(os.path.abspath(path))
    lines = ['<?xml version="1.0" encoding="utf-8"?>',
             '<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">', "  <Collection>"]
    for t, f in frames:
        rel = os.path.relpath(os.path.abspath(f), base)
        lines.append('    <DataSet timestep="%s" part="0" file="%s"/>' % (julia_float(t), rel))
    lines += ["  </Collection>", "</VTKFile>"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_csv(path, columns):
    """CSV.write(path, DataFrame(columns)) for a dict of equal-length numeric columns"""
    names = sorted(columns)
    n = len(columns[names[0]]) if names else 0
    with open(path, "w") as fh:
        fh.write(",".join(names) + "\n")
        for i in range(n):
            fh.write(",".join(julia_float(columns[k][i]) for k in names) + "\n")


def run(sim, dt, t_end, step, nframes=100, path="results", save_points=False, save_grid=True, save_csv=True,
        vtp_vars=(), csv_vars=(), postproc=donothing):
    """run!(sim, dt, t_end, step!; nframes, path, save_points, save_grid, save_csv, vtp_vars,
    csv_vars, postproc!), src/simulation.jl:35-94"""
    if not os.path.isdir(path):
        os.makedirs(path)
        print("created a new path: %s" % path)
    csv_data = {"time": []}
    for var in csv_vars:
        if not isinstance(var, str):
            raise TypeError("csv_vars must be Symbols")
        if var == "time":
            raise ValueError("csv_vars cannot be :time")
        csv_data[var] = []
    k = 0
    nframe = 0
    k_frame = round(t_end / (nframes * dt)) if nframes > 0 else 2 ** 62  # round half to even, as round(Int, .)
    k_frame = max(k_frame, 1)
    cells, points = [], []
    t = 0.0
    while t < t_end:
        k += 1
        step(sim, t)
        if k % k_frame == 0:
            postproc(sim, t)
            csv_data["time"].append(t)
            for var in csv_vars:
                csv_data[var].append(getattr(sim, var))
            if save_grid:
                cells.append((t, export_grid(sim.grid, os.path.join(path, "cframe%d.vtp" % nframe), *vtp_vars)))
            if save_points:
                points.append((t, export_points(sim.grid, os.path.join(path, "pframe%d.vtp" % nframe), *vtp_vars)))
            nframe += 1
        t += dt
    if save_grid:
        write_pvd(os.path.join(path, "cells.pvd"), cells)
    if save_points:
        write_pvd(os.path.join(path, "points.pvd"), points)
    if save_csv:
        write_csv(path + "/simdata.csv", csv_data)


def movingavg(x, radius=1):
    """movingavg(x, radius), src/simulation.jl:96-110: the mean over the window i-radius..i+radius
    clipped to the ends"""
    n = len(x)
    y = [0.0] * n
    for i in range(n):
        lo, hi = max(i - radius, 0), min(i + radius, n - 1)
        s = 0.0
        for j in range(lo, hi + 1):
            s += x[j]
        y[i] = s / (hi - lo + 1)
    return y
