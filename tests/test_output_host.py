"""Host side of the output path (src/simulation.jl, src/IO.jl): Julia's Float64 printing in
simdata.csv and the .pvd collections, the CSV/PVD writers, movingavg, and the VTP reader the GPU
tests use, checked on a file written by hand to the VTK XML specification."""
import os
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np
import pytest

from vtp import read_vtp


@pytest.fixture(scope="module")
def sim_mod(lv):
    from importlib import import_module

    return import_module(lv.__name__ + ".simulation")


@pytest.mark.parametrize("x,s", [(0.1, "0.1"), (1e-5, "1.0e-5"), (2.5e-4, "0.00025"), (0.0001, "0.0001"),
                                 (123456.0, "123456.0"), (1e6, "1.0e6"), (1234567.8, "1.2345678e6"), (-3.0, "-3.0"),
                                 (1.5e-12, "1.5e-12"), (100.0, "100.0"), (0.0, "0.0"), (-0.0, "-0.0"),
                                 (1.0 / 3.0, "0.3333333333333333"), (2.0 ** 60, "1.152921504606847e18"),
                                 (float("nan"), "NaN"), (float("inf"), "Inf")])
def test_julia_float_printing(sim_mod, x, s):
    """string(x) of Julia 1.10 for these values (Base.Ryu shortest round-trip digits)"""
    assert sim_mod.julia_float(x) == s
    if np.isfinite(x):
        assert float(s) == x  # round trip


def test_csv_and_pvd_writers(sim_mod, tmp_path):
    p = str(tmp_path / "simdata.csv")
    sim_mod.write_csv(p, {"time": [0.0, 0.0125], "E_err": [1e-12, -3.5e-13], "P_err": [0.1, 0.2]})
    lines = open(p).read().splitlines()
    assert lines[0] == "E_err,P_err,time"  # DataFrame(::Dict) sorts its columns by name
    assert lines[1] == "1.0e-12,0.1,0.0" and lines[2] == "-3.5e-13,0.2,0.0125"
    q = str(tmp_path / "cells.pvd")
    sim_mod.write_pvd(q, [(0.0, str(tmp_path / "cframe0.vtp")), (0.01, str(tmp_path / "cframe1.vtp"))])
    root = ET.parse(q).getroot()
    assert root.attrib["type"] == "Collection"
    ds = root.findall("Collection/DataSet")
    assert [(d.attrib["timestep"], d.attrib["file"]) for d in ds] == [("0.0", "cframe0.vtp"), ("0.01", "cframe1.vtp")]


def test_movingavg(lv):
    x = np.random.default_rng(0).standard_normal(17)
    for r in (0, 1, 3):
        y = lv.movingavg(list(x), r)
        ref = [x[max(i - r, 0):min(i + r, 16) + 1].mean() for i in range(17)]
        assert np.allclose(y, ref, rtol=0, atol=1e-15)


@pytest.mark.parametrize("compress", [False, True])
def test_vtp_reader_on_a_specification_example(tmp_path, compress):
    """a two-triangle PolyData file assembled byte by byte from the VTK XML specification"""
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=np.float64)
    conn = np.array([0, 1, 2, 1, 3, 2], dtype=np.int64)
    offs = np.array([3, 6], dtype=np.int64)
    P = np.array([0.5, -2.0])

    def block(a):
        b = a.tobytes()
        if not compress:
            return struct.pack("<Q", len(b)) + b
        c = zlib.compress(b)
        return struct.pack("<QQQQ", 1, len(b), len(b), len(c)) + c

    blobs = [block(pts), block(conn), block(offs), block(P)]
    offsets = np.cumsum([0] + [len(b) for b in blobs[:-1]])
    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    head = ('<?xml version="1.0"?><VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" '
            'header_type="UInt64"%s><PolyData><Piece NumberOfPoints="4" NumberOfPolys="2">'
            '<Points><DataArray type="Float64" NumberOfComponents="3" format="appended" offset="%d"/></Points>'
            '<Polys><DataArray type="Int64" Name="connectivity" format="appended" offset="%d"/>'
            '<DataArray type="Int64" Name="offsets" format="appended" offset="%d"/></Polys>'
            '<CellData><DataArray type="Float64" Name="P" format="appended" offset="%d"/></CellData>'
            '</Piece></PolyData><AppendedData encoding="raw">_' % ((comp,) + tuple(offsets)))
    f = tmp_path / "t.vtp"
    f.write_bytes(head.encode() + b"".join(blobs) + b"</AppendedData></VTKFile>")
    d = read_vtp(str(f))
    assert np.array_equal(d["points"], pts) and np.array_equal(d["cells"]["Polys"]["connectivity"], conn)
    assert np.array_equal(d["cells"]["Polys"]["offsets"], offs) and np.array_equal(d["celldata"]["P"], P)
