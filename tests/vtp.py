"""Reader of the VTK XML PolyData files the library writes (appended raw data, UInt64 headers,
optionally one zlib block per array), for the output tests.  Follows the VTK XML file format
specification, independently of the writer."""
import xml.etree.ElementTree as ET
import zlib

import numpy as np

_TYPES = {"Float64": np.float64, "Int64": np.int64, "Float32": np.float32, "Int32": np.int32}


def read_vtp(path):
    raw = open(path, "rb").read()
    mark = raw.index(b"<AppendedData")
    start = raw.index(b"_", mark) + 1
    end = raw.rindex(b"</AppendedData>")
    head = raw[:start - 1].decode() + "</AppendedData></VTKFile>"
    root = ET.fromstring(head)
    assert root.attrib["type"] == "PolyData" and root.attrib["header_type"] == "UInt64"
    compressed = root.attrib.get("compressor") == "vtkZLibDataCompressor"
    blob = raw[start:end]

    def array(da):
        off = int(da.attrib["offset"])
        if compressed:
            nb, usize, psize = np.frombuffer(blob, np.uint64, 3, off)
            csizes = np.frombuffer(blob, np.uint64, int(nb), off + 24)
            p = off + 24 + 8 * int(nb)
            data = b""
            for c in csizes:
                data += zlib.decompress(blob[p:p + int(c)])
                p += int(c)
            assert len(data) == (int(usize) * (int(nb) - 1) + (int(psize) if psize else int(usize)) if nb else 0)
        else:
            (nbytes,) = np.frombuffer(blob, np.uint64, 1, off)
            data = blob[off + 8:off + 8 + int(nbytes)]
        a = np.frombuffer(data, _TYPES[da.attrib["type"]])
        nc = int(da.attrib.get("NumberOfComponents", "1"))
        return a.reshape(-1, nc) if nc > 1 else a

    piece = root.find("PolyData/Piece")
    out = {"piece": dict(piece.attrib), "compressed": compressed,
           "points": array(piece.find("Points/DataArray")), "cells": {}, "celldata": {}}
    for kind in ("Polys", "Verts"):
        node = piece.find(kind)
        if node is not None:
            out["cells"][kind] = {da.attrib["Name"]: array(da) for da in node.findall("DataArray")}
    cd = piece.find("CellData")
    if cd is not None:
        for da in cd.findall("DataArray"):
            out["celldata"][da.attrib["Name"]] = array(da)
    return out
