"""Minimal VTK XML PolyData (.vtp) writer and reader for particle backups.

The reference writes backups with pyvista (``vpmModule.py:492-516``: points
plus point arrays Velocity, Strength, Radius, Group_ID, Viscosity,
Viscosity_t) and reads them back with ``pv.read`` (:466-490).  pyvista is not a
dependency here, so this module writes the same file layout itself: one
``<Piece>`` with ``NumberOfPoints`` points, one vertex cell per point and the
point arrays, in ``binary`` (base64) or ``ascii`` format.  The reader accepts
what this writer produces and what VTK/pyvista write for such files: ascii or
base64 inline arrays, UInt32 or UInt64 headers, optional zlib compression.
"""
from __future__ import annotations

import base64
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np

_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
              np.dtype(np.int64): "Int64", np.dtype(np.int32): "Int32",
              np.dtype(np.uint8): "UInt8"}
_NP_TYPES = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64, "Int32": np.int32,
             "UInt8": np.uint8, "UInt32": np.uint32, "UInt64": np.uint64, "Int8": np.int8,
             "Int16": np.int16, "UInt16": np.uint16}


def _data_array(parent, name, arr, fmt):
    arr = np.ascontiguousarray(arr)
    ncomp = 1 if arr.ndim == 1 else arr.shape[1]
    el = ET.SubElement(parent, "DataArray", type=_VTK_TYPES[arr.dtype], format=fmt)
    if name:
        el.set("Name", name)
    if ncomp > 1:
        el.set("NumberOfComponents", str(ncomp))
    flat = arr.reshape(-1)
    if fmt == "ascii":
        el.text = " ".join(repr(float(v)) if arr.dtype.kind == "f" else str(int(v)) for v in flat)
    else:
        raw = flat.astype(flat.dtype.newbyteorder("<")).tobytes()
        el.text = base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")
    return el


def write_vtp(path, points, point_data: dict, binary=True):
    """Writes points (N,3) and point arrays (N,) or (N,k) to a .vtp file."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    n = len(pts)
    fmt = "binary" if binary else "ascii"
    root = ET.Element("VTKFile", type="PolyData", version="1.0", byte_order="LittleEndian",
                      header_type="UInt64")
    poly = ET.SubElement(root, "PolyData")
    piece = ET.SubElement(poly, "Piece", NumberOfPoints=str(n), NumberOfVerts=str(n),
                          NumberOfLines="0", NumberOfStrips="0", NumberOfPolys="0")
    pd = ET.SubElement(piece, "PointData")
    for name, arr in point_data.items():
        _data_array(pd, name, np.asarray(arr), fmt)
    ET.SubElement(piece, "CellData")
    p = ET.SubElement(piece, "Points")
    _data_array(p, "Points", pts, fmt)
    verts = ET.SubElement(piece, "Verts")
    _data_array(verts, "connectivity", np.arange(n, dtype=np.int64), fmt)
    _data_array(verts, "offsets", np.arange(1, n + 1, dtype=np.int64), fmt)
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


def _decode(el, root, appended):
    dtype = np.dtype(_NP_TYPES[el.get("type")]).newbyteorder("<")
    ncomp = int(el.get("NumberOfComponents", "1"))
    fmt = el.get("format", "ascii")
    htype = np.uint64 if root.get("header_type", "UInt32") == "UInt64" else np.uint32
    hsize = np.dtype(htype).itemsize
    compressed = root.get("compressor", "") != ""
    if fmt == "ascii":
        vals = np.array((el.text or "").split(), dtype=dtype)
    else:
        if fmt == "appended":
            off = int(el.get("offset"))
            raw_all = appended
            if raw_all is None:
                raise ValueError("appended data section missing")
            blob = raw_all[off:]
            b64 = False
        else:
            blob = base64.b64decode("".join((el.text or "").split()))
            b64 = True
        if compressed:
            nb = int(np.frombuffer(blob[:hsize], dtype=htype)[0])
            head = np.frombuffer(blob[:hsize * (3 + nb)], dtype=htype)
            csizes = head[3:3 + nb]
            pos = hsize * (3 + nb)
            parts = []
            for cs in csizes:
                parts.append(zlib.decompress(blob[pos:pos + int(cs)]))
                pos += int(cs)
            data = b"".join(parts)
        else:
            size = int(np.frombuffer(blob[:hsize], dtype=htype)[0])
            data = blob[hsize:hsize + size]
        del b64
        vals = np.frombuffer(data, dtype=dtype)
    vals = vals.astype(dtype.newbyteorder("="))
    return vals.reshape(-1, ncomp) if ncomp > 1 else vals


def read_vtp(path):
    """Returns (points (N,3), {name: array}) from a .vtp file."""
    with open(path, "rb") as f:
        raw = f.read()
    appended = None
    tag = raw.find(b"<AppendedData")
    if tag >= 0:  # raw appended section: bytes after '_' up to the closing tag
        start = raw.index(b"_", raw.index(b">", tag)) + 1
        end = raw.rindex(b"</AppendedData>")
        appended = raw[start:end].rstrip(b"\n ")
        enc = raw[tag:raw.index(b">", tag)]
        if b'encoding="base64"' in enc:
            appended = base64.b64decode(b"".join(appended.split()))
        xml_text = raw[:tag] + b"</VTKFile>"
    else:
        xml_text = raw
    root = ET.fromstring(xml_text)
    piece = root.find("PolyData/Piece")
    if piece is None:
        raise ValueError(f"{path}: not a PolyData file")
    pts = _decode(piece.find("Points/DataArray"), root, appended).reshape(-1, 3)
    data = {}
    pd = piece.find("PointData")
    if pd is not None:
        for el in pd.findall("DataArray"):
            data[el.get("Name")] = _decode(el, root, appended)
    return pts.astype(np.float64), data
