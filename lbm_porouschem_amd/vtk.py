"""VTK output without pyvista: XML StructuredGrid (.vts) files and a ParaView .pvd
collection, the formats the reference writes through pyvista
(LBM3D/_io.py:688-751 export_VTK_pyvista / get_data_pyvista; PVD list of
visualization_tool/PVD.py).

Point data follow the reference's names: solid, rho, velocity, and with the thermal
module rho_solid, Tf, Ts (physical K), dTf, dTs, solid_fluid_exchange_surface,
solid_fluid_exchange_coef; with chemistry every species S (FIX species as
"name(S)") and "d name" for fluid species.  Arrays are written as raw little-endian
binary in an appended data block, points at the cell centres in physical units.
"""
import os

import numpy as np

_VTK_TYPES = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64"}


def write_vts(path, shape, dx, point_data):
    nx, ny, nz = shape
    x, y, z = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dx, np.arange(nz) * dx, indexing="ij")
    pts = np.stack([x.ravel(order="F"), y.ravel(order="F"), z.ravel(order="F")], axis=1).astype(np.float64)
    blocks = []
    header = ['<?xml version="1.0"?>',
              '<VTKFile type="StructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
              f'<StructuredGrid WholeExtent="0 {nx - 1} 0 {ny - 1} 0 {nz - 1}">',
              f'<Piece Extent="0 {nx - 1} 0 {ny - 1} 0 {nz - 1}">', "<PointData>"]
    offset = 0

    def add(arr):
        nonlocal offset
        a = np.ascontiguousarray(arr)
        raw = a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
        blocks.append(np.uint64(len(raw)).tobytes() + raw)
        o = offset
        offset += 8 + len(raw)
        return o

    for name, arr in point_data.items():
        a = np.asarray(arr)
        if a.dtype not in _VTK_TYPES:
            a = a.astype(np.float64)
        ncomp = 1 if a.ndim == 1 else a.shape[1]
        o = add(a)
        header.append(f'<DataArray type="{_VTK_TYPES[a.dtype]}" Name="{name}" NumberOfComponents="{ncomp}" '
                      f'format="appended" offset="{o}"/>')
    header.append("</PointData>")
    o = add(pts)
    header += ["<Points>", f'<DataArray type="Float64" NumberOfComponents="3" format="appended" offset="{o}"/>',
               "</Points>", "</Piece>", "</StructuredGrid>", '<AppendedData encoding="raw">']
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n_").encode("ascii"))
        for b in blocks:
            f.write(b)
        f.write(b"\n</AppendedData>\n</VTKFile>\n")


def read_vts(path):
    """Point-data arrays of a file written by write_vts (for tests and tooling)."""
    with open(path, "rb") as f:
        data = f.read()
    head, rest = data.split(b'<AppendedData encoding="raw">', 1)
    blob = rest[rest.index(b"_") + 1:]
    out = {}
    import re
    for m in re.finditer(rb'<DataArray type="(\w+)" Name="([^"]+)" NumberOfComponents="(\d+)" format="appended" '
                         rb'offset="(\d+)"/>', head):
        typ, name, ncomp, off = m.group(1).decode(), m.group(2).decode(), int(m.group(3)), int(m.group(4))
        n = int(np.frombuffer(blob[off:off + 8], dtype="<u8")[0])
        dt = np.dtype("<f4") if typ == "Float32" else np.dtype("<f8")
        a = np.frombuffer(blob[off + 8:off + 8 + n], dtype=dt)
        out[name] = a.reshape(-1, ncomp) if ncomp > 1 else a
    return out


class PVD:
    """Time series collection (.pvd) of the exported files."""

    def __init__(self, path, name):
        self.exportPath = path
        self.name = name
        self.entries = []

    def addVTK(self, time, filename):
        self.entries.append((float(time), filename))

    def writePVD(self):
        lines = ['<?xml version="1.0"?>', '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">',
                 "<Collection>"]
        for t, fn in self.entries:
            lines.append(f'<DataSet timestep="{t!r}" group="" part="0" file="{fn}"/>')
        lines += ["</Collection>", "</VTKFile>"]
        with open(os.path.join(self.exportPath, self.name + ".pvd"), "w") as f:
            f.write("\n".join(lines) + "\n")
