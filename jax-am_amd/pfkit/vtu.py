"""Minimal VTK XML UnstructuredGrid writer for the regular hex mesh (no meshio).

The reference writes its solution files with meshio (allen_cahn.py:199-210) on the
Neper hex mesh; this writes the same cell layout (C order z, y, x) as hexahedra on the
(Nx+1)(Ny+1)(Nz+1) grid points, with base64-encoded binary arrays.
"""
import base64
import struct

import numpy as np

_VTK_TYPES = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64", np.dtype(np.int32): "Int32",
              np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8"}


def _b64(a):
    a = np.ascontiguousarray(a)
    raw = a.tobytes()
    return base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")


def _array(name, a, ncomp=1):
    return (f'<DataArray type="{_VTK_TYPES[a.dtype]}" Name="{name}" NumberOfComponents="{ncomp}" '
            f'format="binary">{_b64(a)}</DataArray>')


def write_hex_vtu(path, Nx, Ny, Nz, h_xyz, cell_data):
    z, y, x = np.meshgrid(np.arange(Nz + 1), np.arange(Ny + 1), np.arange(Nx + 1), indexing="ij")
    pts = np.stack((x.reshape(-1) * h_xyz[0], y.reshape(-1) * h_xyz[1], z.reshape(-1) * h_xyz[2]), axis=1)
    pid = np.arange((Nz + 1) * (Ny + 1) * (Nx + 1)).reshape(Nz + 1, Ny + 1, Nx + 1)
    c = [pid[:-1, :-1, :-1], pid[:-1, :-1, 1:], pid[:-1, 1:, 1:], pid[:-1, 1:, :-1],
         pid[1:, :-1, :-1], pid[1:, :-1, 1:], pid[1:, 1:, 1:], pid[1:, 1:, :-1]]
    conn = np.stack([v.reshape(-1) for v in c], axis=1).astype(np.int64)
    n = conn.shape[0]
    offsets = (np.arange(1, n + 1) * 8).astype(np.int64)
    types = np.full(n, 12, dtype=np.uint8)  # VTK_HEXAHEDRON
    cd = []
    for name, a in cell_data.items():
        a = np.asarray(a)
        if a.dtype not in _VTK_TYPES:
            a = a.astype(np.float64)
        ncomp = 1 if a.ndim == 1 else a.shape[1]
        cd.append(_array(name, a.reshape(-1), ncomp))
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
                'header_type="UInt64">\n<UnstructuredGrid>\n')
        f.write(f'<Piece NumberOfPoints="{pts.shape[0]}" NumberOfCells="{n}">\n')
        f.write("<Points>" + _array("Points", pts.reshape(-1).astype(np.float64), 3) + "</Points>\n")
        f.write("<Cells>" + _array("connectivity", conn.reshape(-1)) + _array("offsets", offsets)
                + _array("types", types) + "</Cells>\n")
        f.write("<CellData>" + "".join(cd) + "</CellData>\n")
        f.write("</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")


def read_cell_data(path):
    """Cell arrays of a file written by write_hex_vtu (name -> (N,) or (N, ncomp) array), as the
    reference's process_eta reads them back with meshio (utils.py:181-189)."""
    import re
    rev = {v: k for k, v in _VTK_TYPES.items()}
    txt = open(path).read()
    body = txt[txt.index("<CellData>"):txt.index("</CellData>")]
    out = {}
    for m in re.finditer(r'<DataArray type="(\w+)" Name="([^"]+)" NumberOfComponents="(\d+)" format="binary">'
                         r'([^<]*)</DataArray>', body):
        typ, name, ncomp, data = m.group(1), m.group(2), int(m.group(3)), m.group(4)
        raw = base64.b64decode(data)
        nbytes = struct.unpack("<Q", raw[:8])[0]
        a = np.frombuffer(raw[8:8 + nbytes], dtype=rev[typ]).copy()
        out[name] = a if ncomp == 1 else a.reshape(-1, ncomp)
    return out
