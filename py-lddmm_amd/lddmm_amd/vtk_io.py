"""Legacy ASCII VTK POLYDATA reader / writer (no vtk library needed).

Reads the format of the reference's TestData surfaces ("# vtk DataFile",
"ASCII", "DATASET POLYDATA", POINTS / POLYGONS, optional POINT_DATA
scalars named 'weights').  The writer emits the same layout as the
reference's Surface.saveVTK (py-lddmm base/surfaces.py:1794-1926): POINTS,
POLYGONS, and POINT_DATA scalars / vectors.
"""
import numpy as np


def _tokens(path):
    with open(path) as fh:
        for line in fh:
            yield line.rstrip("\n")


def read_polydata(path):
    lines = list(_tokens(path))
    if not lines or not lines[0].startswith("# vtk"):
        raise ValueError(f"{path}: not a legacy VTK file")
    if lines[2].strip().upper() != "ASCII":
        raise NotImplementedError(f"{path}: only ASCII VTK is supported")
    body = " ".join(lines[3:]).split()
    i = 0
    v = f = w = None
    npts = 0
    while i < len(body):
        tok = body[i].upper()
        if tok == "DATASET":
            if body[i + 1].upper() != "POLYDATA":
                raise NotImplementedError(f"{path}: dataset {body[i + 1]} is not POLYDATA")
            i += 2
        elif tok == "POINTS":
            npts = int(body[i + 1])
            vals = np.array(body[i + 3:i + 3 + 3 * npts], dtype=np.float64)
            v = vals.reshape(npts, 3)
            i += 3 + 3 * npts
        elif tok == "POLYGONS":
            nf, size = int(body[i + 1]), int(body[i + 2])
            vals = np.array(body[i + 3:i + 3 + size], dtype=np.int64)
            if not np.all(vals[0::4] == 3) or size != 4 * nf:
                raise NotImplementedError(f"{path}: only triangles are supported")
            f = vals.reshape(nf, 4)[:, 1:]
            i += 3 + size
        elif tok == "POINT_DATA":
            i += 2
        elif tok == "SCALARS":
            name = body[i + 1]
            j = i + 3
            if j < len(body) and body[j].isdigit() and len(body[j]) < 3:
                j += 1
            if body[j].upper() == "LOOKUP_TABLE":
                j += 2
            vals = np.array(body[j:j + npts], dtype=np.float64)
            if name == "weights":
                w = vals
            i = j + npts
        elif tok in ("VECTORS", "NORMALS"):
            i += 3 + 3 * npts
        else:
            i += 1
    if v is None or f is None:
        raise ValueError(f"{path}: missing POINTS or POLYGONS")
    return v, f, w


def write_polydata(path, vertices, faces, weights=None, scalars=None, scal_name='scalars',
                   vectors=None, vect_name='vectors'):
    v = np.asarray(vertices, dtype=np.float64)
    f = np.asarray(faces, dtype=np.int64)
    with open(path, "w") as fh:
        fh.write("# vtk DataFile Version 3.0\nSurface Data\nASCII\nDATASET POLYDATA\n\n")
        fh.write(f"POINTS {v.shape[0]} float\n")
        for p in v:
            fh.write(f"{p[0]} {p[1]} {p[2]}\n")
        fh.write(f"POLYGONS {f.shape[0]} {4 * f.shape[0]}\n")
        for t in f:
            fh.write(f"3 {t[0]} {t[1]} {t[2]}\n")
        fh.write(f"\nPOINT_DATA {v.shape[0]}\n")
        if weights is not None:
            fh.write("SCALARS weights float 1\nLOOKUP_TABLE default\n")
            for x in np.asarray(weights, dtype=np.float64):
                fh.write(f"{x}\n")
        if scalars is not None:
            fh.write(f"SCALARS {scal_name} float 1\nLOOKUP_TABLE default\n")
            for x in np.asarray(scalars, dtype=np.float64).ravel():
                fh.write(f"{x}\n")
        if vectors is not None:
            fh.write(f"VECTORS {vect_name} float\n")
            for p in np.asarray(vectors, dtype=np.float64):
                fh.write(f"{p[0]} {p[1]} {p[2]}\n")
