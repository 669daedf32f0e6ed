"""Surface and point-set file formats without the vtk library.

Legacy ASCII VTK POLYDATA (the format of the reference's TestData and of
Surface.saveVTK / saveVTK2, py-lddmm base/surfaces.py:1794-1926): POINTS,
POLYGONS (triangles), CELL_DATA 'labels' (the face component) and POINT_DATA
scalars / vectors / normals / tensors.  BYU (base/surfaces.py:1697-1787),
OFF (:1658-1695), OBJ and STL (:1948-2009, read through vtk readers in the
reference: vertices and the first three point ids of each polygon, STL
duplicate points merged).  Point sets (base/pointSets.py:48-237): the
'N dim' vector format, .lmk landmarks, savePoints and saveTrajectories.
"""
import struct

import numpy as np


def _g(x):
    """Shortest text that reads back to the same float64."""
    return repr(float(x))


_DATA_KEYS = {"SCALARS", "VECTORS", "NORMALS", "TENSORS", "POINT_DATA", "CELL_DATA", "LOOKUP_TABLE",
              "FIELD", "TEXTURE_COORDINATES", "COLOR_SCALARS", "POLYGONS", "LINES", "VERTICES",
              "TRIANGLE_STRIPS", "POINTS", "DATASET", "METADATA"}


def read_polydata(path, with_labels=False):
    """(vertices [N,3], faces [F,3] int64, weights [N] or None[, labels [F] or None])."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    if not lines or not lines[0].startswith("# vtk"):
        raise ValueError(f"{path}: not a legacy VTK file")
    if lines[2].strip().upper() != "ASCII":
        raise NotImplementedError(f"{path}: only ASCII VTK is supported")
    body = " ".join(lines[3:]).split()
    i = 0
    v = f = w = lab = None
    npts = nfaces = 0
    section = "POINT_DATA"
    while i < len(body):
        tok = body[i].upper()
        if tok == "DATASET":
            if body[i + 1].upper() != "POLYDATA":
                raise NotImplementedError(f"{path}: dataset {body[i + 1]} is not POLYDATA")
            i += 2
        elif tok == "POINTS":
            npts = int(body[i + 1])
            v = np.array(body[i + 3:i + 3 + 3 * npts], dtype=np.float64).reshape(npts, 3)
            i += 3 + 3 * npts
        elif tok == "POLYGONS":
            nfaces, size = int(body[i + 1]), int(body[i + 2])
            vals = np.array(body[i + 3:i + 3 + size], dtype=np.int64)
            if not np.all(vals[0::4] == 3) or size != 4 * nfaces:
                raise NotImplementedError(f"{path}: only triangles are supported")
            f = vals.reshape(nfaces, 4)[:, 1:]
            i += 3 + size
        elif tok in ("POINT_DATA", "CELL_DATA"):
            section = tok
            i += 2
        elif tok == "SCALARS":
            name = body[i + 1]
            j = i + 3
            if j < len(body) and body[j].isdigit():
                j += 1                                   # numComp
            if j < len(body) and body[j].upper() == "LOOKUP_TABLE":
                j += 2
            count = npts if section == "POINT_DATA" else nfaces
            vals = np.array(body[j:j + count], dtype=np.float64)
            if section == "POINT_DATA" and name == "weights":
                w = vals
            if section == "CELL_DATA" and name == "labels":
                lab = vals.astype(np.int64)
            i = j + count
        else:
            # any other block (VECTORS / NORMALS / TENSORS / FIELD ...): skip its numbers
            i += 1
            while i < len(body) and (body[i].upper() not in _DATA_KEYS):
                i += 1
    if v is None or f is None:
        raise ValueError(f"{path}: missing POINTS or POLYGONS")
    return (v, f, w, lab) if with_labels else (v, f, w)


class vtkFields:
    """Point-data fields of saveVTK2 (base/surfaces.py:53-58): each list holds
    name, array, name, array, ..."""

    def __init__(self):
        self.scalars = []
        self.vectors = []
        self.normals = []
        self.tensors = []


def write_polydata2(path, vertices, faces, component=None, fields=None):
    """Layout of Surface.saveVTK2 (base/surfaces.py:1815-1877): POINTS,
    POLYGONS, CELL_DATA 'labels' (int), then POINT_DATA scalars, vectors,
    normals and tensors (two rows of three values per point, as the
    reference writes them).  Values are written at full float64 precision."""
    V = np.asarray(vertices, dtype=np.float64)
    F = np.asarray(faces, dtype=np.int64)
    comp = np.zeros(F.shape[0], dtype=np.int64) if component is None else np.asarray(component)
    with open(path, "w") as fh:
        fh.write("# vtk DataFile Version 3.0\nSurface Data\nASCII\nDATASET POLYDATA\n")
        fh.write(f"\nPOINTS {V.shape[0]} float")
        for p in V:
            fh.write(f"\n{_g(p[0])} {_g(p[1])} {_g(p[2])}")
        fh.write(f"\nPOLYGONS {F.shape[0]} {4 * F.shape[0]}")
        for t in F:
            fh.write(f"\n3 {t[0]} {t[1]} {t[2]}")
        fh.write(f"\nCELL_DATA {F.shape[0]}\nSCALARS labels int 1\nLOOKUP_TABLE default")
        for c in comp:
            fh.write(f"\n {int(c)}")
        if fields is not None:
            hdr = False

            def header():
                nonlocal hdr
                if not hdr:
                    fh.write(f"\nPOINT_DATA {V.shape[0]}")
                    hdr = True
            for k in range(len(fields.scalars) // 2):
                header()
                fh.write(f"\nSCALARS {fields.scalars[2 * k]} float 1\nLOOKUP_TABLE default")
                for x in np.asarray(fields.scalars[2 * k + 1], dtype=np.float64).ravel():
                    fh.write(f"\n {_g(x)}")
            for kind, lst in (("VECTORS", fields.vectors), ("NORMALS", fields.normals)):
                for k in range(len(lst) // 2):
                    header()
                    fh.write(f"\n{kind} {lst[2 * k]} float")
                    for p in np.asarray(lst[2 * k + 1], dtype=np.float64):
                        fh.write(f"\n {_g(p[0])} {_g(p[1])} {_g(p[2])}")
            for k in range(len(fields.tensors) // 2):
                header()
                fh.write(f"\nTENSORS {fields.tensors[2 * k]} float")
                for tt in np.asarray(fields.tensors[2 * k + 1], dtype=np.float64):
                    for kk in range(2):
                        fh.write(f"\n {_g(tt[kk, 0])} {_g(tt[kk, 1])} {_g(tt[kk, 2])}")
        fh.write("\n")


def write_polydata(path, vertices, faces, weights=None, scalars=None, scal_name='scalars',
                   vectors=None, vect_name='vectors', normals=None, tensors=None, component=None):
    """Surface.saveVTK (base/surfaces.py:1794-1812): weights first, then the
    optional scalars / vectors / normals / tensors, through the saveVTK2 layout."""
    vf = vtkFields()
    if weights is not None:
        vf.scalars += ['weights', weights]
    if scalars is not None:
        vf.scalars += [scal_name, scalars]
    if vectors is not None:
        vf.vectors += [vect_name, vectors]
    if normals is not None:
        vf.normals += ['normals', normals]
    if tensors is not None:
        vf.tensors += ['tensors', tensors]
    write_polydata2(path, vertices, faces, component, vf)


# ----------------------------------------------------------------- BYU / OFF
def read_byu(path):
    """base/surfaces.py:1697-1747: header 'ncomp npoints nfaces ...', component
    lines, vertices (two per line possible), 1-based faces with the last
    index of each face negated."""
    with open(path) as fh:
        head = fh.readline().split()
        ncomp, npts, nfaces = int(head[0]), int(head[1]), int(head[2])
        for _ in range(ncomp):
            fh.readline()
        rest = fh.read().split()
    v = np.array(rest[:3 * npts], dtype=np.float64).reshape(npts, 3)
    f = np.abs(np.array(rest[3 * npts:3 * npts + 3 * nfaces], dtype=np.int64)).reshape(nfaces, 3) - 1
    return v, f


def write_byu(path, vertices, faces):
    """base/surfaces.py:1750-1787 layout (full-precision values)."""
    V = np.asarray(vertices, dtype=np.float64)
    F = np.asarray(faces, dtype=np.int64)
    with open(path, "w") as fh:
        fh.write(f"1 {V.shape[0]} {F.shape[0]} {3 * F.shape[0]} 0\n1 {F.shape[0]}\n")
        for k in range(0, V.shape[0], 2):
            row = V[k:k + 2].ravel()
            fh.write(" ".join(_g(x) for x in row) + "\n")
        toks = []
        for t in F:
            toks += [str(t[0] + 1), str(t[1] + 1), str(-(t[2] + 1))]
        for k in range(0, len(toks), 16):
            fh.write(" ".join(toks[k:k + 16]) + "\n")


def _lines_skip(fh, comment='#'):
    for line in fh:
        s = line.strip()
        if s and not s.startswith(comment):
            yield s


def read_off(path):
    """base/surfaces.py:1658-1695: 'OFF', counts, vertices, '3 i j k' faces."""
    with open(path) as fh:
        it = _lines_skip(fh)
        if next(it).split()[0].lower() != 'off':
            raise ValueError(f"{path}: not an OFF file")
        npts, nfaces = (int(x) for x in next(it).split()[:2])
        v = np.array([next(it).split()[:3] for _ in range(npts)], dtype=np.float64).reshape(npts, 3)
        rows = [next(it).split() for _ in range(nfaces)]
    if any(int(r[0]) != 3 for r in rows):
        raise NotImplementedError(f"{path}: only triangulated OFF surfaces are read")
    f = np.array([r[1:4] for r in rows], dtype=np.int64).reshape(nfaces, 3)
    return v, f


def write_off(path, vertices, faces):
    V = np.asarray(vertices, dtype=np.float64)
    F = np.asarray(faces, dtype=np.int64)
    with open(path, "w") as fh:
        fh.write(f"OFF\n{V.shape[0]} {F.shape[0]} 0\n")
        for p in V:
            fh.write(f"{_g(p[0])} {_g(p[1])} {_g(p[2])}\n")
        for t in F:
            fh.write(f"3 {t[0]} {t[1]} {t[2]}\n")


# ----------------------------------------------------------------- OBJ / STL
def read_obj(path):
    """Wavefront OBJ: 'v x y z' vertices and 'f a b c ...' polygons (1-based,
    'a/t/n' forms and negative relative indices accepted); the first three
    point ids of each polygon make the face, as the reference's
    vtkOBJReader-based readOBJ keeps them."""
    verts, faces = [], []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == 'v':
                verts.append([float(x) for x in tok[1:4]])
            elif tok[0] == 'f':
                ids = []
                for t in tok[1:4]:
                    k = int(t.split('/')[0])
                    ids.append(k - 1 if k > 0 else len(verts) + k)
                faces.append(ids)
    return np.array(verts, dtype=np.float64).reshape(-1, 3), np.array(faces, dtype=np.int64).reshape(-1, 3)


def read_stl(path):
    """ASCII or binary STL; identical corner points are merged (vtkSTLReader's
    default), faces in file order."""
    with open(path, "rb") as fh:
        data = fh.read()
    tri = None
    if data[:5].lower() == b"solid" and b"facet" in data[:1024]:
        vals = [float(t) for line in data.decode(errors="replace").splitlines()
                for t in line.split()[1:4] if line.strip().lower().startswith("vertex")]
        tri = np.array(vals, dtype=np.float64).reshape(-1, 3, 3)
    if tri is None:
        n = struct.unpack("<I", data[80:84])[0]
        rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                            count=n, offset=84)
        tri = rec["v"].astype(np.float64)
    pts = tri.reshape(-1, 3)
    uniq, first, inv = np.unique(pts, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)                   # points in order of first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return uniq[order], rank[inv.ravel()].reshape(-1, 3).astype(np.int64)


# --------------------------------------------------------------- point sets
def read_vector(path):
    """base/pointSets.py:80-101: 'N dim' header, then 'x y z w' rows."""
    with open(path) as fh:
        head = fh.readline().split()
        n = int(head[0])
        rows = [fh.readline().split() for _ in range(n)]
    v = np.array([r[:3] for r in rows], dtype=np.float64).reshape(n, 3)
    w = np.array([r[3] for r in rows], dtype=np.float64).reshape(n, 1)
    return v, w


def loadlmk(path, dim=3):
    """base/pointSets.py:104-163: .lmk landmarks (label line + coordinate line
    per point); version >= 6 files end with centre and scale lines that map
    the coordinates."""
    with open(path) as fh:
        first = fh.readline()
        version = 1
        parts = first.split("-")
        if len(parts) == 2:
            try:
                version = int(float(parts[1]))
            except ValueError:
                pass
        n = int(fh.readline().split()[0])
        x = np.zeros((n, dim))
        labels = []
        for i in range(n):
            labels.append(fh.readline())
            row = fh.readline().split()
            x[i] = [float(row[k]) for k in range(dim)]
        if version >= 6:
            tail = [ln for ln in fh.read().splitlines()]
            if len(tail) >= 2:
                centers, scales = tail[-2].split(','), tail[-1].split(',')
                if len(centers) == dim and len(scales) == dim:
                    x = x * np.array([float(s) for s in scales]) + np.array([float(c) for c in centers])
    return x, labels


def savelmk(x, path):
    """base/pointSets.py:166-181"""
    x = np.asarray(x, dtype=np.float64)
    with open(path, "w") as fh:
        fh.write(f"Landmarks-1.0\n {x.shape[0]}\n")
        for i in range(x.shape[0]):
            fh.write(f'"L-{i}"\n' + " ".join(_g(c) for c in x[i]) + " \n")
        fh.write("1 1 \n")


def savePoints(path, x, vectors=None, scalars=None):
    """base/pointSets.py:185-209: UNSTRUCTURED_GRID points with optional
    named point scalars / vectors (dicts)."""
    x = np.asarray(x, dtype=np.float64)
    if x.shape[1] < 3:
        x = np.concatenate((x, np.zeros((x.shape[0], 3 - x.shape[1]))), axis=1)
    with open(path, "w") as fh:
        fh.write("# vtk DataFile Version 3.0\nSurface Data\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        fh.write(f"\nPOINTS {x.shape[0]} float")
        for p in x:
            fh.write(f"\n{_g(p[0])} {_g(p[1])} {_g(p[2])}")
        if vectors is None and scalars is None:
            fh.write("\n")
            return
        fh.write(f"\nPOINT_DATA {x.shape[0]}")
        for k, val in (scalars or {}).items():
            fh.write(f"\nSCALARS {k} float 1\nLOOKUP_TABLE default")
            for s in np.asarray(val, dtype=np.float64).ravel():
                fh.write(f"\n {_g(s)} ")
        for k, val in (vectors or {}).items():
            fh.write(f"\nVECTORS {k} float")
            for p in np.asarray(val, dtype=np.float64):
                fh.write(f"\n {_g(p[0])} {_g(p[1])} {_g(p[2])}")
        fh.write("\n")


def read_points(path):
    """Reads what savePoints writes: (points [N,3], {name: scalars}, {name: vectors})."""
    with open(path) as fh:
        body = " ".join(fh.read().splitlines()[3:]).split()
    i, x, sc, vec, n = 0, None, {}, {}, 0
    while i < len(body):
        tok = body[i].upper()
        if tok == "POINTS":
            n = int(body[i + 1])
            x = np.array(body[i + 3:i + 3 + 3 * n], dtype=np.float64).reshape(n, 3)
            i += 3 + 3 * n
        elif tok == "SCALARS":
            name, j = body[i + 1], i + 4
            if body[j].upper() == "LOOKUP_TABLE":
                j += 2
            sc[name] = np.array(body[j:j + n], dtype=np.float64)
            i = j + n
        elif tok == "VECTORS":
            vec[body[i + 1]] = np.array(body[i + 3:i + 3 + 3 * n], dtype=np.float64).reshape(n, 3)
            i += 3 + 3 * n
        else:
            i += 1
    return x, sc, vec


def saveTrajectories(path, xt):
    """base/pointSets.py:212-236: trajectories [T+1, N, d] as POLYDATA lines
    from each point to itself one step later, with a 'time' point scalar."""
    xt = np.asarray(xt, dtype=np.float64)
    if xt.shape[2] == 2:
        xt = np.concatenate((xt, np.zeros(xt.shape[:2] + (1,))), axis=2)
    T1, N = xt.shape[0], xt.shape[1]
    with open(path, "w") as fh:
        fh.write("# vtk DataFile Version 3.0\ncurves \nASCII\nDATASET POLYDATA\n")
        fh.write(f"\nPOINTS {T1 * N} float")
        for p in xt.reshape(-1, 3):
            fh.write(f"\n{_g(p[0])} {_g(p[1])} {_g(p[2])}")
        nl = (T1 - 1) * N
        fh.write(f"\nLINES {nl} {3 * nl}")
        for t in range(T1 - 1):
            for ll in range(N):
                fh.write(f"\n2 {t * N + ll} {(t + 1) * N + ll}")
        fh.write(f"\nPOINT_DATA {T1 * N}\nSCALARS time int 1\nLOOKUP_TABLE default")
        for t in range(T1):
            fh.write(f"\n{t}" * N)
        fh.write("\n")
