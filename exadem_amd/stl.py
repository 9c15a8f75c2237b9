"""STL meshes as triangle-mesh drivers (register_stl_mesh, src/driver/register_stl_mesh.cpp:30-140).

read_stl follows STLMeshReader (src/driver/include/exaDEM/stl_mesh_reader.hpp:63-135): binary files are an
80-byte header, a uint32 triangle count, then per triangle a float32 normal, three float32 vertices and a
uint16 attribute; ASCII files list "vertex x y z" lines between "outer loop" and "endloop".

build_shape follows build_shape (src/driver/include/exaDEM/stl_mesh_to_driver.hpp:25-126) as it runs on one
thread: the vertices are the face vertices sorted and made unique (lexicographic x, y, z), faces keep the
file order, and each face lists its edges (v0 v1), (v1 v2), (v2 v0) with only consecutive repeats removed,
so an edge shared by two faces is listed twice.
"""
import numpy as np


def read_stl(path, binary=False):
    """Triangles of an STL file as float64 [nf][3][3]."""
    if binary:
        with open(path, "rb") as f:
            data = f.read()
        if len(data) < 84:
            raise ValueError(f"{path}: truncated binary STL")
        n = int(np.frombuffer(data, np.uint32, 1, 80)[0])
        rec = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
        if len(data) < 84 + n * rec.itemsize:
            raise ValueError(f"{path}: {n} triangles announced, file too short")
        tri = np.frombuffer(data, rec, n, 84)
        return tri["v"].astype(np.float64)
    faces, cur, inside = [], [], False
    with open(path) as f:
        tok = f.read().split()
    k = 0
    while k < len(tok):
        t = tok[k]
        if t == "outer":
            inside, cur = True, []
        elif t == "vertex" and inside:
            cur.append([float(tok[k + 1]), float(tok[k + 2]), float(tok[k + 3])])
            k += 3
        elif t == "endloop" and inside:
            if len(cur) != 3:
                raise ValueError(f"{path}: a face with {len(cur)} vertices (triangles only)")
            faces.append(cur)
            inside = False
        k += 1
    return np.array(faces, np.float64).reshape(-1, 3, 3)


def build_shape(triangles):
    """(vertices [nv][3], edges [ne][2], faces [nf][3]) of the mesh."""
    tri = np.ascontiguousarray(triangles, np.float64).reshape(-1, 3, 3)
    flat = tri.reshape(-1, 3)
    order = np.lexsort((flat[:, 2], flat[:, 1], flat[:, 0]))
    s = flat[order]
    keep = np.ones(len(s), bool)
    keep[1:] = np.any(s[1:] != s[:-1], axis=1)
    verts = s[keep]
    # index of each face vertex in the sorted unique list
    idx = np.empty(len(flat), np.int64)
    idx[order] = np.cumsum(keep) - 1
    faces = idx.reshape(-1, 3).astype(np.uint32)
    e = np.stack([faces, np.roll(faces, -1, axis=1)], axis=2).reshape(-1, 2)
    ek = np.ones(len(e), bool)
    ek[1:] = np.any(e[1:] != e[:-1], axis=1)
    return verts, e[ek].astype(np.uint32), faces


def load(path, binary=False, scale=None):
    """read_stl + build_shape; scale multiplies the vertices (shape::rescale before the Minkowski radius)."""
    v, e, f = build_shape(read_stl(path, binary))
    if scale is not None:
        if scale <= 0:
            raise ValueError("scale <= 0")
        v = v * scale
    return v, e, f


def write_stl(path, triangles, binary=False):
    """Writes triangles [nf][3][3] as an STL file (normals from the vertex order)."""
    tri = np.ascontiguousarray(triangles, np.float64).reshape(-1, 3, 3)
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = nrm / np.where(ln > 0, ln, 1.0)
    if binary:
        rec = np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
        out = np.zeros(len(tri), rec)
        out["n"] = nrm
        out["v"] = tri
        with open(path, "wb") as f:
            f.write(b"exadem_amd".ljust(80, b" "))
            f.write(np.uint32(len(tri)).tobytes())
            f.write(out.tobytes())
        return
    with open(path, "w") as f:
        f.write("solid mesh\n")
        for t, n in zip(tri, nrm):
            f.write("facet normal %r %r %r\n outer loop\n" % (float(n[0]), float(n[1]), float(n[2])))
            for v in t:
                f.write("  vertex %r %r %r\n" % (float(v[0]), float(v[1]), float(v[2])))
            f.write(" endloop\nendfacet\n")
        f.write("endsolid mesh\n")
