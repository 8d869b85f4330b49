"""Reading (and writing) gmsh ``.msh`` files, so meshes made by the reference's gmsh
scripts (``mesh_gen/``) on a machine that has gmsh can be run here unchanged.

No gmsh is needed: the ASCII formats MSH 4.1 and 2.2 are parsed directly.  The result is
the tuple the reference's mesh generators return and its drivers consume
(run_fvm.py:21-34):

    points [nv, d] float64, cells [N, d+1] int32,
    (interior facets [F_i, d] int64, boundary facets [F_b, d] int64), boundary tags

with interior facets as ``extract_interor_edges`` orders them (sorted vertex tuples,
lexicographic, create_mesh.py:11-19) and each boundary facet tagged with the name of the
physical group of the line (2-D) or triangle (3-D) element lying on it.  Boundary facets
keep the order of those boundary elements in the file.  Cells are the triangles (2-D) or
tetrahedra (3-D); element types 1 (line), 2 (triangle) and 4 (tetrahedron).
"""
from __future__ import annotations

import numpy as np

class MshFormatError(ValueError):
    pass


def _sections(text: str) -> dict:
    out, lines, i = {}, text.splitlines(), 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("$") and not ln.startswith("$End"):
            name = ln[1:]
            j = i + 1
            while j < len(lines) and lines[j].strip() != f"$End{name}":
                j += 1
            if j == len(lines):
                raise MshFormatError(f"section ${name} is not closed")
            out[name] = lines[i + 1:j]
            i = j
        i += 1
    return out


def _physical_names(sec) -> dict:
    names = {}
    if sec is None:
        return names
    for ln in sec[1:]:
        p = ln.split(maxsplit=2)
        names[(int(p[0]), int(p[1]))] = p[2].strip().strip('"')
    return names


def _read_v41(S):
    """Nodes (tag -> xyz) and element blocks [(dim, physical tags, type, node tags)]."""
    ent_phys = {}
    if "Entities" in S:
        lines = S["Entities"]
        counts = [int(x) for x in lines[0].split()]
        k = 1
        for dim, n in enumerate(counts):
            for _ in range(n):
                p = lines[k].split()
                k += 1
                tag = int(p[0])
                off = 4 if dim == 0 else 7
                nphys = int(p[off])
                ent_phys[(dim, tag)] = [int(x) for x in p[off + 1:off + 1 + nphys]]
    ln = S["Nodes"]
    nblk = int(ln[0].split()[0])
    tags, xyz, k = [], [], 1
    for _ in range(nblk):
        edim, etag, param, n = (int(x) for x in ln[k].split())
        k += 1
        tags.extend(int(ln[k + i]) for i in range(n))
        k += n
        for i in range(n):
            xyz.append([float(x) for x in ln[k + i].split()[:3]])
        k += n
    le = S["Elements"]
    nblk = int(le[0].split()[0])
    blocks, k = [], 1
    for _ in range(nblk):
        edim, etag, etype, n = (int(x) for x in le[k].split())
        k += 1
        rows = np.array([[int(x) for x in le[k + i].split()] for i in range(n)], np.int64).reshape(n, -1)
        k += n
        blocks.append((edim, ent_phys.get((edim, etag), []), etype, rows[:, 1:]))
    return np.array(tags, np.int64), np.array(xyz, np.float64), blocks


def _read_v22(S):
    ln = S["Nodes"]
    n = int(ln[0])
    arr = np.array([[float(x) for x in ln[1 + i].split()] for i in range(n)])
    tags, xyz = arr[:, 0].astype(np.int64), arr[:, 1:4]
    le = S["Elements"]
    m = int(le[0])
    blocks = []
    for i in range(m):
        p = [int(x) for x in le[1 + i].split()]
        etype, ntag = p[1], p[2]
        phys = [p[3]] if ntag >= 1 else []
        nodes = np.array(p[3 + ntag:], np.int64)[None]
        blocks.append((None, phys, etype, nodes))
    return tags, xyz, blocks


def read_msh(path: str, dim: int | None = None):
    """(points, cells, (interior, boundary), tags) of a gmsh ASCII mesh file."""
    text = open(path, "r").read()
    S = _sections(text)
    if "MeshFormat" not in S:
        raise MshFormatError("not a gmsh mesh file ($MeshFormat missing)")
    version, ftype = S["MeshFormat"][0].split()[:2]
    if ftype != "0":
        raise MshFormatError("binary .msh files are not supported; save as ASCII")
    names = _physical_names(S.get("PhysicalNames"))
    if version.startswith("4"):
        ntags, xyz, blocks = _read_v41(S)
    elif version.startswith("2"):
        ntags, xyz, blocks = _read_v22(S)
    else:
        raise MshFormatError(f"unsupported MSH version {version}")
    types = {b[2] for b in blocks}
    if dim is None:
        dim = 3 if 4 in types else 2
    cell_t, face_t = (4, 2) if dim == 3 else (2, 1)
    lut = np.full(int(ntags.max()) + 1, -1, np.int64)
    lut[ntags] = np.arange(len(ntags))

    def idx(rows):
        r = lut[rows]
        if np.any(r < 0):
            raise MshFormatError("an element refers to an undefined node")
        return r

    cells = np.concatenate([idx(b[3]) for b in blocks if b[2] == cell_t]) if cell_t in types else None
    if cells is None or not len(cells):
        raise MshFormatError(f"no {'tetrahedra' if dim == 3 else 'triangles'} in {path}")
    faces, ftags = [], []
    for edim, phys, etype, rows in blocks:
        if etype != face_t:
            continue
        nm = [names.get((dim - 1, p), names.get((edim, p), str(p))) for p in phys] or [""]
        faces.append(idx(rows))
        ftags.extend([nm[0]] * len(rows))
    # keep only the vertices the cells use
    used = np.unique(cells)
    new = np.full(len(xyz), -1, np.int64)
    new[used] = np.arange(len(used))
    pts = xyz[used][:, :dim]
    cells = new[cells]
    # facets of the cells
    pats = [(0, 1), (1, 2), (2, 0)] if dim == 2 else [(0, 1, 2), (0, 2, 3), (0, 3, 1), (1, 3, 2)]
    allf = np.sort(np.concatenate([cells[:, list(p)] for p in pats]), axis=1)
    uniq, counts = np.unique(allf, axis=0, return_counts=True)
    if np.any(counts > 2):
        raise MshFormatError("non-manifold facet (shared by more than two cells)")
    interior = uniq[counts == 2]
    bset = uniq[counts == 1]
    # tag boundary facets from the boundary elements, in file order
    if faces:
        fb = np.sort(new[np.concatenate(faces)], axis=1)
        ftags = np.asarray(ftags, object)
    else:
        fb, ftags = np.zeros((0, dim), np.int64), np.zeros(0, object)
    fb_keys = [tuple(r) for r in fb]
    b_keys = {tuple(r) for r in bset}
    missing = len(b_keys - set(fb_keys))
    if missing:
        raise MshFormatError(f"{missing} boundary facets carry no boundary element (physical group)")
    on_boundary = np.array([k in b_keys for k in fb_keys], bool) if len(fb) else np.zeros(0, bool)
    boundary, tags = fb[on_boundary], [str(t) for t in ftags[on_boundary]]
    return pts, cells.astype(np.int32), (interior.astype(np.int64), boundary.astype(np.int64)), tags


def write_msh(path: str, points, cells, boundary, tags):
    """Write a mesh tuple as MSH 4.1 ASCII: one physical group per boundary tag (boundary
    elements in the given order) and one for the cells."""
    pts = np.asarray(points, np.float64)
    cells = np.asarray(cells)
    boundary = np.asarray(boundary)
    dim = cells.shape[1] - 1
    names = list(dict.fromkeys(tags))
    cell_t, face_t = (4, 2) if dim == 3 else (2, 1)
    xyz = np.zeros((len(pts), 3))
    xyz[:, :dim] = pts
    out = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", str(len(names) + 1)]
    out += [f'{dim - 1} {i + 1} "{n}"' for i, n in enumerate(names)] + [f'{dim} {len(names) + 1} "domain"']
    out.append("$EndPhysicalNames")
    # one entity per boundary name (dim - 1) and one for the domain (dim)
    counts = [0, 0, 0, 0]
    counts[dim - 1] = len(names)
    counts[dim] = 1
    out += ["$Entities", " ".join(map(str, counts))]
    lo, hi = xyz.min(0), xyz.max(0)
    box = " ".join(f"{v:.17g}" for v in (*lo, *hi))
    for i in range(len(names)):
        out.append(f"{i + 1} {box} 1 {i + 1} 0")
    out.append(f"1 {box} 1 {len(names) + 1} 0")
    out.append("$EndEntities")
    out += ["$Nodes", f"1 {len(pts)} 1 {len(pts)}", f"{dim} 1 0 {len(pts)}"]
    out += [str(i + 1) for i in range(len(pts))]
    out += [f"{x:.17g} {y:.17g} {z:.17g}" for x, y, z in xyz]
    out.append("$EndNodes")
    tags = np.asarray(tags, object)
    nel = len(boundary) + len(cells)
    out += ["$Elements", f"{len(names) + 1} {nel} 1 {nel}"]
    e = 1
    for i, n in enumerate(names):
        rows = boundary[tags == n]
        out.append(f"{dim - 1} {i + 1} {face_t} {len(rows)}")
        for r in rows:
            out.append(f"{e} " + " ".join(str(int(v) + 1) for v in r))
            e += 1
    out.append(f"{dim} 1 {cell_t} {len(cells)}")
    for r in cells:
        out.append(f"{e} " + " ".join(str(int(v) + 1) for v in r))
        e += 1
    out.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
