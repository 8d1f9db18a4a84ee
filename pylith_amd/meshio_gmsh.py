"""Reader for Gmsh 4.1 mesh files (binary or ASCII), the format of the reference's MMS fixture
meshes (tests/mmstests/linearelasticity/nofaults-3d/data/{hex,tet}.msh, read there through
PETSc's DMPlexCreateGmsh via meshio/MeshIOPetsc.cc).

Volume cells become the mesh (8-node hexahedra -> Q1 in PyLith vertex order, 4-node tetrahedra
-> P2 with edge-midpoint nodes appended after the vertices); a cell's material label is the
physical tag of the volume entity it belongs to (``material-id:24`` -> 24).  Boundary groups
are the node sets of the physical surfaces, by physical name.  Cells are oriented so that the
Jacobian determinant is positive (vertex order flipped where it is not).
"""
import struct

import numpy as np

from .meshgen import HEX_Q1, TET_P2, Mesh

# nodes per element of the element types read here (Gmsh numbering)
_NODES = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 15: 1}
_TET_EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


class _Reader:
    def __init__(self, data, binary, size_t):
        self.d, self.binary, self.pos = data, binary, 0
        self.st = "<q" if size_t == 8 else "<i"
        self.st_n = size_t

    # binary primitives
    def _unpack(self, fmt, n):
        v = struct.unpack_from(fmt, self.d, self.pos)
        self.pos += n
        return v

    def size(self):
        if self.binary:
            return self._unpack(self.st, self.st_n)[0]
        return int(self.word())

    def int32(self):
        if self.binary:
            return self._unpack("<i", 4)[0]
        return int(self.word())

    def doubles(self, n):
        if self.binary:
            a = np.frombuffer(self.d, "<f8", n, self.pos).copy()
            self.pos += 8 * n
            return a
        return np.array([float(self.word()) for _ in range(n)])

    def sizes(self, n):
        if self.binary:
            a = np.frombuffer(self.d, "<i8" if self.st_n == 8 else "<i4", n, self.pos).astype(np.int64)
            self.pos += self.st_n * n
            return a
        return np.array([int(self.word()) for _ in range(n)], np.int64)

    def ints(self, n):
        if self.binary:
            a = np.frombuffer(self.d, "<i4", n, self.pos).astype(np.int64)
            self.pos += 4 * n
            return a
        return np.array([int(self.word()) for _ in range(n)], np.int64)

    # ASCII token
    def word(self):
        d = self.d
        while d[self.pos:self.pos + 1].isspace():
            self.pos += 1
        start = self.pos
        while self.pos < len(d) and not d[self.pos:self.pos + 1].isspace():
            self.pos += 1
        return d[start:self.pos].decode()


def _section(data, name):
    start = data.find(b"$" + name + b"\n")
    if start < 0:
        return None
    return start + len(name) + 2


def _entities(r):
    """Physical tags per (dim, entity tag)."""
    counts = [r.size() for _ in range(4)]
    phys = {}
    for dim, count in enumerate(counts):
        for _ in range(count):
            tag = r.int32()
            r.doubles(3 if dim == 0 else 6)
            nphys = r.size()
            phys[(dim, tag)] = list(r.ints(nphys))
            if dim > 0:
                nb = r.size()
                r.ints(nb)
    return phys


def read(path):
    """Returns (mesh, groups): mesh a meshgen.Mesh (Q1 for hexahedra, P2 for tetrahedra), groups a
    dict physical-surface name -> sorted node ids."""
    data = open(path, "rb").read()
    p = _section(data, b"MeshFormat")
    if p is None:
        raise ValueError(f"{path}: not a Gmsh mesh file")
    end = data.index(b"\n", p)
    version, file_type, size_t = data[p:end].split()
    if not version.startswith(b"4."):
        raise ValueError(f"{path}: Gmsh format {version.decode()} (4.1 is read)")
    binary = file_type == b"1"
    if binary and struct.unpack_from("<i", data, end + 1)[0] != 1:
        raise ValueError(f"{path}: big-endian binary files are not read")
    names = {}
    p = _section(data, b"PhysicalNames")
    if p is not None:
        lines = data[p:data.index(b"$EndPhysicalNames", p)].decode().strip().splitlines()
        for line in lines[1:]:
            dim, tag, name = line.split(maxsplit=2)
            names[(int(dim), int(tag))] = name.strip('"')
    phys = {}
    p = _section(data, b"Entities")
    if p is not None:
        r = _Reader(data, binary, int(size_t))
        r.pos = p
        phys = _entities(r)
    # nodes
    r = _Reader(data, binary, int(size_t))
    r.pos = _section(data, b"Nodes")
    nblocks, nnodes, _, max_tag = (r.size() for _ in range(4))
    xyz = np.zeros((max_tag + 1, 3))
    present = np.zeros(max_tag + 1, bool)
    for _ in range(nblocks):
        r.int32(), r.int32()
        parametric = r.int32()
        n = r.size()
        tags = r.sizes(n)
        if parametric:
            raise ValueError(f"{path}: parametric nodes are not read")
        xyz[tags] = r.doubles(3 * n).reshape(n, 3)
        present[tags] = True
    # elements
    r.pos = _section(data, b"Elements")
    nblocks = r.size()
    r.size(), r.size(), r.size()
    vols = {4: [], 5: []}
    vol_lab = {4: [], 5: []}
    surf = {}
    for _ in range(nblocks):
        dim, etag, etype = r.int32(), r.int32(), r.int32()
        n = r.size()
        if etype not in _NODES:
            raise ValueError(f"{path}: element type {etype} is not read")
        k = _NODES[etype]
        conn = r.sizes(n * (1 + k)).reshape(n, 1 + k)[:, 1:]
        ptags = phys.get((dim, etag), [])
        if dim == 3:
            vols[etype].append(conn)
            vol_lab[etype].append(np.full(n, ptags[0] if ptags else 0, np.int32))
        elif dim == 2:
            for t in ptags:
                surf.setdefault(names.get((2, t), str(t)), []).append(conn)
    used_types = [t for t in (4, 5) if vols[t]]
    if len(used_types) != 1:
        raise ValueError(f"{path}: expected one volume cell type (hexahedra or tetrahedra)")
    etype = used_types[0]
    conn = np.concatenate(vols[etype])
    labels = np.concatenate(vol_lab[etype])
    # compact vertex numbering (tags in order)
    vert_tags = np.nonzero(present)[0]
    remap = -np.ones(max_tag + 1, np.int64)
    remap[vert_tags] = np.arange(vert_tags.size)
    coords = xyz[vert_tags]
    cells = remap[conn]
    if np.any(cells < 0):
        raise ValueError(f"{path}: cells reference undefined nodes")
    groups = {name: np.unique(np.concatenate([remap[c].ravel() for c in parts])) for name, parts in surf.items()}
    if etype == 5:
        # Gmsh and PyLith share the hexahedron order (0-3 one face, 4-7 the opposite face,
        # vertex i+4 across from vertex i); flip faces of negatively oriented cells
        x = coords[cells]
        det = np.einsum("ij,ij->i", x[:, 1] - x[:, 0], np.cross(x[:, 3] - x[:, 0], x[:, 4] - x[:, 0]))
        neg = det < 0
        cells[neg] = cells[neg][:, [0, 3, 2, 1, 4, 7, 6, 5]]
        closure = cells.astype(np.int32)
        mesh = Mesh(HEX_Q1, closure, closure.copy(), coords, coords, None, labels)
        return mesh, groups
    x = coords[cells]
    det = np.einsum("ij,ij->i", x[:, 1] - x[:, 0], np.cross(x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]))
    neg = det < 0
    cells[neg] = cells[neg][:, [0, 1, 3, 2]]
    # P2 nodes: the vertices, then one node per edge at its midpoint
    nv = coords.shape[0]
    edges = np.stack([np.sort(cells[:, [a, b]], axis=1) for a, b in _TET_EDGES], 1)  # [nc, 6, 2]
    uniq, inv = np.unique(edges.reshape(-1, 2), axis=0, return_inverse=True)
    node_coords = np.concatenate([coords, 0.5 * (coords[uniq[:, 0]] + coords[uniq[:, 1]])])
    closure = np.concatenate([cells, nv + inv.reshape(-1, 6)], 1).astype(np.int32)
    # boundary groups gain the nodes of the edges of their triangles
    for name, parts in surf.items():
        tris = [remap[c] for c in parts if c.shape[1] == 3]
        if not tris:
            continue
        t = np.concatenate(tris)
        fe = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [0, 2]]]), axis=1)
        fe = np.unique(fe, axis=0)
        # locate each face edge among the cell edges (both lists sorted lexicographically)
        key = lambda e: e[:, 0] * (nv + 1) + e[:, 1]  # noqa: E731
        idx = np.searchsorted(key(uniq), key(fe))
        if np.any(idx >= len(uniq)) or np.any(key(uniq)[np.minimum(idx, len(uniq) - 1)] != key(fe)):
            raise ValueError(f"{path}: boundary triangles of {name} are not faces of the tetrahedra")
        groups[name] = np.concatenate([groups[name], nv + idx])
    mesh = Mesh(TET_P2, closure, cells.astype(np.int32), coords, node_coords, None, labels)
    return mesh, groups
