"""Synthetic meshes, materials and fields for the benchmark configurations.

PyLith reads meshes from files (meshio/MeshBuilder.cc:108-160); this module builds the
same kind of arrays for box domains instead:

* ``hex_box``: Q1 hexahedra with PyLith's vertex order (0-3 one face cyclic, 4-7 the opposite
  face, vertex i+4 above vertex i), positively oriented.
* ``tet_p2_box``: each cube split into 6 tetrahedra along its main diagonal (Kuhn split), P2
  nodes at vertices and edge midpoints.  Every half-grid point of the box is a P2 node.

Cells and nodes are numbered brick by brick (bricks of 8^3 hex cells, 4^3 cubes for tets) so
that consecutive cells form compact tiles and a tile's nodes occupy short contiguous runs, the
locality PyLith gets from its mesh reordering option (``reorder_mesh``).
"""
from __future__ import annotations

import dataclasses

import numpy as np

HEX_Q1 = 1
TET_P2 = 2


@dataclasses.dataclass
class Mesh:
    cell_type: int
    closure: np.ndarray      # [ncells, nb] int32 node ids
    cone: np.ndarray         # [ncells, nv] int32 vertex ids
    coords: np.ndarray       # [nverts, 3] float64
    node_coords: np.ndarray  # [nnodes, 3] float64 (nodal positions, for fields / BCs)
    shape: tuple             # cells per direction
    labels: np.ndarray = None
    colors: np.ndarray = None  # [ncells] colouring: cells of one colour share no node

    @property
    def num_cells(self):
        return self.closure.shape[0]

    @property
    def num_nodes(self):
        return self.node_coords.shape[0]

    @property
    def num_vertices(self):
        return self.coords.shape[0]


def _brick_order(dims, brick):
    """Permutation array: lexicographic index (i fastest) -> brick-major rank."""
    nx, ny, nz = dims
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(order="F"), j.ravel(order="F"), k.ravel(order="F")  # i fastest
    bx, by, bz = i // brick, j // brick, k // brick
    nbx, nby = -(-nx // brick), -(-ny // brick)
    key_brick = (bz * nby + by) * nbx + bx
    key_in = ((k % brick) * brick + (j % brick)) * brick + (i % brick)
    order = np.lexsort((key_in, key_brick))      # lexicographic indices sorted by brick
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return rank


def hex_box(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), brick=8, ordered=True, shear=None):
    """Q1 hex mesh of an n[0] x n[1] x n[2] box.  ``shear`` optionally maps (x, y, z) ->
    (x + shear * z, y, z) so planes of constant x dip (parallelepiped cells)."""
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    nx, ny, nz = n
    vx, vy, vz = nx + 1, ny + 1, nz + 1
    xs = [np.linspace(lo[d], hi[d], n[d] + 1) for d in range(3)]
    vrank = _brick_order((vx, vy, vz), brick) if ordered else np.arange(vx * vy * vz)
    I, J, K = np.meshgrid(np.arange(vx), np.arange(vy), np.arange(vz), indexing="ij")
    lex = (I + vx * (J + vy * K)).ravel(order="F")
    coords = np.empty((vx * vy * vz, 3))
    pts = np.stack([xs[0][I], xs[1][J], xs[2][K]], axis=-1).reshape(-1, 3, order="F")
    coords[vrank[lex]] = pts
    if shear is not None:
        coords[:, 0] += shear * coords[:, 2]
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(order="F"), cj.ravel(order="F"), ck.ravel(order="F")

    def v(a, b, c):
        return vrank[(ci + a) + vx * ((cj + b) + vy * (ck + c))]

    cl = np.stack([v(0, 0, 0), v(1, 0, 0), v(1, 1, 0), v(0, 1, 0),
                   v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1)], axis=1).astype(np.int32)
    crank = _brick_order(n, brick) if ordered else np.arange(nx * ny * nz)
    closure = np.empty_like(cl)
    closure[crank] = cl
    colors = np.empty(cl.shape[0], np.int32)
    colors[crank] = (ci % 2) + 2 * (cj % 2) + 4 * (ck % 2)
    return Mesh(HEX_Q1, closure, closure.copy(), coords, coords, n,
                np.zeros(closure.shape[0], np.int32), colors)


def jittered(mesh, amp, seed):
    """The same mesh with every node off the bounding box moved by a seeded uniform offset of up
    to `amp` (in coordinate units) per component: hexes become non-affine (Q1 only: the nodes
    are the vertices)."""
    if mesh.cell_type != HEX_Q1:
        raise ValueError("jittered: Q1 hex meshes only")
    rng = np.random.default_rng(seed)
    x = mesh.coords.copy()
    lo, hi = x.min(axis=0), x.max(axis=0)
    tol = 1e-9 * float(np.max(hi - lo))
    inner = np.all((x > lo + tol) & (x < hi - tol), axis=1)
    x[inner] += rng.uniform(-amp, amp, (int(inner.sum()), 3))
    return Mesh(mesh.cell_type, mesh.closure, mesh.cone, x, x, mesh.shape, mesh.labels, mesh.colors)


def renumbered(mesh, seed):
    """The same mesh with cells, nodes and vertices renumbered by seeded random permutations (a
    mesh whose numbering carries no locality, as a general mesh file or partitioner may give)."""
    rng = np.random.default_rng(seed)
    cperm = rng.permutation(mesh.num_cells)           # new cell i = old cell cperm[i]
    nnew = np.empty(mesh.num_nodes, np.int64)
    nnew[rng.permutation(mesh.num_nodes)] = np.arange(mesh.num_nodes)   # old node -> new id
    isoparametric = mesh.cell_type == HEX_Q1
    if isoparametric:
        vnew = nnew
    else:
        vnew = np.empty(mesh.num_vertices, np.int64)
        vnew[rng.permutation(mesh.num_vertices)] = np.arange(mesh.num_vertices)
    closure = nnew[mesh.closure[cperm]].astype(np.int32)
    cone = vnew[mesh.cone[cperm]].astype(np.int32)
    node_coords = np.empty_like(mesh.node_coords)
    node_coords[nnew] = mesh.node_coords
    coords = node_coords if isoparametric else np.empty_like(mesh.coords)
    if not isoparametric:
        coords[vnew] = mesh.coords
    labels = None if mesh.labels is None else mesh.labels[cperm]
    colors = None if mesh.colors is None else mesh.colors[cperm]
    return Mesh(mesh.cell_type, closure, cone, coords, node_coords, mesh.shape, labels, colors)


# Kuhn split: the 6 vertex paths (0,0,0) -> ... -> (1,1,1) through the cube corners.
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
_TET_EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def tet_p2_box(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), brick=4, ordered=True):
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    nx, ny, nz = n
    vx, vy, vz = nx + 1, ny + 1, nz + 1
    hx, hy, hz = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1  # half-grid
    xs = [np.linspace(lo[d], hi[d], n[d] + 1) for d in range(3)]
    hs = [np.linspace(lo[d], hi[d], 2 * n[d] + 1) for d in range(3)]
    vrank = _brick_order((vx, vy, vz), brick) if ordered else np.arange(vx * vy * vz)
    hrank = _brick_order((hx, hy, hz), 2 * brick) if ordered else np.arange(hx * hy * hz)
    I, J, K = np.meshgrid(np.arange(vx), np.arange(vy), np.arange(vz), indexing="ij")
    lex = (I + vx * (J + vy * K)).ravel(order="F")
    coords = np.empty((vx * vy * vz, 3))
    coords[vrank[lex]] = np.stack([xs[0][I], xs[1][J], xs[2][K]], -1).reshape(-1, 3, order="F")
    I, J, K = np.meshgrid(np.arange(hx), np.arange(hy), np.arange(hz), indexing="ij")
    lex = (I + hx * (J + hy * K)).ravel(order="F")
    node_coords = np.empty((hx * hy * hz, 3))
    node_coords[hrank[lex]] = np.stack([hs[0][I], hs[1][J], hs[2][K]], -1).reshape(-1, 3, order="F")

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(order="F"), cj.ravel(order="F"), ck.ravel(order="F")
    ncube = ci.size
    tets_v = np.empty((ncube, 6, 4, 3), np.int64)  # integer vertex-grid positions
    for t, p in enumerate(_PERMS):
        pos = np.zeros((ncube, 3), np.int64)
        base = np.stack([ci, cj, ck], 1)
        path = [base.copy()]
        for d in p:
            pos = pos.copy()
            pos[:, d] += 1
            path.append(base + pos)
        tets_v[:, t] = np.stack(path, 1)
    # orientation: det of (v1-v0, v2-v0, v3-v0) must be positive; swap v2,v3 otherwise
    e = tets_v[:, :, 1:] - tets_v[:, :, :1]
    det = np.linalg.det(e.astype(float))
    neg = det < 0
    tmp = tets_v[neg, 2].copy()
    tets_v[neg, 2] = tets_v[neg, 3]
    tets_v[neg, 3] = tmp
    tets_v = tets_v.reshape(-1, 4, 3)
    cone = vrank[tets_v[..., 0] + vx * (tets_v[..., 1] + vy * tets_v[..., 2])]
    half = 2 * tets_v
    hnode = lambda h: hrank[h[..., 0] + hx * (h[..., 1] + hy * h[..., 2])]
    cols = [hnode(half[:, a]) for a in range(4)]
    cols += [hnode((half[:, a] + half[:, b]) // 2) for a, b in _TET_EDGES]
    cl = np.stack(cols, 1).astype(np.int32)
    cone = cone.astype(np.int32)
    par = (ci % 2) + 2 * (cj % 2) + 4 * (ck % 2)
    colors = (par[:, None] * 6 + np.arange(6)[None, :]).ravel().astype(np.int32)
    if ordered:
        crank = _brick_order(n, brick)
        cube_of = np.empty(ncube, np.int64)
        cube_of[crank] = np.arange(ncube)
        sel = (cube_of[:, None] * 6 + np.arange(6)[None, :]).ravel()
        cl, cone, colors = cl[sel], cone[sel], colors[sel]
    return Mesh(TET_P2, np.ascontiguousarray(cl), np.ascontiguousarray(cone), coords, node_coords,
                n, np.zeros(cl.shape[0], np.int32), np.ascontiguousarray(colors))


# ------------------------------------------------------------------------------------------
# Materials (materials/Query.cc:204-270: mu = rho Vs^2, K = rho (Vp^2 - 4/3 Vs^2)).
# ------------------------------------------------------------------------------------------
def moduli(density, vs, vp):
    density, vs, vp = (np.asarray(a, dtype=np.float64) for a in (density, vs, vp))
    if np.any(vs <= 0) or np.any(vp <= 0) or np.any(density <= 0):
        raise ValueError("density, Vs and Vp must be positive")
    mu = density * vs * vs
    K = density * (vp * vp - 4.0 / 3.0 * vs * vs)
    return mu, K


def elastic_aux(num_cells, density=2500.0, vs=3000.0, vp=5292.0, seed=None, vs_sigma=None,
                vp_vs=None):
    """Aux field [density, shear_modulus, bulk_modulus] per cell (order-0 discretization,
    subfield order of materials/Elasticity.cc:234-241)."""
    if vs_sigma is not None:
        rng = np.random.Generator(np.random.PCG64(seed))
        vs = np.exp(np.log(vs) + vs_sigma * rng.standard_normal(num_cells))
        vp = vp_vs * vs
    rho = np.full(num_cells, density)
    vs = np.broadcast_to(np.asarray(vs, float), (num_cells,))
    vp = np.broadcast_to(np.asarray(vp, float), (num_cells,))
    mu, K = moduli(rho, vs, vp)
    return np.ascontiguousarray(np.stack([rho, mu, K], 1)), np.array([1, 1, 1], np.int32)


YEAR = 365.25 * 24.0 * 3600.0
GENMAX_SIZES = np.array([1, 1, 1, 3, 3, 18, 6], np.int32)


def genmax_aux(num_cells, nq=8, density=2500.0, vs=3000.0, vp=5292.0,
               viscosity=(9.46728e17, 4.73364e17, 1.893456e18), ratios=(0.25, 0.25, 0.25),
               seed=None, log_eta_range=None, viscous_strain=None, total_strain=None):
    """Generalized Maxwell aux field per cell and quadrature point, [density, shear_modulus,
    bulk_modulus, maxwell_time(3), shear_modulus_ratio(3), viscous_strain(18), total_strain(6)]
    (IsotropicLinearGenMaxwell.hh:109-165; maxwell_time = viscosity / (shear_modulus * ratio),
    materials/Query.cc).  ``log_eta_range`` = (lo, hi) draws each cell's three viscosities
    log-uniformly (seeded); viscous/total strain give the initial state (6 or 18 / 6 values,
    broadcast)."""
    rho = np.full(num_cells, density)
    mu, K = moduli(rho, np.full(num_cells, vs, float), np.full(num_cells, vp, float))
    r = np.broadcast_to(np.asarray(ratios, float), (num_cells, 3))
    if log_eta_range is not None:
        rng = np.random.Generator(np.random.PCG64(seed))
        lo, hi = np.log10(log_eta_range[0]), np.log10(log_eta_range[1])
        eta = 10.0 ** rng.uniform(lo, hi, (num_cells, 3))
    else:
        eta = np.broadcast_to(np.asarray(viscosity, float), (num_cells, 3))
    tau = eta / (mu[:, None] * r)
    cell = np.zeros((num_cells, 33))
    cell[:, 0], cell[:, 1], cell[:, 2] = rho, mu, K
    cell[:, 3:6], cell[:, 6:9] = tau, r
    if viscous_strain is not None:
        vs_ = np.asarray(viscous_strain, float)
        cell[:, 9:27] = np.tile(vs_, 3) if vs_.size == 6 else vs_
    if total_strain is not None:
        cell[:, 27:33] = np.asarray(total_strain, float)
    aux = np.repeat(cell, nq, axis=0)
    return np.ascontiguousarray(aux), GENMAX_SIZES.copy()


# ------------------------------------------------------------------------------------------
# Fields and boundary conditions.
# ------------------------------------------------------------------------------------------
def random_field(num_nodes, seed, amp=1e-3):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(-amp, amp, size=3 * num_nodes)


def sinusoid_field(node_coords, seed, amp=1e-3, count=8, length=None):
    """Sum of ``count`` seeded sinusoids per component, each of amplitude ``amp``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = node_coords
    L = length if length is not None else float(np.max(x.max(0) - x.min(0)))
    u = np.zeros((x.shape[0], 3))
    for _ in range(count):
        kv = 2.0 * np.pi * rng.integers(1, 5, size=3) / L
        ph = rng.uniform(0.0, 2.0 * np.pi, size=3)
        arg = x @ kv
        for c in range(3):
            u[:, c] += amp * np.sin(arg + ph[c])
    return u.ravel()


def shear_bcs(node_coords, lo_x, hi_x, shear=1.0, tol=1e-9):
    """-x face: all components 0; +x face: u_y = shear (examples/box-3d/step02_sheardisp.cfg).
    Returns (local dof indices, values)."""
    x = node_coords[:, 0]
    scale = max(abs(lo_x), abs(hi_x), 1.0)
    neg = np.nonzero(np.abs(x - lo_x) <= tol * scale)[0]
    pos = np.nonzero(np.abs(x - hi_x) <= tol * scale)[0]
    dofs = np.concatenate([3 * neg, 3 * neg + 1, 3 * neg + 2, 3 * pos + 1]).astype(np.int32)
    vals = np.concatenate([np.zeros(3 * neg.size), np.full(pos.size, shear)])
    order = np.argsort(dofs, kind="stable")
    return dofs[order], vals[order]


# ------------------------------------------------------------------------------------------
# Faults: cohesive cells on a through-going mesh surface (faults/TopologyOps.cc inserts them
# in PyLith; here the structured generator splits the nodes of the plane i = i_fault).
# ------------------------------------------------------------------------------------------
@dataclasses.dataclass
class Fault:
    closure_neg: np.ndarray   # [ncoh, 4] node ids on the negative side (face vertex order)
    closure_pos: np.ndarray   # [ncoh, 4] node ids on the positive side
    closure_lag: np.ndarray   # [ncoh, 4] Lagrange-multiplier node ids (0 .. num_lag - 1)
    face_cone: np.ndarray     # [ncoh, 4] vertex ids (negative side) for the face geometry
    neg_node: np.ndarray      # [num_lag] negative-side node of each Lagrange node
    pos_node: np.ndarray      # [num_lag] positive-side node of each Lagrange node
    lag_coords: np.ndarray    # [num_lag, 3]
    lag_jk: np.ndarray        # [num_lag, 2] (j, k) grid position on the fault
    colors: np.ndarray        # [ncoh] 4-colouring of the cohesive cells (no shared nodes)
    ref_dir_1: tuple = (0.0, 0.0, 1.0)   # FaultCohesive.py defaults
    ref_dir_2: tuple = (0.0, 1.0, 0.0)

    @property
    def num_lag(self):
        return self.neg_node.size

    @property
    def num_cohesive(self):
        return self.closure_neg.shape[0]


def hex_box_with_fault(n, lo, hi, i_fault, brick=8, shear=None):
    """Q1 box with a through-going fault on the vertex plane i = i_fault (0 < i_fault < nx).
    Nodes on the plane are split: the negative side (-x, cells i < i_fault) keeps the original
    ids, the positive side gets ids num_vertices + (fault vertex index).  Returns (mesh, fault).
    The face vertex order makes d x/d xi (along +y) cross d x/d eta (along +z) point from the
    negative to the positive side."""
    n = (n, n, n) if np.isscalar(n) else tuple(n)
    nx, ny, nz = n
    if not 0 < i_fault < nx:
        raise ValueError("the fault plane must be interior")
    m = hex_box(n, lo, hi, brick=brick, shear=shear)
    vx, vy, vz = nx + 1, ny + 1, nz + 1
    vrank = _brick_order((vx, vy, vz), brick)
    J, K = np.meshgrid(np.arange(vy), np.arange(vz), indexing="ij")
    J, K = J.ravel(order="F"), K.ravel(order="F")
    fault_v = vrank[i_fault + vx * (J + vy * K)]            # fault vertex ids (j fastest)
    nf = fault_v.size
    nv = m.num_vertices
    pos_id = nv + np.arange(nf)
    # cells on the positive side: cell index i >= i_fault
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.ravel(order="F")
    crank = _brick_order(n, brick)
    cell_i = np.empty(ci.size, np.int64)
    cell_i[crank] = ci
    remap = np.arange(nv + nf)
    remap[fault_v] = pos_id
    closure = m.closure.copy()
    posc = cell_i >= i_fault
    closure[posc] = remap[closure[posc]].astype(np.int32)
    coords = np.vstack([m.coords, m.coords[fault_v]])
    # cohesive cells on faces (j, k)
    fj, fk = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")
    fj, fk = fj.ravel(order="F"), fk.ravel(order="F")
    lag = lambda j, k: j + vy * k                           # Lagrange node = fault vertex index
    corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
    cl_lag = np.stack([lag(fj + a, fk + b) for a, b in corners], 1).astype(np.int32)
    cl_neg = fault_v[cl_lag].astype(np.int32)
    cl_pos = pos_id[cl_lag].astype(np.int32)
    colors = ((fj % 2) + 2 * (fk % 2)).astype(np.int32)
    mesh = Mesh(HEX_Q1, closure, closure.copy(), coords, coords, n,
                np.zeros(closure.shape[0], np.int32), m.colors)
    fault = Fault(cl_neg, cl_pos, cl_lag, cl_neg.copy(), fault_v.astype(np.int32),
                  pos_id.astype(np.int32), m.coords[fault_v], np.stack([J, K], 1), colors)
    return mesh, fault


def fault_normal_frame(fault, coords):
    """Unit normal of the (planar) fault and the tangent directions of
    BoundaryDirections::tangential_directions (ref dirs of the fault)."""
    X = coords[fault.face_cone[0]]
    dxi = 0.25 * (-X[0] + X[1] + X[2] - X[3])
    deta = 0.25 * (-X[0] - X[1] + X[2] + X[3])
    nrm = np.cross(dxi, deta)
    nrm /= np.linalg.norm(nrm)
    r1, r2 = np.asarray(fault.ref_dir_1, float), np.asarray(fault.ref_dir_2, float)
    r = r1 if abs(r1 @ nrm) <= 0.999 else r2
    t1 = np.cross(r, nrm)
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(nrm, t1)
    t2 /= np.linalg.norm(t2)
    if r2[0] * nrm[0] + r2[1] * nrm[1] + r2[1] * nrm[1] < 0.0:  # as written at BoundaryDirections.hh:76
        t2 = -t2
    return nrm, t1, t2


def tapered_patch_slip(fault, center, half_width, slip, component=1, taper=None):
    """Slip [num_lag, 3] (opening, left-lateral, reverse): `slip` on the square patch of the
    fault (in fault-plane coordinates (y, z)) |y - yc| <= a, |z - zc| <= a, tapering linearly to
    zero over `taper` outside it."""
    taper = half_width if taper is None else taper
    y, z = fault.lag_coords[:, 1], fault.lag_coords[:, 2]
    dy = np.maximum(np.abs(y - center[0]) - half_width, 0.0)
    dz = np.maximum(np.abs(z - center[1]) - half_width, 0.0)
    w = np.clip(1.0 - np.maximum(dy, dz) / taper, 0.0, 1.0)
    s = np.zeros((fault.num_lag, 3))
    s[:, component] = slip * w
    return s


def fault_vertex_impulses(fault, j_range, k_range, component=1, amplitude=1.0):
    """Unit-slip impulses of FaultCohesiveImpulses (faults/FaultCohesiveImpulses.cc:271-285): one
    impulse per fault vertex of the region j in j_range, k in k_range (fault grid positions),
    `amplitude` on slip component `component` (1 = left-lateral).  Returns (slip
    [num_impulses, num_lag, 3], the Lagrange-node id of each impulse)."""
    jk = fault.lag_jk
    sel = np.nonzero((jk[:, 0] >= j_range[0]) & (jk[:, 0] < j_range[1]) &
                     (jk[:, 1] >= k_range[0]) & (jk[:, 1] < k_range[1]))[0]
    order = np.lexsort((jk[sel, 0], jk[sel, 1]))  # strike fastest
    sel = sel[order]
    slip = np.zeros((sel.size, fault.num_lag, 3))
    slip[np.arange(sel.size), sel, component] = amplitude
    return slip, sel
