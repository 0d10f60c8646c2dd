"""FEM side of the coupled FEM-BEM solve, on the host (NumPy / SciPy sparse).

The reference keeps its FEM on the CPU in every backend (the ROCm coupled path is "CPU sparse Schur + GPU
dense LU", src/blab/solvers/julia_local/coupled_solver.jl:2094-2096); so does this build. Restated from
src/blab/solvers/julia_local/src/BeatEngineCoupled.jl:

* ``VolumeMesh`` / ``read_gmsh41_volume``      :51-100, 214-243, 2630-2757 (Gmsh 4.1 ASCII, tetrahedra + boundary
                                                 triangles, physical tags of their entities)
* ``restrict_volume_mesh``                    :143-205
* ``p1_fem_matrices``                         :260-312 (P1 stiffness and consistent mass)
* ``dynamic_stiffness``                       :441-462 (K - k^2 M - i k^2 M_loss)
* ``miki_admittance``                         :464-494 (rigid-backed Miki lining, exp(-i w t) convention)
* ``boundary_mass_matrix``                    :544-578
* ``prescribed_velocity_load``                :580-609
* ``outward_boundary_normals``                :2770-2803
* ``build_conforming_interface_map``          :658-719
* ``interface_operators``                     :721-773
* ``transducer_operators``                    :775-914

Everything is computed in float64 and vectorised over elements; duplicate (row, column) entries are summed
as Julia's ``sparse(I, J, V)`` sums them. Quadratic (10-node) tetrahedra are read but not assembled here: the
reference assembles them only in its interior-only FEM solve, which is not part of this build.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

_REF_GRAD = np.array([[-1.0, 1.0, 0.0, 0.0], [-1.0, 0.0, 1.0, 0.0], [-1.0, 0.0, 0.0, 1.0]])
_REF_MASS_TET = np.array([[2.0, 1, 1, 1], [1, 2, 1, 1], [1, 1, 2, 1], [1, 1, 1, 2]])
_REF_MASS_TRI = np.array([[2.0, 1, 1], [1, 2, 1], [1, 1, 2]])
_NODES_PER_TYPE = {1: 2, 2: 3, 4: 4, 9: 6, 11: 10, 15: 1}


@dataclass
class VolumeMesh:
    """Tetrahedral volume mesh (0-based indices): tetrahedra with their physical volume tags and boundary
    triangles with their physical surface tags."""

    vertices: np.ndarray  # (n, 3)
    tetrahedra: np.ndarray  # (T, 4)
    tetra_tags: np.ndarray  # (T,)
    boundary_faces: np.ndarray  # (B, 3)
    boundary_tags: np.ndarray  # (B,)
    physical_names: dict = field(default_factory=dict)  # (dimension, tag) -> name
    quadratic_tetrahedra: np.ndarray | None = None  # (T, 10) when the mesh is second order
    quadratic_boundary_faces: np.ndarray | None = None  # (B, 6)

    def __post_init__(self):
        self.vertices = np.asarray(self.vertices, dtype=np.float64).reshape(-1, 3)
        self.tetrahedra = np.asarray(self.tetrahedra, dtype=np.int64).reshape(-1, 4)
        self.tetra_tags = np.asarray(self.tetra_tags, dtype=np.int64).reshape(-1)
        self.boundary_faces = np.asarray(self.boundary_faces, dtype=np.int64).reshape(-1, 3)
        self.boundary_tags = np.asarray(self.boundary_tags, dtype=np.int64).reshape(-1)

    @property
    def n_vertices(self) -> int:
        return len(self.vertices)

    @property
    def is_quadratic(self) -> bool:
        return self.quadratic_tetrahedra is not None and len(self.quadratic_tetrahedra) > 0

    def physical_tag(self, dimension: int, name: str) -> int:
        """BeatEngineCoupled.jl:245-258."""
        for (d, tag), nm in self.physical_names.items():
            if d == dimension and nm == name:
                return int(tag)
        avail = sorted(nm for (d, _), nm in self.physical_names.items() if d == dimension)
        raise ValueError(f"Physical group {name!r} was not found in dimension {dimension}. Available groups: "
                         f"{', '.join(avail) if avail else '(none)'}.")


# ---------------------------------------------------------------------------------------------
# Gmsh 4.1 ASCII reader
# ---------------------------------------------------------------------------------------------


def _section(lines, name):
    try:
        start = lines.index("$" + name)
    except ValueError:
        return None
    end = lines.index("$End" + name, start)
    return start, end


def read_gmsh41_volume(path, scale: float = 1.0, translation=(0.0, 0.0, 0.0)) -> VolumeMesh:
    """Tetrahedra (type 4, or 11 second order) and boundary triangles (type 2, or 9) of a Gmsh 4.1 ASCII mesh,
    each tagged with the first physical tag of its entity (0 when the entity has none); vertices in ascending
    node-tag order, scaled then translated (load_gmsh41_volume, translated_volume_mesh: coupled_solver.jl:24-38)."""
    with open(path, "r") as fh:
        lines = [ln.strip() for ln in fh.read().splitlines()]
    fmt = _section(lines, "MeshFormat")
    if fmt is None:
        raise ValueError("Gmsh mesh is missing $MeshFormat.")
    head = lines[fmt[0] + 1].split()
    if not head[0].startswith("4.1"):
        raise ValueError("Volume FEM import currently requires a Gmsh 4.1 ASCII mesh.")
    if len(head) < 2 or head[1] != "0":
        raise ValueError("Volume FEM import currently requires an ASCII Gmsh mesh.")
    names = {}
    pn = _section(lines, "PhysicalNames")
    if pn is not None:
        count = int(lines[pn[0] + 1])
        for ln in lines[pn[0] + 2 : pn[0] + 2 + count]:
            parts = ln.split(maxsplit=2)
            names[(int(parts[0]), int(parts[1]))] = parts[2].strip('"')
    ent = _section(lines, "Entities")
    if ent is None:
        raise ValueError("Gmsh 4.1 volume mesh is missing $Entities.")
    counts = [int(x) for x in lines[ent[0] + 1].split()]
    if len(counts) != 4:
        raise ValueError("Invalid Gmsh $Entities header.")
    entity_tag = {}
    li = ent[0] + 2
    for dim in range(4):
        for _ in range(counts[dim]):
            parts = lines[li].split()
            k = 4 if dim == 0 else 7
            if int(parts[k]) > 0:
                entity_tag[(dim, int(parts[0]))] = int(parts[k + 1])
            li += 1
    nodes = _section(lines, "Nodes")
    if nodes is None:
        raise ValueError("Gmsh volume mesh is missing $Nodes.")
    tok = np.array(" ".join(lines[nodes[0] + 1 : nodes[1]]).split(), dtype=np.float64)
    n_blocks, n_nodes = int(tok[0]), int(tok[1])
    cur = 4
    tags_all, xyz_all = [], []
    for _ in range(n_blocks):
        edim, param, nb = int(tok[cur]), int(tok[cur + 2]), int(tok[cur + 3])
        cur += 4
        tags_all.append(tok[cur : cur + nb].astype(np.int64))
        cur += nb
        width = 3 + (edim if param == 1 else 0)
        blk = tok[cur : cur + nb * width].reshape(nb, width)
        xyz_all.append(blk[:, :3])
        cur += nb * width
    tags = np.concatenate(tags_all) if tags_all else np.zeros(0, np.int64)
    xyz = np.concatenate(xyz_all) if xyz_all else np.zeros((0, 3))
    uniq, first = np.unique(tags, return_index=True)
    if len(uniq) != n_nodes:
        raise ValueError("Gmsh node count does not match parsed coordinates.")
    vertices = xyz[first] * float(scale) + np.asarray(translation, dtype=np.float64)[None, :]

    def index_of(node_tags):
        pos = np.searchsorted(uniq, node_tags)
        pos = np.minimum(pos, len(uniq) - 1)
        if not np.all(uniq[pos] == node_tags):
            raise ValueError("Gmsh element references an unknown node.")
        return pos

    els = _section(lines, "Elements")
    if els is None:
        raise ValueError("Gmsh volume mesh is missing $Elements.")
    tok = np.array(" ".join(lines[els[0] + 1 : els[1]]).split(), dtype=np.float64)
    n_blocks = int(tok[0])
    cur = 4
    parts = {2: [], 4: [], 9: [], 11: []}
    for _ in range(n_blocks):
        edim, etag, etype, ne = int(tok[cur]), int(tok[cur + 1]), int(tok[cur + 2]), int(tok[cur + 3])
        cur += 4
        npe = _NODES_PER_TYPE.get(etype, 0)
        if npe <= 0:
            raise ValueError(f"Unsupported Gmsh element type {etype} in FEM mesh.")
        blk = tok[cur : cur + ne * (1 + npe)].reshape(ne, 1 + npe)[:, 1:].astype(np.int64)
        cur += ne * (1 + npe)
        if etype in parts:
            parts[etype].append((index_of(blk), np.full(ne, entity_tag.get((edim, etag), 0), np.int64)))

    def cat(key, width):
        if not parts[key]:
            return np.zeros((0, width), np.int64), np.zeros(0, np.int64)
        return (np.concatenate([p[0] for p in parts[key]]), np.concatenate([p[1] for p in parts[key]]))

    tet, tet_tags = cat(4, 4)
    tri, tri_tags = cat(2, 3)
    tet2, tet2_tags = cat(11, 10)
    tri2, tri2_tags = cat(9, 6)
    if len(tet2) and len(tet):
        raise ValueError("FEM mesh mixes first- and second-order tetrahedra.")
    if len(tri2) and len(tri):
        raise ValueError("FEM mesh mixes first- and second-order boundary triangles.")
    if bool(len(tet2)) != bool(len(tri2)):
        raise ValueError("FEM mesh must use the same interpolation order for tetrahedra and boundary triangles.")
    quad_tet = quad_tri = None
    if len(tet2):
        quad_tet, quad_tri = tet2, tri2
        tet, tet_tags = tet2[:, :4], tet2_tags
        tri, tri_tags = tri2[:, :3], tri2_tags
    if not len(tet):
        raise ValueError("FEM volume mesh contains no supported tetrahedra.")
    if not len(tri):
        raise ValueError("FEM volume mesh contains no supported boundary triangles.")
    return VolumeMesh(vertices, tet, tet_tags, tri, tri_tags, names, quad_tet, quad_tri)


def write_gmsh41_volume(path, mesh: VolumeMesh, names: dict | None = None) -> None:
    """A Gmsh 4.1 ASCII file of a first-order volume mesh: one volume entity per tetrahedron tag and one surface
    entity per boundary tag (entity tag = physical tag), nodes numbered from 1. Used to build test inputs."""
    names = dict(names if names is not None else mesh.physical_names)
    vt = sorted(set(int(t) for t in mesh.tetra_tags))
    st = sorted(set(int(t) for t in mesh.boundary_tags))
    out = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat"]
    if names:
        out += ["$PhysicalNames", str(len(names))]
        out += [f'{d} {t} "{nm}"' for (d, t), nm in sorted(names.items())]
        out += ["$EndPhysicalNames"]
    lo, hi = mesh.vertices.min(0), mesh.vertices.max(0)
    box = " ".join(f"{x:.17g}" for x in (*lo, *hi))
    out += ["$Entities", f"0 0 {len(st)} {len(vt)}"]
    out += [f"{t} {box} 1 {t} 0" for t in st]
    out += [f"{t} {box} 1 {t} 0" for t in vt]
    out += ["$EndEntities", "$Nodes"]
    n = mesh.n_vertices
    out += [f"1 {n} 1 {n}", f"3 {vt[0]} 0 {n}"]
    out += [str(i + 1) for i in range(n)]
    out += [f"{x:.17g} {y:.17g} {z:.17g}" for x, y, z in mesh.vertices]
    out += ["$EndNodes", "$Elements"]
    blocks = []
    eid = 1
    for t in st:
        f = mesh.boundary_faces[mesh.boundary_tags == t] + 1
        rows = []
        for r in f:
            rows.append(f"{eid} {r[0]} {r[1]} {r[2]}")
            eid += 1
        blocks.append((f"2 {t} 2 {len(f)}", rows))
    for t in vt:
        f = mesh.tetrahedra[mesh.tetra_tags == t] + 1
        rows = []
        for r in f:
            rows.append(f"{eid} {r[0]} {r[1]} {r[2]} {r[3]}")
            eid += 1
        blocks.append((f"3 {t} 4 {len(f)}", rows))
    out += [f"{len(blocks)} {eid - 1} 1 {eid - 1}"]
    for head, rows in blocks:
        out.append(head)
        out += rows
    out += ["$EndElements"]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


# ---------------------------------------------------------------------------------------------
# topology
# ---------------------------------------------------------------------------------------------


def _face_keys(faces: np.ndarray, n: int) -> np.ndarray:
    s = np.sort(faces, axis=1)
    return (s[:, 0] * n + s[:, 1]) * n + s[:, 2]


def _tet_faces(tet: np.ndarray):
    """The four faces of every tetrahedron in the reference's order, (T, 4, 3), and the opposite vertices."""
    f = np.stack([tet[:, [0, 1, 2]], tet[:, [0, 1, 3]], tet[:, [0, 2, 3]], tet[:, [1, 2, 3]]], axis=1)
    opp = tet[:, [3, 2, 1, 0]]
    return f, opp


def restrict_volume_mesh(mesh: VolumeMesh, selected_tags):
    """The tetrahedra of the selected volume tags, the boundary triangles that are exterior facets of them, and
    the compacted vertex numbering (BeatEngineCoupled.jl:143-205). Returns (mesh, vertex_index_map,
    boundary_face_index_map), maps from old 0-based index to new."""
    tags = set(int(t) for t in selected_tags)
    if not tags:
        raise ValueError("Bounded FEM region must select at least one physical volume group.")
    sel = np.nonzero(np.isin(mesh.tetra_tags, list(tags)))[0]
    if not len(sel):
        raise ValueError("Selected FEM volume groups contain no tetrahedra. Requested tags: "
                         f"{', '.join(str(t) for t in sorted(tags))}.")
    tets = mesh.tetrahedra[sel]
    n = mesh.n_vertices
    f, _ = _tet_faces(tets)
    keys, cnt = np.unique(_face_keys(f.reshape(-1, 3), n), return_counts=True)
    exterior = keys[cnt == 1]
    bidx = np.nonzero(np.isin(_face_keys(mesh.boundary_faces, n), exterior))[0]
    conn = mesh.quadratic_tetrahedra[sel] if mesh.is_quadratic else tets
    active = np.unique(conn.reshape(-1))
    remap = np.full(n, -1, np.int64)
    remap[active] = np.arange(len(active))
    qt = qf = None
    if mesh.is_quadratic:
        qt = remap[mesh.quadratic_tetrahedra[sel]]
        qf = remap[mesh.quadratic_boundary_faces[bidx]]
    out = VolumeMesh(mesh.vertices[active], remap[tets], mesh.tetra_tags[sel], remap[mesh.boundary_faces[bidx]],
                     mesh.boundary_tags[bidx], dict(mesh.physical_names), qt, qf)
    vmap = {int(v): i for i, v in enumerate(active)}
    fmap = {int(b): i for i, b in enumerate(bidx)}
    return out, vmap, fmap


def outward_boundary_normals(mesh: VolumeMesh) -> np.ndarray:
    """Unit normal of every boundary triangle pointing out of its tetrahedron (BeatEngineCoupled.jl:2770-2803)."""
    n = mesh.n_vertices
    f, opp = _tet_faces(mesh.tetrahedra)
    k = _face_keys(f.reshape(-1, 3), n)
    order = np.argsort(k, kind="stable")
    ks = k[order]
    uniq, start, cnt = np.unique(ks, return_index=True, return_counts=True)
    # the last occurrence's opposite vertex, as the reference's dictionary keeps the last write
    last_opp = opp.reshape(-1)[order[start + cnt - 1]]
    bk = _face_keys(mesh.boundary_faces, n)
    pos = np.searchsorted(uniq, bk)
    pos = np.minimum(pos, len(uniq) - 1)
    if not np.all((uniq[pos] == bk) & (cnt[pos] == 1)):
        raise ValueError("FEM boundary triangle is not an exterior tetrahedral facet.")
    x = mesh.vertices[mesh.boundary_faces]
    nrm = np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0])
    mag = np.linalg.norm(nrm, axis=1)
    if np.any(~(mag > 0)):
        raise ValueError("Interface contains a degenerate triangle.")
    nrm = nrm / mag[:, None]
    inward = mesh.vertices[last_opp[pos]] - x.mean(axis=1)
    flip = np.einsum("ij,ij->i", nrm, inward) > 0
    nrm[flip] *= -1.0
    return nrm


def triangle_areas(vertices: np.ndarray, faces: np.ndarray) -> np.ndarray:
    x = vertices[faces]
    return 0.5 * np.linalg.norm(np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), axis=1)


def triangle_normals(vertices: np.ndarray, faces: np.ndarray) -> np.ndarray:
    x = vertices[faces]
    nrm = np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0])
    mag = np.linalg.norm(nrm, axis=1)
    if np.any(~(mag > 0)):
        raise ValueError("Interface contains a degenerate triangle.")
    return nrm / mag[:, None]


# ---------------------------------------------------------------------------------------------
# matrices
# ---------------------------------------------------------------------------------------------


def p1_fem_matrices(mesh: VolumeMesh, eps: float = np.finfo(np.float64).eps):
    """P1 stiffness K and consistent mass M (CSR, float64): per tetrahedron K_e = V G^T G with G the gradients of
    the barycentric functions, M_e = V / 20 (1 + delta_ij) (BeatEngineCoupled.jl:260-312). A tetrahedron with
    |det J| <= eps * max|J|^3 is degenerate (eps: machine epsilon of the request's precision)."""
    if mesh.is_quadratic:
        raise ValueError("Quadratic tetrahedra are currently supported only by the pure interior FEM solve.")
    x = mesh.vertices[mesh.tetrahedra]  # (T,4,3)
    J = np.transpose(x[:, 1:, :] - x[:, :1, :], (0, 2, 1))  # columns x1-x0, x2-x0, x3-x0
    det = np.linalg.det(J)
    scale = np.abs(J).max(axis=(1, 2))
    if np.any(~(np.abs(det) > eps * scale**3)):
        raise ValueError("FEM volume mesh contains a numerically degenerate tetrahedron.")
    vol = np.abs(det) / 6.0
    G = np.linalg.solve(np.transpose(J, (0, 2, 1)), np.broadcast_to(_REF_GRAD, (len(J), 3, 4)))  # (T,3,4)
    Ke = vol[:, None, None] * np.einsum("tci,tcj->tij", G, G)
    Me = (vol / 20.0)[:, None, None] * _REF_MASS_TET[None]
    rows = np.repeat(mesh.tetrahedra[:, :, None], 4, axis=2).reshape(-1)
    cols = np.repeat(mesh.tetrahedra[:, None, :], 4, axis=1).reshape(-1)
    n = mesh.n_vertices
    K = sp.coo_matrix((Ke.reshape(-1), (rows, cols)), shape=(n, n)).tocsr()
    M = sp.coo_matrix((Me.reshape(-1), (rows, cols)), shape=(n, n)).tocsr()
    K.sum_duplicates()
    M.sum_duplicates()
    return K, M


def dynamic_stiffness(K, M, k: float, bulk_loss_mass=None, bulk_loss_factor: float = 0.0):
    """K - k^2 M - i k^2 M_loss (M_loss = bulk_loss_mass, or bulk_loss_factor * M) in complex128
    (assemble_fem_dynamic_stiffness, BeatEngineCoupled.jl:441-462)."""
    if not (np.isfinite(k) and k >= 0.0):
        raise ValueError("FEM wavenumber must be finite and nonnegative.")
    if not (np.isfinite(bulk_loss_factor) and bulk_loss_factor >= 0.0):
        raise ValueError("FEM bulk loss factor must be finite and nonnegative.")
    k2 = float(k) ** 2
    A = K.astype(np.complex128) - k2 * M
    if bulk_loss_mass is None:
        return (A - 1j * bulk_loss_factor * k2 * M).tocsr()
    return (A - 1j * k2 * bulk_loss_mass).tocsr()


def miki_admittance(frequency_hz: float, sound_speed: float, density: float, thickness_m: float,
                    flow_resistivity: float) -> complex:
    """Normal surface admittance of a rigid-backed porous lining after Miki's model, written for exp(+i w t)
    and conjugated into this code's exp(-i w t) convention (BeatEngineCoupled.jl:464-494)."""
    if not frequency_hz > 0.0:
        raise ValueError("Miki wall impedance requires a positive frequency.")
    if not (sound_speed > 0.0 and density > 0.0):
        raise ValueError("Miki wall impedance requires a positive medium.")
    if not thickness_m > 0.0:
        raise ValueError("Wall-lining thickness must be positive.")
    if not flow_resistivity > 0.0:
        raise ValueError("Wall-lining airflow resistivity must be positive.")
    x = frequency_hz / flow_resistivity
    pz, pk = x ** -0.632, x ** -0.618
    zc = density * sound_speed * complex(1.0 + 0.070 * pz, -0.107 * pz)
    kp = (2.0 * np.pi * frequency_hz / sound_speed) * complex(1.0 + 0.109 * pk, -0.160 * pk)
    zs = -1j * zc / np.tan(kp * thickness_m)
    y = 1.0 / np.conj(zs)
    if not y.real >= 0.0:
        raise ValueError("Miki wall impedance produced a non-passive admittance.")
    return complex(y)


def boundary_mass_matrix(mesh: VolumeMesh, face_indices, boundary_vertex_indices):
    """(n_vertices x n_boundary) sparse P1 boundary mass: rows are mesh vertices, columns the positions of the
    face vertices in ``boundary_vertex_indices`` (BeatEngineCoupled.jl:544-578)."""
    if mesh.is_quadratic:
        raise ValueError("Quadratic tetrahedra are currently supported only by the pure interior FEM solve.")
    faces = mesh.boundary_faces[np.asarray(face_indices, dtype=np.int64)]
    bvi = np.asarray(boundary_vertex_indices, dtype=np.int64)
    col_of = np.full(mesh.n_vertices, -1, np.int64)
    col_of[bvi] = np.arange(len(bvi))  # the last position of a repeated vertex, as a dictionary keeps it
    if len(faces) and np.any(col_of[faces] < 0):
        raise ValueError("Boundary face references a vertex outside the requested boundary space.")
    area = triangle_areas(mesh.vertices, faces)
    loc = (area / 12.0)[:, None, None] * _REF_MASS_TRI[None]
    rows = np.repeat(faces[:, :, None], 3, axis=2).reshape(-1)
    cols = col_of[np.repeat(faces[:, None, :], 3, axis=1).reshape(-1)]
    out = sp.coo_matrix((loc.reshape(-1), (rows, cols)), shape=(mesh.n_vertices, len(bvi))).tocsr()
    out.sum_duplicates()
    return out


def prescribed_velocity_load(mesh: VolumeMesh, tag: int, density: float, omega: float,
                             velocity: complex) -> np.ndarray:
    """f_v += i rho omega v * area / 3 for every vertex of the faces tagged ``tag`` (BeatEngineCoupled.jl:580-609)."""
    sel = mesh.boundary_tags == int(tag)
    faces = mesh.boundary_faces[sel]
    load = np.zeros(mesh.n_vertices, dtype=np.complex128)
    val = (1j * density * omega * complex(velocity)) * triangle_areas(mesh.vertices, faces) / 3.0
    for c in range(3):
        np.add.at(load, faces[:, c], val)
    return load


# ---------------------------------------------------------------------------------------------
# FEM-BEM interface and transducers
# ---------------------------------------------------------------------------------------------


@dataclass
class InterfaceMap:
    """ConformingInterfaceMap (BeatEngineCoupled.jl:102-124), 0-based."""

    fem_vertex_indices: np.ndarray
    fem_to_bem_vertex_indices: np.ndarray
    fem_face_indices: np.ndarray
    bem_face_indices: np.ndarray
    normal_sign: np.ndarray

    def __post_init__(self):
        for name in ("fem_vertex_indices", "fem_to_bem_vertex_indices", "fem_face_indices", "bem_face_indices",
                     "normal_sign"):
            setattr(self, name, np.asarray(getattr(self, name), dtype=np.int64).reshape(-1))

    @property
    def count(self) -> int:
        return len(self.fem_vertex_indices)

    def offset(self, fem_vertex=0, fem_face=0, bem_vertex=0, bem_face=0) -> "InterfaceMap":
        return InterfaceMap(self.fem_vertex_indices + fem_vertex, self.fem_to_bem_vertex_indices + bem_vertex,
                            self.fem_face_indices + fem_face, self.bem_face_indices + bem_face, self.normal_sign)

    @staticmethod
    def concatenate(maps) -> "InterfaceMap":
        maps = list(maps)
        if not maps:
            z = np.zeros(0, np.int64)
            return InterfaceMap(z, z, z, z, z)
        return InterfaceMap(*(np.concatenate([getattr(m, n) for m in maps]) for n in (
            "fem_vertex_indices", "fem_to_bem_vertex_indices", "fem_face_indices", "bem_face_indices",
            "normal_sign")))


def build_conforming_interface_map(fem_mesh: VolumeMesh, bem_mesh, fem_tag: int, bem_tag: int,
                                   tol: float = 1e-10) -> InterfaceMap:
    """Match the FEM boundary triangles tagged fem_tag with the BEM triangles tagged bem_tag vertex by vertex
    (nearest within tol, one-to-one) and face by face (same vertex set); normal_sign = +1 where the FEM face's
    vertex-order normal agrees with the BEM normal (BeatEngineCoupled.jl:658-719)."""
    ff = np.nonzero(fem_mesh.boundary_tags == int(fem_tag))[0]
    bf = np.nonzero(np.asarray(bem_mesh.tags) == int(bem_tag))[0]
    if len(ff) != len(bf):
        raise ValueError(f"FEM and BEM interface triangle counts differ: {len(ff)} and {len(bf)}.")
    fv = np.unique(fem_mesh.boundary_faces[ff].reshape(-1))
    bv = np.unique(np.asarray(bem_mesh.faces)[bf].reshape(-1))
    if len(fv) != len(bv):
        raise ValueError(f"FEM and BEM interface vertex counts differ: {len(fv)} and {len(bv)}.")
    mapped = np.empty(len(fv), np.int64)
    used = set()
    bx = np.asarray(bem_mesh.vertices)[bv]
    for i, v in enumerate(fv):
        d = np.linalg.norm(bx - fem_mesh.vertices[v][None, :], axis=1)
        j = int(np.argmin(d))
        if not d[j] <= tol:
            raise ValueError(f"FEM interface vertex {v + 1} has no BEM match within {tol} m.")
        if int(bv[j]) in used:
            raise ValueError("FEM-to-BEM interface vertex mapping is not one-to-one.")
        used.add(int(bv[j]))
        mapped[i] = bv[j]
    to_bem = dict(zip(fv.tolist(), mapped.tolist()))
    bfaces = np.asarray(bem_mesh.faces)
    by_key = {tuple(sorted(bfaces[i].tolist())): int(i) for i in bf}
    fn = triangle_normals(fem_mesh.vertices, fem_mesh.boundary_faces[ff]) if len(ff) else np.zeros((0, 3))
    bn = np.asarray(bem_mesh.normals)
    out_faces, signs = [], []
    for r, fi in enumerate(ff):
        key = tuple(sorted(to_bem[int(v)] for v in fem_mesh.boundary_faces[fi]))
        b = by_key.get(key, -1)
        if b < 0:
            raise ValueError("FEM and BEM interface triangle connectivity differs.")
        out_faces.append(b)
        signs.append(1 if float(np.dot(fn[r], bn[b])) >= 0.0 else -1)
    return InterfaceMap(fv, mapped, ff, np.asarray(out_faces, np.int64), np.asarray(signs, np.int64))


@dataclass
class InterfaceOperators:
    fem_load: sp.csr_matrix  # (n_fem, n_iface) boundary mass on the interface
    bem_flux: sp.csr_matrix  # (F_bem, n_iface) DP0 Neumann data of each interface flux dof
    fem_trace: sp.csr_matrix  # (n_iface, n_fem)
    bem_trace: sp.csr_matrix  # (n_iface, V_bem)


def interface_operators(fem_mesh: VolumeMesh, bem_mesh, imap: InterfaceMap) -> InterfaceOperators:
    """BeatEngineCoupled.jl:721-773."""
    fem_load = boundary_mass_matrix(fem_mesh, imap.fem_face_indices, imap.fem_vertex_indices)
    n_if = imap.count
    dof = np.full(fem_mesh.n_vertices, -1, np.int64)
    dof[imap.fem_vertex_indices] = np.arange(n_if)
    faces = fem_mesh.boundary_faces[imap.fem_face_indices]
    rows = np.repeat(imap.bem_face_indices, 3)
    cols = dof[faces.reshape(-1)]
    vals = np.repeat(imap.normal_sign.astype(np.float64) / 3.0, 3)
    F = len(np.asarray(bem_mesh.faces))
    bem_flux = sp.coo_matrix((vals, (rows, cols)), shape=(F, n_if)).tocsr()
    bem_flux.sum_duplicates()
    fem_trace = sp.coo_matrix((np.ones(n_if), (np.arange(n_if), imap.fem_vertex_indices)),
                              shape=(n_if, fem_mesh.n_vertices)).tocsr()
    bem_trace = sp.coo_matrix((np.ones(n_if), (np.arange(n_if), imap.fem_to_bem_vertex_indices)),
                              shape=(n_if, len(np.asarray(bem_mesh.vertices)))).tocsr()
    return InterfaceOperators(fem_load, bem_flux, fem_trace, bem_trace)


@dataclass
class Transducer:
    """ElectrodynamicTransducer (BeatEngineCoupled.jl:126-141): a rigid piston on FEM and/or BEM moving
    boundaries (tags with signed weights) with its Thiele-Small parameters."""

    id: str
    fem_tags: list
    fem_signs: list
    bem_tags: list
    bem_signs: list
    motion_axis: np.ndarray
    surface_completion_factor: float
    physical_driver_orbit_count: int
    re_ohm: float
    le_h: float
    bl_n_per_a: float
    mmd_kg: float
    cms_m_per_n: float
    rms_n_s_per_m: float


def transducer_operators(fem_mesh: VolumeMesh, bem_mesh, transducers) -> dict:
    """Per transducer column: FEM nodal projected areas (fem_surface) and their force version (times the surface
    completion factor), the same on the BEM side, and the BEM faces' normal velocity per unit piston velocity
    (BeatEngineCoupled.jl:775-914)."""
    n_fem = fem_mesh.n_vertices
    bv = np.asarray(bem_mesh.vertices)
    bfaces = np.asarray(bem_mesh.faces)
    btags = np.asarray(bem_mesh.tags)
    V, F, nt = len(bv), len(bfaces), len(transducers)
    if nt == 0:
        return {"fem_surface": sp.csr_matrix((n_fem, 0)), "fem_force": sp.csr_matrix((n_fem, 0)),
                "bem_surface": sp.csr_matrix((V, 0)), "bem_force": sp.csr_matrix((V, 0)),
                "bem_normal_velocity": sp.csr_matrix((F, 0))}
    normals = outward_boundary_normals(fem_mesh)
    fem_area = triangle_areas(fem_mesh.vertices, fem_mesh.boundary_faces)
    bem_n = np.asarray(bem_mesh.normals)
    bem_area = np.asarray(bem_mesh.areas)
    fr, fc, fv, ffv = [], [], [], []
    br, bc, bvv, bfv = [], [], [], []
    vr, vc, vv = [], [], []
    for col, td in enumerate(transducers):
        if len(td.fem_tags) != len(td.fem_signs):
            raise ValueError(f"Transducer {td.id!r} FEM boundary tags and signs differ in length.")
        if len(td.bem_tags) != len(td.bem_signs):
            raise ValueError(f"Transducer {td.id!r} BEM boundary tags and signs differ in length.")
        if not td.fem_tags and not td.bem_tags:
            raise ValueError(f"Transducer {td.id!r} has no acoustic moving boundary.")
        axis = np.asarray(td.motion_axis, dtype=np.float64)
        na = float(np.linalg.norm(axis))
        if not na > 0.0:
            raise ValueError(f"Transducer {td.id!r} has a zero motion axis.")
        axis = axis / na
        if not td.surface_completion_factor >= 1.0:
            raise ValueError(f"Transducer {td.id!r} has an invalid surface completion factor.")
        if not td.physical_driver_orbit_count >= 1:
            raise ValueError(f"Transducer {td.id!r} has an invalid physical driver orbit count.")
        fsign = dict(zip([int(t) for t in td.fem_tags], [float(s) for s in td.fem_signs]))
        if len(fsign) != len(td.fem_tags):
            raise ValueError(f"Transducer {td.id!r} repeats a FEM boundary tag.")
        sign = np.array([fsign.get(int(t), 0.0) for t in fem_mesh.boundary_tags])
        sel = np.nonzero(sign != 0.0)[0]
        missing = set(fsign) - set(int(t) for t in fem_mesh.boundary_tags[sel])
        if missing:
            raise ValueError(f"Transducer {td.id!r} FEM boundary tags contain no selected faces: "
                             + ", ".join(str(t) for t in sorted(missing)))
        proj = sign[sel] * (normals[sel] @ axis)
        nodal = proj * fem_area[sel] / 3.0
        fr.append(fem_mesh.boundary_faces[sel].reshape(-1))
        fc.append(np.full(3 * len(sel), col))
        fv.append(np.repeat(nodal, 3))
        ffv.append(np.repeat(td.surface_completion_factor * nodal, 3))
        bsign_map = dict(zip([int(t) for t in td.bem_tags], [float(s) for s in td.bem_signs]))
        if len(bsign_map) != len(td.bem_tags):
            raise ValueError(f"Transducer {td.id!r} repeats a BEM boundary tag.")
        bsign = np.array([bsign_map.get(int(t), 0.0) for t in btags])
        bsel = np.nonzero(bsign != 0.0)[0]
        missing = set(bsign_map) - set(int(t) for t in btags[bsel])
        if missing:
            raise ValueError(f"Transducer {td.id!r} BEM boundary tags contain no faces: "
                             + ", ".join(str(t) for t in sorted(missing)))
        bproj = bsign[bsel] * (bem_n[bsel] @ axis)
        bnodal = bproj * bem_area[bsel] / 3.0
        br.append(bfaces[bsel].reshape(-1))
        bc.append(np.full(3 * len(bsel), col))
        bvv.append(np.repeat(bnodal, 3))
        bfv.append(np.repeat(td.surface_completion_factor * bnodal, 3))
        vr.append(bsel)
        vc.append(np.full(len(bsel), col))
        vv.append(bproj)

    def mk(vals, rows, cols, shape):
        m = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=shape).tocsr()
        m.sum_duplicates()
        return m

    return {"fem_surface": mk(fv, fr, fc, (n_fem, nt)), "fem_force": mk(ffv, fr, fc, (n_fem, nt)),
            "bem_surface": mk(bvv, br, bc, (V, nt)), "bem_force": mk(bfv, br, bc, (V, nt)),
            "bem_normal_velocity": mk(vv, vr, vc, (F, nt))}
