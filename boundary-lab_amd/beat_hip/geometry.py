"""Host-side mesh preparation for the HIP solver (once per mesh, vectorised NumPy).

Restates the reference's mesh / space / quadrature set-up in the layout the gfx950 kernels consume:

* ``BoundaryMesh``           BeatEngineCore.jl:122-130,236-256 (normals from cross(v2-v1, v3-v1), no welding)
* Gmsh 2.2 reader            BeatEngineCore.jl:193-234, combined meshes solver.jl:252-269
* order-4 triangle rule      BeatEngineCore.jl:517-549 (the same tabulated constants)
* Duffy rules and remaps     BeatEngineCore.jl:551-655,751-771,820-873
* vertex-sharing pair cache  BeatEngineCore.jl:441-479,784-811,875-915
* surface curls              BeatEngineCore.jl:685-702
* mass blocks                BeatEngineCore.jl:731-749

New for the HIP layout: trial elements are put in Morton (Z-curve) order of their centroids and cut into
chunks of 64 (one lane each). Each chunk's distinct vertices are "slots"; the regular kernel adds each slot's
value into that vertex's column of the test-element strip row (``slot_vertex``; a vertex shared by several chunks
keeps its partial sums in LDS until its last chunk stores them, ``slot_pend``), and the row gather sums the strip
rows of each P1 row's star (``vstar_ptr/vstar_row``).
"""

from __future__ import annotations

import concurrent.futures
import math
import threading
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------------------------
# Mesh
# ---------------------------------------------------------------------------------------------


@dataclass
class BoundaryMesh:
    vertices: np.ndarray  # (V,3) float64 [m]
    faces: np.ndarray  # (F,3) int64 0-based
    tags: np.ndarray  # (F,) physical tags
    mesh_ids: np.ndarray = None  # (F,) 1-based input-mesh id
    normals: np.ndarray = field(init=False)
    areas: np.ndarray = field(init=False)

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float64)
        self.faces = np.ascontiguousarray(self.faces, dtype=np.int64)
        self.tags = np.asarray(self.tags, dtype=np.int64)
        if self.mesh_ids is None:
            self.mesh_ids = np.ones(len(self.faces), dtype=np.int64)
        fv = self.vertices[self.faces]
        cross = np.cross(fv[:, 1] - fv[:, 0], fv[:, 2] - fv[:, 0])
        norm = np.linalg.norm(cross, axis=1)
        if np.any(norm <= 0.0):
            raise ValueError("mesh contains degenerate (zero-area) triangles")
        self.areas = norm / 2.0
        self.normals = cross / norm[:, None]
        self.face_vertices = fv

    @property
    def n_faces(self) -> int:
        return int(self.faces.shape[0])

    @property
    def n_vertices(self) -> int:
        return int(self.vertices.shape[0])


def read_gmsh22(path, scale: float = 1.0, translation=(0.0, 0.0, 0.0)):
    """Gmsh 2.2 ASCII: nodes and type-2 (triangle) elements; tag = first tag token (BeatEngineCore.jl:193-234)."""
    with open(path, "r", encoding="utf-8") as handle:
        text = handle.read()
    fast = _gmsh22_text(text, scale)
    if fast is not None:
        v, faces, tags = fast
        return v + np.asarray(translation, dtype=np.float64), faces, tags
    lines = text.splitlines()
    stripped = [line.strip() for line in lines]
    try:
        ns, ne = stripped.index("$Nodes"), stripped.index("$EndNodes")
        es, ee = stripped.index("$Elements"), stripped.index("$EndElements")
    except ValueError as exc:
        raise ValueError("Only Gmsh 2.2 ASCII meshes with Nodes and Elements sections are supported.") from exc
    fast = _gmsh22_blocks(lines[ns + 2 : ne], lines[es + 2 : ee], scale)
    if fast is not None:
        v, faces, tags = fast
        return v + np.asarray(translation, dtype=np.float64), faces, tags
    node_index = {}
    coords = []
    for line in lines[ns + 2 : ne]:
        parts = line.split()
        if len(parts) < 4:
            continue
        node_index[int(parts[0])] = len(coords)
        coords.append((float(parts[1]) * scale, float(parts[2]) * scale, float(parts[3]) * scale))
    faces, tags = [], []
    for line in lines[es + 2 : ee]:
        parts = line.split()
        if len(parts) < 8 or parts[1] != "2":
            continue
        a = node_index.get(int(parts[-3]))
        b = node_index.get(int(parts[-2]))
        c = node_index.get(int(parts[-1]))
        if a is None or b is None or c is None:
            continue
        faces.append((a, b, c))
        tags.append(int(parts[3]))
    v = np.asarray(coords, dtype=np.float64).reshape(-1, 3) + np.asarray(translation, dtype=np.float64)
    return v, np.asarray(faces, dtype=np.int64).reshape(-1, 3), np.asarray(tags, dtype=np.int64)


def _uniform_block(lines, dtype):
    """The whitespace-separated numbers of `lines` as a (len(lines), k) array when every line holds k of them,
    else None."""
    if not lines:
        return None
    k = len(lines[0].split())
    flat = np.fromstring(" ".join(lines), dtype=dtype, sep=" ")  # noqa: NPY201 (text mode)
    if k == 0 or flat.size != k * len(lines):
        return None
    return flat.reshape(len(lines), k)


def _marker_line(text, name):
    """(start, end) of the first line of `text` that is `$name` between blanks, or None."""
    key = "$" + name
    i = text.find(key)
    while i >= 0:
        ls = text.rfind("\n", 0, i) + 1
        le = text.find("\n", i)
        le = len(text) if le < 0 else le
        if not text[ls:i].strip() and not text[i + len(key) : le].strip():
            return ls, le
        i = text.find(key, i + 1)
    return None


def _text_block(text, begin, end):
    """The data lines of a section (after its marker and count lines, before its end marker) as one string and
    their number, or None."""
    mb, me = _marker_line(text, begin), _marker_line(text, end)
    if mb is None or me is None:
        return None
    count_end = text.find("\n", mb[1] + 1)  # the count line follows the marker line
    if count_end < 0:
        return None
    if count_end >= me[0]:
        return "", 0
    block = text[count_end + 1 : me[0]]
    return block, block.count("\n")


def _uniform_text(block, nlines, dtype):
    """_uniform_block for a section given as one string of nlines lines."""
    if nlines == 0:
        return None
    first = block[: block.find("\n")]
    k = len(first.split())
    flat = np.fromstring(block, dtype=dtype, sep=" ")  # noqa: NPY201 (text mode)
    if k == 0 or flat.size != k * nlines:
        return None
    return flat.reshape(nlines, k)


def _gmsh22_text(text, scale):
    """The vectorised read of _gmsh22_blocks on the whole file text (no split into lines); None when a section is
    missing or irregular (the caller then reads line by line and reports what is wrong)."""
    import warnings

    nb, eb = _text_block(text, "Nodes", "EndNodes"), _text_block(text, "Elements", "EndElements")
    if nb is None or eb is None:
        return None
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", DeprecationWarning)
        nodes = _uniform_text(nb[0], nb[1], np.float64)
        elems = _uniform_text(eb[0], eb[1], np.int64) if eb[1] else np.zeros((0, 8), np.int64)
    return _gmsh22_arrays(nodes, elems, scale)


def _gmsh22_blocks(node_lines, elem_lines, scale):
    """Vectorised read of the Nodes and Elements blocks (the same semantics as the line loop of read_gmsh22:
    triangles only, element nodes are the last three tokens, tag is the first tag token). None when a block has
    lines of different lengths or an element names an unknown node; the caller then reads line by line."""
    import warnings

    with warnings.catch_warnings():
        warnings.simplefilter("ignore", DeprecationWarning)
        nodes = _uniform_block(node_lines, np.float64)
        elems = _uniform_block(elem_lines, np.int64) if elem_lines else np.zeros((0, 8), np.int64)
    return _gmsh22_arrays(nodes, elems, scale)


def _gmsh22_arrays(nodes, elems, scale):
    if nodes is None or elems is None or nodes.shape[1] != 4 or elems.shape[1] < 8:
        return None
    ids = nodes[:, 0].astype(np.int64)
    if not np.array_equal(ids, nodes[:, 0]):
        return None
    tri = elems[elems[:, 1] == 2]
    order = np.argsort(ids, kind="stable")
    sorted_ids = ids[order]
    if len(sorted_ids) > 1 and (np.diff(sorted_ids) == 0).any():
        return None  # repeated node ids: the line loop keeps the last definition
    ref = tri[:, -3:]
    pos = np.searchsorted(sorted_ids, ref)
    pos_c = np.minimum(pos, len(sorted_ids) - 1)
    if len(ref) and not (sorted_ids[pos_c] == ref).all():
        return None
    faces = order[pos_c].reshape(-1, 3)
    return nodes[:, 1:4] * scale, faces.astype(np.int64), tri[:, 3].astype(np.int64)


def combine(parts) -> BoundaryMesh:
    """Concatenate (V, F, tags) parts without welding; face mesh ids are 1-based part indices."""
    vs, fs, ts, ids = [], [], [], []
    off = 0
    for i, (v, f, t) in enumerate(parts, start=1):
        vs.append(np.asarray(v, dtype=np.float64))
        fs.append(np.asarray(f, dtype=np.int64) + off)
        ts.append(np.asarray(t, dtype=np.int64))
        ids.append(np.full(len(f), i, dtype=np.int64))
        off += len(v)
    return BoundaryMesh(np.concatenate(vs), np.concatenate(fs), np.concatenate(ts), np.concatenate(ids))


# ---------------------------------------------------------------------------------------------
# Quadrature
# ---------------------------------------------------------------------------------------------

# Order-4 six-point rule on the reference triangle (xi, eta); weights sum to 1/2.
# Positions closer than this (model units, metres) are the same point for image-pair detection
# (BeatEngineCore.jl:917, geometric tolerance of image_singular_candidates).
GEOMETRY_TOLERANCE = 1e-8

RULE_POINTS = np.array(
    [
        [0.4459484909159651, 0.4459484909159651],
        [0.0915762135097710, 0.0915762135097700],
        [0.1081030181680700, 0.4459484909159651],
        [0.4459484909159651, 0.1081030181680700],
        [0.8168475729804590, 0.0915762135097700],
        [0.0915762135097710, 0.8168475729804580],
    ]
)
RULE_WEIGHTS = 0.5 * np.array(
    [0.2233815896780110, 0.1099517436553220, 0.2233815896780110, 0.2233815896780110, 0.1099517436553220,
     0.1099517436553220]
)


def p1_basis(points):
    p = np.asarray(points, dtype=np.float64)
    return np.stack([1.0 - p[..., 0] - p[..., 1], p[..., 0], p[..., 1]], axis=-1)


def gauss01(order: int):
    """Gauss-Legendre nodes/weights on [0,1] for orders 1..4."""
    if order == 1:
        return np.array([0.5]), np.array([1.0])
    if order == 2:
        a = 0.5 / math.sqrt(3.0)
        return np.array([0.5 - a, 0.5 + a]), np.array([0.5, 0.5])
    if order == 3:
        a = math.sqrt(0.6) / 2.0
        return np.array([0.5 - a, 0.5, 0.5 + a]), np.array([5.0 / 18.0, 4.0 / 9.0, 5.0 / 18.0])
    if order == 4:
        inner = 2.0 / 7.0 * math.sqrt(1.2)
        x1 = math.sqrt(3.0 / 7.0 - inner) / 2.0
        x2 = math.sqrt(3.0 / 7.0 + inner) / 2.0
        w1 = (18.0 + math.sqrt(30.0)) / 72.0
        w2 = (18.0 - math.sqrt(30.0)) / 72.0
        return np.array([0.5 - x2, 0.5 - x1, 0.5 + x1, 0.5 + x2]), np.array([w2, w1, w1, w2])
    raise ValueError("singular_order must be 1..4")


def duffy_base(order: int, kind: int):
    """Duffy rule on T x T for kind 1 (coincident, 6 maps), 2 (edge, 5), 3 (vertex, 2).

    Returns (test (n,2), trial (n,2), w (n,)). The simplex maps are those of the reference
    (BeatEngineCore.jl:590-620); points are converted with (x, y) -> (x - y, y).
    """
    x, w = gauss01(order)
    # tensor grid: outer index i -> second coordinate, inner j -> first coordinate
    gx = np.tile(x, len(x))
    gy = np.repeat(x, len(x))
    gw = np.repeat(w, len(x)) * np.tile(w, len(x))
    xsi = np.repeat(gx, len(gx))
    e1 = np.repeat(gy, len(gx))
    e2 = np.tile(gx, len(gx))
    e3 = np.tile(gy, len(gx))
    base = np.repeat(gw, len(gx)) * np.tile(gw, len(gx))
    e12 = e1 * e2
    e123 = e12 * e3
    one = np.ones_like(xsi)
    if kind == 1:
        wt = base * xsi**3 * e1**2 * e2
        maps = [
            (xsi, xsi * (1 - e1 + e12), xsi * (1 - e123), xsi * (1 - e1), wt),
            (xsi * (1 - e123), xsi * (1 - e1), xsi, xsi * (1 - e1 + e12), wt),
            (xsi, xsi * (e1 - e12 + e123), xsi * (1 - e12), xsi * (e1 - e12), wt),
            (xsi * (1 - e12), xsi * (e1 - e12), xsi, xsi * (e1 - e12 + e123), wt),
            (xsi * (1 - e123), xsi * (e1 - e123), xsi, xsi * (e1 - e12), wt),
            (xsi, xsi * (e1 - e12), xsi * (1 - e123), xsi * (e1 - e123), wt),
        ]
    elif kind == 2:
        wt = base * xsi**3 * e1**2
        maps = [
            (xsi, xsi * e1 * e3, xsi * (1 - e12), xsi * e1 * (1 - e2), wt),
            (xsi, xsi * e1, xsi * (1 - e123), xsi * e12 * (1 - e3), wt * e2),
            (xsi * (1 - e12), xsi * e1 * (1 - e2), xsi * one, xsi * e123, wt * e2),
            (xsi * (1 - e123), xsi * e12 * (1 - e3), xsi * one, xsi * e1, wt * e2),
            (xsi * (1 - e123), xsi * e1 * (1 - e2 * e3), xsi * one, xsi * e12, wt * e2),
        ]
    elif kind == 3:
        wt = base * xsi**3 * e2
        maps = [
            (xsi, xsi * e1, xsi * e2, xsi * e2 * e3, wt),
            (xsi * e2, xsi * e2 * e3, xsi * one, xsi * e1, wt),
        ]
    else:
        raise ValueError(kind)
    # interleave maps per tensor point (point-major, map-minor) like the reference's push order
    tx = np.stack([m[0] for m in maps], axis=1).reshape(-1)
    ty = np.stack([m[1] for m in maps], axis=1).reshape(-1)
    sx = np.stack([m[2] for m in maps], axis=1).reshape(-1)
    sy = np.stack([m[3] for m in maps], axis=1).reshape(-1)
    ww = np.stack([m[4] for m in maps], axis=1).reshape(-1)
    return np.stack([tx - ty, ty], axis=1), np.stack([sx - sy, sy], axis=1), ww


_REF = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])


def remap_vertex(p, vid):
    """Move the Duffy singular corner (origin) to local vertex vid (0-based)."""
    if vid == 0:
        return p.copy()
    if vid == 1:
        return np.stack([1.0 - p[:, 0] - p[:, 1], p[:, 1]], axis=1)
    return np.stack([p[:, 0], 1.0 - p[:, 0] - p[:, 1]], axis=1)


def remap_edge(p, va, vb):
    """Affine map sending the reference edge (0,0)-(1,0) to local vertices va -> vb (0-based)."""
    vr = 3 - va - vb
    o, a, b = _REF[va], _REF[vb], _REF[vr]
    return o[None, :] + p[:, :1] * (a - o)[None, :] + p[:, 1:2] * (b - o)[None, :]


def regular_rule(order: int = 4):
    """Points (n, 2) and weights (n,) of the regular triangle rule of ``order`` (BeatEngineCore.jl:517-549): the
    centroid rule for order <= 1, the six-point rule for order 4, the three-point rule (1/6, 1/6), (2/3, 1/6),
    (1/6, 2/3) with weights 1/6 for any other order."""
    order = int(order)
    if order <= 1:
        return np.array([[1.0 / 3.0, 1.0 / 3.0]]), np.array([0.5])
    if order == 4:
        return RULE_POINTS, RULE_WEIGHTS
    return (np.array([[1.0 / 6.0, 1.0 / 6.0], [2.0 / 3.0, 1.0 / 6.0], [1.0 / 6.0, 2.0 / 3.0]]),
            np.full(3, 1.0 / 6.0))


def mass_blocks(order: int = 4):
    """P1 x P1 and P1 x DP0 identity blocks per unit 2A under the regular rule: the reference assembles its
    identity operators with that rule (assemble_l2_identity_matrix, BeatEngineCore.jl:1138-1180)."""
    pts, wts = regular_rule(order)
    b = p1_basis(pts)
    m11 = np.einsum("qi,qj,q->ij", b, b, wts)
    m10 = b.T @ wts
    return m11, m10


# ---------------------------------------------------------------------------------------------
# Per-face geometry
# ---------------------------------------------------------------------------------------------


def surface_curls(face_vertices, normals):
    """n x grad(phi_i) for the three P1 functions via the inverse Gram matrix. (F,3,3)."""
    e1 = face_vertices[:, 1] - face_vertices[:, 0]
    e2 = face_vertices[:, 2] - face_vertices[:, 0]
    a = (e1 * e1).sum(1)
    b = (e1 * e2).sum(1)
    d = (e2 * e2).sum(1)
    det = a * d - b * b
    g1 = (e1 * d[:, None] - e2 * b[:, None]) / det[:, None]
    g2 = (e2 * a[:, None] - e1 * b[:, None]) / det[:, None]
    grads = np.stack([-g1 - g2, g1, g2], axis=1)
    return np.cross(normals[:, None, :], grads)


def quadrature_points(face_vertices, order: int = 4):
    """(F, n, 3) points of the regular rule of ``order``."""
    return np.matmul(p1_basis(regular_rule(order)[0]), face_vertices)  # (n, 3) @ (F, 3, 3) -> (F, n, 3)


# ---------------------------------------------------------------------------------------------
# Singular pairs
# ---------------------------------------------------------------------------------------------


@dataclass
class SingularPairs:
    test: np.ndarray
    trial: np.ndarray
    rule: np.ndarray
    ptr: np.ndarray  # (F+1,) CSR by test element
    rule_off: np.ndarray
    rule_len: np.ndarray
    rule_pts: np.ndarray  # (n,5)
    keys: list


def _vertex_incidence(faces, n_vertices):
    F = len(faces)
    v = faces.reshape(-1)
    e = np.repeat(np.arange(F), 3)
    o = np.argsort(v, kind="stable")
    ptr = np.zeros(n_vertices + 1, dtype=np.int64)
    np.add.at(ptr, v + 1, 1)
    return np.cumsum(ptr), e[o], v[o]


def _topological_candidates(mesh: BoundaryMesh):
    """(t, s) pairs sharing at least one vertex id (incl. t == s) and their 3x3 vertex match matrices."""
    faces = mesh.faces
    F = len(faces)
    ptr, inc, _ = _vertex_incidence(faces, mesh.n_vertices)
    counts = ptr[faces + 1] - ptr[faces]  # (F,3)
    tot = counts.sum(1)
    t_idx = np.repeat(np.arange(F), tot)
    starts = ptr[faces].reshape(-1)
    lens = counts.reshape(-1)
    seg = np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens)
    s_idx = inc[np.arange(lens.sum()) + seg]
    key = np.unique(t_idx * F + s_idx)
    t, s = key // F, key % F
    match = faces[t][:, :, None] == faces[s][:, None, :]  # (P,3,3) [i test, j trial]
    return t, s, match


def _first_two_set_bits():
    """For every 9-bit mask: the positions of its first and second set bit in scan order (0 when absent)."""
    lut = np.zeros((512, 2), dtype=np.int64)
    for mask in range(512):
        bits = [b for b in range(9) if (mask >> b) & 1] + [0, 0]
        lut[mask] = bits[:2]
    return lut


_FIRST_TWO = _first_two_set_bits()


def _pair_codes(match):
    """Orientation codes of vertex-sharing pairs (BeatEngineCore.jl:453-507,820-844): kind 1 coincident,
    2 edge (test a, b; trial a, b after ordering the trial indices), 3 vertex (test a, trial a)."""
    nshared = match.sum(axis=(1, 2))
    kinds = np.where(nshared >= 3, 1, np.where(nshared == 2, 2, 3))
    # edge orientation: list matches in (i, j) scan order, then sort the pair by trial index j
    mask = match.reshape(len(match), 9).astype(np.int64) @ (1 << np.arange(9, dtype=np.int64))
    order_ij = _FIRST_TWO[mask]  # first two matches in scan order
    ia, ja = order_ij[:, 0] // 3, order_ij[:, 0] % 3
    ib, jb = order_ij[:, 1] // 3, order_ij[:, 1] % 3
    swap = jb < ja
    ia2, ib2 = np.where(swap, ib, ia), np.where(swap, ia, ib)
    ja2, jb2 = np.where(swap, jb, ja), np.where(swap, ja, jb)
    return np.where(
        kinds == 1,
        100000,
        np.where(kinds == 2, 200000 + ia2 * 1000 + ib2 * 100 + ja2 * 10 + jb2, 300000 + ia * 1000 + ja * 10),
    )


def _build_pairs(F, t, key, codes, order):
    """Pairs grouped by test element (CSR), sorted by trial key, with one remapped Duffy rule per code."""
    o = np.lexsort((key, t))
    t, key, codes = t[o], key[o], codes[o]
    uniq, rule_of_pair = np.unique(codes, return_inverse=True)
    return _pairs_with_rules(F, t, key, rule_of_pair, uniq, order)


def _pairs_with_rules(F, t, key, rule_of_pair, uniq, order):
    """SingularPairs from pairs already sorted by (test, trial key), rule_of_pair indexing the ascending distinct
    orientation codes `uniq`: one remapped Duffy rule per code."""
    bases = {k: duffy_base(order, k) for k in (1, 2, 3)}
    pts, offs, lens_r, keys = [], [], [], []
    off = 0
    for code in uniq:
        kind = int(code // 100000)
        test_p, trial_p, w = bases[kind]
        if kind == 2:
            a, b, c, d = (int(code // 1000) % 10, int(code // 100) % 10, int(code // 10) % 10, int(code) % 10)
            test_p, trial_p = remap_edge(test_p, a, b), remap_edge(trial_p, c, d)
            keys.append((2, a, b, c, d))
        elif kind == 3:
            a, c = int(code // 1000) % 10, int(code // 10) % 10
            test_p, trial_p = remap_vertex(test_p, a), remap_vertex(trial_p, c)
            keys.append((3, a, -1, c, -1))
        else:
            keys.append((1, -1, -1, -1, -1))
        pts.append(np.column_stack([test_p, trial_p, w]))
        offs.append(off)
        lens_r.append(len(w))
        off += len(w)
    ptr_t = np.zeros(F + 1, dtype=np.int64)
    np.add.at(ptr_t, t + 1, 1)
    return SingularPairs(
        test=t.astype(np.int32),
        trial=key.astype(np.int32),
        rule=rule_of_pair.astype(np.int32),
        ptr=np.cumsum(ptr_t).astype(np.int32),
        rule_off=np.asarray(offs, dtype=np.int32),
        rule_len=np.asarray(lens_r, dtype=np.int32),
        rule_pts=np.ascontiguousarray(np.concatenate(pts), dtype=np.float64) if pts else np.zeros((0, 5)),
        keys=keys,
    )


def _native_vertex_pairs(mesh: BoundaryMesh):
    """(test, trial, rule rank, distinct codes) of the vertex-sharing pairs from the library's host builder."""
    import ctypes as C

    from . import _lib

    faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
    deg = np.bincount(faces.reshape(-1), minlength=mesh.n_vertices)
    cap = int(deg[faces].sum())  # each face's candidates are at most the incidences of its three vertices
    t = np.empty(cap, np.int32)
    s = np.empty(cap, np.int32)
    rule = np.empty(cap, np.int32)
    codes = np.empty(64, np.int32)
    nc, n = C.c_int(), C.c_longlong()
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    _lib.check(_lib.load().bem_host_vertex_pairs(mesh.n_faces, mesh.n_vertices, ip(faces), cap, ip(t), ip(s), ip(rule),
                                                  ip(codes), C.byref(nc), C.byref(n)), "bem_host_vertex_pairs")
    n = int(n.value)
    return t[:n].astype(np.int64), s[:n].astype(np.int64), rule[:n].astype(np.int64), codes[: nc.value].astype(np.int64)


def singular_pairs(mesh: BoundaryMesh, order: int = 4, transforms=None, tol: float = GEOMETRY_TOLERANCE,
                   native: bool = True) -> SingularPairs:
    """All (test, trial) pairs sharing at least one vertex id (incl. t == s), with remapped Duffy rules.

    With plane symmetry, ``transforms`` lists the sign vectors of the identity and the image transforms; the
    image pairs (t, s) whose reflected trial element R_r s touches t geometrically (BeatEngineCore.jl:917-945,
    BeatEngineRocmRegular.jl:103-132) are added with trial key ``s + r * F``."""
    F = mesh.n_faces
    images = [r for r in range(1, len(transforms or ()))]
    if native:
        t, s, rule, uniq = _native_vertex_pairs(mesh)
        if not images:
            return _pairs_with_rules(F, t, s, rule, uniq, order)
        ts, keys, codes = [t], [s], [uniq[rule]]
    else:
        t, s, match = _topological_candidates(mesh)
        ts, keys, codes = [t], [s], [_pair_codes(match)]
    for r, signs in enumerate(transforms or ()):
        if r == 0:
            continue
        ti, si, mi = image_candidates(mesh, np.asarray(signs, float), tol)
        if len(ti):
            ts.append(ti)
            keys.append(si + r * F)
            codes.append(_pair_codes(mi))
    return _build_pairs(F, np.concatenate(ts), np.concatenate(keys), np.concatenate(codes), order)


def image_candidates(mesh: BoundaryMesh, signs, tol: float):
    """(t, s, match) with match[i, j] = |x_t,i - R x_s,j| <= tol, for the pairs with at least one such vertex
    (candidates by rounded positions as BeatEngineCore.jl:509-515,917-945)."""
    fv = mesh.face_vertices
    V = mesh.vertices
    keyv = np.rint(V / tol).astype(np.int64)
    keyr = np.rint(V * signs / tol).astype(np.int64)
    both = np.concatenate([keyv, keyr])
    _, inv = np.unique(both, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    gv, gr = inv[: len(V)], inv[len(V):]
    # vertices u whose position equals the reflection of vertex v
    order = np.argsort(gv, kind="stable")
    sorted_g = gv[order]
    lo = np.searchsorted(sorted_g, gr, side="left")
    hi = np.searchsorted(sorted_g, gr, side="right")
    vv = np.repeat(np.arange(len(V)), hi - lo)
    uu = order[np.concatenate([np.arange(a, b) for a, b in zip(lo, hi)]) if len(vv) else np.zeros(0, np.int64)]
    if len(vv) == 0:
        z = np.zeros(0, np.int64)
        return z, z, np.zeros((0, 3, 3), bool)
    ptr, inc, _ = _vertex_incidence(mesh.faces, mesh.n_vertices)
    tt, ss = [], []
    for u, v in zip(uu.tolist(), vv.tolist()):
        tu = inc[ptr[u]:ptr[u + 1]]
        sv = inc[ptr[v]:ptr[v + 1]]
        tt.append(np.repeat(tu, len(sv)))
        ss.append(np.tile(sv, len(tu)))
    key = np.unique(np.concatenate(tt) * mesh.n_faces + np.concatenate(ss))
    t, s = key // mesh.n_faces, key % mesh.n_faces
    d = fv[t][:, :, None, :] - (fv[s] * signs)[:, None, :, :]
    match = np.sqrt((d * d).sum(-1)) <= tol
    keep = match.any(axis=(1, 2))
    return t[keep], s[keep], match[keep]


# ---------------------------------------------------------------------------------------------
# Locality order and slot tables
# ---------------------------------------------------------------------------------------------


def _morton(points, bits=10):
    lo = points.min(0)
    span = np.maximum(points.max(0) - lo, 1e-30)
    q = np.minimum(((points - lo) / span * (2**bits - 1)).astype(np.int64), 2**bits - 1)
    code = np.zeros(len(points), dtype=np.int64)
    for b in range(bits):
        for a in range(3):
            code |= ((q[:, a] >> b) & 1) << (3 * b + a)
    return code


@dataclass
class SlotTables:
    order: np.ndarray  # (n_chunks*64,) int32
    slot_count: np.ndarray
    slot_start: np.ndarray  # (n_chunks, 196) uint16
    slot_inc: np.ndarray  # (n_chunks, 192) uint16
    slot_vertex: np.ndarray  # (n_chunks, 192) int32: 2 * vertex + (an earlier chunk stored the vertex column)
    slot_pend: np.ndarray  # (n_chunks, 192) int32: LDS partial-sum slot | PEND_FROM_LDS | PEND_TO_STRIP
    slot_prev: np.ndarray  # (n_chunks, 192) int32: previous chunk holding the slot's vertex, -1 if none
    elem_pos: np.ndarray  # (F,) int32: position of each element in the trial order
    chunk_start: np.ndarray  # (F,) int32: first chunk holding the element or an element sharing a vertex with it
    vertex_last: np.ndarray  # (V,) int32: last chunk holding each vertex (-1: unused vertex)
    vstar_ptr: np.ndarray
    vstar_row: np.ndarray


PEND_SLOTS = 96  # LDS partial-sum slots per wave in the regular kernels (BEM_PEND_SLOTS)
PEND_NONE = 0xFFFF
PEND_FROM_LDS = 1 << 16  # an earlier chunk left this vertex's partial sums in LDS
PEND_TO_STRIP = 1 << 17  # the vertex's last chunk: the finished sums go to the strip


def _native_slot_tables(mesh: BoundaryMesh, order: np.ndarray, pend_slots: int) -> SlotTables:
    import ctypes as C

    from . import _lib

    F, V = mesh.n_faces, mesh.n_vertices
    n_chunks = (F + 63) // 64
    faces = np.ascontiguousarray(mesh.faces, dtype=np.int32)
    order = np.ascontiguousarray(order, dtype=np.int32)
    out = dict(slot_count=np.empty(n_chunks, np.int32), slot_start=np.empty((n_chunks, 196), np.uint16),
               slot_inc=np.empty((n_chunks, 192), np.uint16), slot_vertex=np.empty((n_chunks, 192), np.int32),
               slot_pend=np.empty((n_chunks, 192), np.int32), slot_prev=np.empty((n_chunks, 192), np.int32),
               elem_pos=np.empty(F, np.int32), chunk_start=np.empty(F, np.int32), vertex_last=np.empty(V, np.int32),
               vstar_ptr=np.empty(V + 1, np.int32), vstar_row=np.empty(3 * F, np.int32))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    up = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint16))  # noqa: E731
    _lib.check(_lib.load().bem_host_slot_tables(
        F, V, ip(faces), ip(order), pend_slots, ip(out["slot_count"]), up(out["slot_start"]), up(out["slot_inc"]),
        ip(out["slot_vertex"]), ip(out["slot_pend"]), ip(out["slot_prev"]), ip(out["elem_pos"]),
        ip(out["chunk_start"]), ip(out["vertex_last"]), ip(out["vstar_ptr"]), ip(out["vstar_row"])),
        "bem_host_slot_tables")
    return SlotTables(order=order, **out)


def slot_tables(mesh: BoundaryMesh, pend_slots: int = PEND_SLOTS, native: bool = True) -> SlotTables:
    """Chunked trial order and per-chunk slot tables. A vertex that appears in several chunks keeps its partial
    strip sums in one of ``pend_slots`` LDS slots from its first to its last chunk (interval allocation in chunk
    order) and is stored once, by its last chunk; when no LDS slot is free it is stored by its first chunk and
    read-modify-written by the later ones (``slot_vertex`` odd)."""
    F, V = mesh.n_faces, mesh.n_vertices
    centroids = mesh.face_vertices.mean(axis=1)
    perm = np.argsort(_morton(centroids), kind="stable")
    n_chunks = (F + 63) // 64
    order = np.full(n_chunks * 64, -1, dtype=np.int32)
    order[:F] = perm
    if native:
        return _native_slot_tables(mesh, order, pend_slots)
    slot_count = np.zeros(n_chunks, dtype=np.int32)
    slot_start = np.zeros((n_chunks, 196), dtype=np.uint16)
    slot_inc = np.zeros((n_chunks, 192), dtype=np.uint16)
    slot_vertex = np.zeros((n_chunks, 192), dtype=np.int32)
    slot_pend = np.full((n_chunks, 192), PEND_NONE | PEND_TO_STRIP, dtype=np.int32)
    chunk_verts = []
    for c in range(n_chunks):
        elems = order[c * 64 : (c + 1) * 64]
        valid = elems >= 0
        lanes = np.nonzero(valid)[0]
        verts = mesh.faces[elems[valid]]  # (n,3)
        uniq, inv = np.unique(verts.reshape(-1), return_inverse=True)
        code = (np.repeat(lanes, 3) << 2) | np.tile(np.arange(3), len(lanes))
        o = np.lexsort((code, inv))
        slot_inc[c, : len(o)] = code[o]
        cnt = np.bincount(inv, minlength=len(uniq))
        slot_start[c, 1 : len(uniq) + 1] = np.cumsum(cnt)
        slot_count[c] = len(uniq)
        chunk_verts.append(uniq)
    first = np.full(V, -1, dtype=np.int64)
    last = np.full(V, -1, dtype=np.int64)
    for c, uniq in enumerate(chunk_verts):
        first[uniq] = np.where(first[uniq] < 0, c, first[uniq])
        last[uniq] = c
    # LDS slots for multi-chunk vertices. The live ranges [first, last] are intervals over chunks; keep the largest
    # set that needs at most pend_slots slots at every chunk (scan by start, and when over capacity drop the
    # interval that ends last), then colour it in chunk order. A slot is reused from the chunk after its
    # vertex's last one (within a chunk every live vertex has its own slot).
    import heapq

    multi = np.nonzero((first >= 0) & (last > first))[0]
    keep = np.zeros(V, dtype=bool)
    active = []  # max-heap by last chunk of the kept intervals that may still cover the scan point
    # every kept interval starts at or before the scan point (the scan runs by start), so the kept intervals
    # covering it are those that end at or after it: `live` counts them, `ends[c]` the kept ones ending at c
    ends = np.zeros(n_chunks + 1, dtype=np.int64)
    live, scan = 0, 0
    for v in multi[np.lexsort((last[multi], first[multi]))]:
        fv, lv = int(first[v]), int(last[v])
        while scan < fv:
            live -= int(ends[scan])
            scan += 1
        keep[v] = True
        heapq.heappush(active, (-lv, int(v)))
        live += 1
        ends[lv] += 1
        if live > pend_slots:  # over capacity: drop the covering interval that ends last
            while not keep[active[0][1]] or last[active[0][1]] < fv:
                heapq.heappop(active)
            _, u = heapq.heappop(active)
            keep[u] = False
            live -= 1
            ends[last[u]] -= 1
    pend = np.full(V, -1, dtype=np.int64)
    slot_prev = np.full((n_chunks, 192), -1, dtype=np.int32)
    last_seen = np.full(V, -1, dtype=np.int64)
    free = list(range(pend_slots - 1, -1, -1))
    release = [[] for _ in range(n_chunks)]
    for c, uniq in enumerate(chunk_verts):
        for v in uniq:
            if first[v] == c and keep[v]:
                pend[v] = free.pop()
                release[last[v]].append(pend[v])
        seen_before = first[uniq] < c
        in_lds = pend[uniq] >= 0
        code = np.where(in_lds, pend[uniq], PEND_NONE)
        code = code | np.where(in_lds & seen_before, PEND_FROM_LDS, 0)
        code = code | np.where(~in_lds | (last[uniq] == c), PEND_TO_STRIP, 0)
        slot_pend[c, : len(uniq)] = code
        slot_vertex[c, : len(uniq)] = 2 * uniq + (seen_before & ~in_lds)
        slot_prev[c, : len(uniq)] = last_seen[uniq]
        last_seen[uniq] = c
        free.extend(release[c])
    # paired assembly: element positions, the first chunk a test element's wave must visit (its own or that of an
    # element sharing a vertex with it), and each vertex's last chunk (the strip columns a wave stores)
    elem_pos = np.empty(F, dtype=np.int64)
    elem_pos[perm] = np.arange(F)
    echunk = elem_pos // 64
    vmin = np.full(V, n_chunks, dtype=np.int64)
    np.minimum.at(vmin, mesh.faces.reshape(-1), np.repeat(echunk, 3))
    chunk_start = np.minimum(echunk, vmin[mesh.faces].min(axis=1))
    rows = np.arange(3 * F)
    rv = mesh.faces.reshape(-1)
    o2 = np.lexsort((rows, rv))
    vs_ptr = np.zeros(V + 1, dtype=np.int64)
    np.add.at(vs_ptr, rv + 1, 1)
    return SlotTables(
        order=order,
        slot_count=slot_count,
        slot_start=slot_start,
        slot_inc=slot_inc,
        slot_vertex=slot_vertex,
        slot_pend=slot_pend,
        slot_prev=slot_prev,
        elem_pos=elem_pos.astype(np.int32),
        chunk_start=chunk_start.astype(np.int32),
        vertex_last=last.astype(np.int32),
        vstar_ptr=np.cumsum(vs_ptr).astype(np.int32),
        vstar_row=rows[o2].astype(np.int32),
    )


# ---------------------------------------------------------------------------------------------
# Everything the device needs for one mesh
# ---------------------------------------------------------------------------------------------


@dataclass
class PreparedMesh:
    """Device-ready mesh. ``mesh`` is in locality order: ``fperm[i]`` / ``vperm[i]`` are the original face /
    vertex ids of internal face / vertex ``i``; per-face inputs are permuted with ``fperm`` and per-vertex
    outputs restored with ``unpermute_vertices``."""

    mesh: BoundaryMesh
    curls: np.ndarray
    qpoints: np.ndarray
    pairs: SingularPairs
    slots: SlotTables
    m11: np.ndarray
    m10: np.ndarray
    fperm: np.ndarray
    vperm: np.ndarray
    original: BoundaryMesh
    symmetry: str = "off"
    transforms: np.ndarray = None  # (R,3) sign vectors, identity first
    row_weights: np.ndarray = None  # (V,) P1 orbit weights (internal order)
    plane_mask: np.ndarray = None  # (F,) bit r: face takes part in an image pair of transform r
    regular_order: int = 4  # order of the regular rule (qpoints holds its points)
    singular_order: int = 4

    @property
    def n_faces(self):
        return self.mesh.n_faces

    @property
    def n_vertices(self):
        return self.mesh.n_vertices

    def permute_faces(self, per_face):
        """Original face order -> internal order (first axis)."""
        return np.asarray(per_face)[self.fperm]

    def unpermute_vertices(self, per_vertex, axis=0):
        """Internal vertex order -> original order along ``axis``."""
        a = np.asarray(per_vertex)
        out = np.empty_like(a)
        idx = [slice(None)] * a.ndim
        idx[axis] = self.vperm
        out[tuple(idx)] = a
        return out


def locality_order(mesh: BoundaryMesh):
    """Faces in Morton order of their centroids; vertices in order of last use by those faces. With vertices
    numbered by last use, the vertices whose last chunk of 64 trial faces is chunk c are a contiguous range of
    ids, so the strip stores of each chunk (made by a vertex's last chunk) are contiguous."""
    fperm = np.argsort(_morton(mesh.face_vertices.mean(axis=1)), kind="stable")
    flat = mesh.faces[fperm].reshape(-1)
    last = np.full(mesh.n_vertices, -1, dtype=np.int64)
    np.maximum.at(last, flat, np.arange(len(flat)))
    ids = np.nonzero(last >= 0)[0]
    used = ids[np.argsort(last[ids], kind="stable")]
    unused = np.setdiff1d(np.arange(mesh.n_vertices), ids)
    vperm = np.concatenate([used, unused]).astype(np.int64)
    return fperm.astype(np.int64), vperm


# ---------------------------------------------------------------------------------------------
# Plane symmetry (BeatEngineCore.jl:317-439)
# ---------------------------------------------------------------------------------------------

SYMMETRY_SIGNS = {
    "off": ((1, 1, 1),),
    "x": ((1, 1, 1), (-1, 1, 1)),
    "xy": ((1, 1, 1), (-1, 1, 1), (1, -1, 1), (-1, -1, 1)),
}


def normalized_symmetry_mode(mode) -> str:
    m = str(mode).strip().lower()
    if m not in SYMMETRY_SIGNS:
        raise ValueError(f"Unsupported symmetry mode: {mode}. Expected off, x, or xy.")
    return m


def symmetry_reduction_factor(mode) -> int:
    return len(SYMMETRY_SIGNS[normalized_symmetry_mode(mode)])


def _active_axes(mode):
    return {"off": (), "x": (0,), "xy": (0, 1)}[normalized_symmetry_mode(mode)]


def symmetry_plane_tolerance(vertices) -> float:
    """max(1e-9, 1e-6 x bounding-box diagonal) (BeatEngineCore.jl:358-375)."""
    v = np.asarray(vertices, dtype=np.float64)
    scale = float(np.linalg.norm(v.max(0) - v.min(0))) if len(v) else 0.0
    return max(1e-9, scale * 1e-6)


def snap_symmetry_planes(mesh: BoundaryMesh, mode) -> BoundaryMesh:
    """Coordinates within the plane tolerance of an active symmetry plane become exactly zero
    (BeatEngineCore.jl:377-403)."""
    axes = _active_axes(mode)
    if not axes:
        return mesh
    tol = symmetry_plane_tolerance(mesh.vertices)
    v = np.array(mesh.vertices, dtype=np.float64)
    for a in axes:
        v[np.abs(v[:, a]) <= tol, a] = 0.0
    if np.array_equal(v, mesh.vertices):
        return mesh
    return BoundaryMesh(v, mesh.faces, mesh.tags, mesh.mesh_ids)


def validate_symmetry_domain(mesh: BoundaryMesh, mode) -> None:
    """The mesh must lie in the positive half-space of every active plane (BeatEngineCore.jl:405-423)."""
    tol = symmetry_plane_tolerance(mesh.vertices)
    for a in _active_axes(mode):
        bad = np.nonzero(mesh.vertices[:, a] < -tol)[0]
        if bad.size:
            name = "XYZ"[a]
            raise ValueError(
                f"Mesh is not in the positive {name} fundamental domain for {normalized_symmetry_mode(mode).upper()} "
                f"symmetry. Vertex {bad[0] + 1} has {name.lower()}={mesh.vertices[bad[0], a]} m.")


def p1_orbit_weights(mesh: BoundaryMesh, mode) -> np.ndarray:
    """Row weight 2 per active plane a vertex lies on (BeatEngineCore.jl:425-439)."""
    tol = symmetry_plane_tolerance(mesh.vertices)
    w = np.ones(mesh.n_vertices)
    for a in _active_axes(mode):
        w[np.abs(mesh.vertices[:, a]) <= tol] *= 2.0
    return w


def plane_masks(pairs: SingularPairs, n_faces: int, n_transforms: int) -> np.ndarray:
    """Bit r of face f is set when f takes part in an image pair of transform r (as test or trial)."""
    mask = np.zeros(n_faces, dtype=np.int32)
    r = pairs.trial // n_faces
    img = r > 0
    np.bitwise_or.at(mask, pairs.test[img], (1 << r[img]).astype(np.int32))
    np.bitwise_or.at(mask, pairs.trial[img] % n_faces, (1 << r[img]).astype(np.int32))
    return mask


_HELPER = None
_HELPER_LOCK = threading.Lock()


def _helper():
    """One process-wide helper thread for the independent parts of mesh preparation."""
    global _HELPER
    with _HELPER_LOCK:
        if _HELPER is None:
            _HELPER = concurrent.futures.ThreadPoolExecutor(max_workers=1, thread_name_prefix="bem-prepare")
        return _HELPER


def prepare(mesh: BoundaryMesh, singular_order: int = 4, reorder: bool = True, symmetry: str = "off",
            regular_order: int = 4) -> PreparedMesh:
    """Device-ready tables. With ``symmetry`` x / xy the mesh is the fundamental domain (already snapped and
    validated, as the request driver does); image pairs join the singular pairs with trial key s + r F.
    ``regular_order`` picks the regular rule (4: six points, 2: three, 1: the centroid) and ``singular_order``
    the Gauss order of the Duffy rules (1..4)."""
    if not 1 <= int(singular_order) <= 4:
        raise ValueError("Duffy 1D Gauss order must be between 1 and 4.")
    symmetry = normalized_symmetry_mode(symmetry)
    if reorder:
        fperm, vperm = locality_order(mesh)
    else:
        fperm, vperm = np.arange(mesh.n_faces), np.arange(mesh.n_vertices)
    vinv = np.empty_like(vperm)
    vinv[vperm] = np.arange(len(vperm))
    work = BoundaryMesh(mesh.vertices[vperm], vinv[mesh.faces[fperm]], mesh.tags[fperm], mesh.mesh_ids[fperm])
    signs = SYMMETRY_SIGNS[symmetry]
    # the singular pairs (mostly the library's native builder, which releases the GIL) on a helper thread while
    # this one forms the curls, rule points and slot tables
    pending = _helper().submit(singular_pairs, work, singular_order, transforms=signs if len(signs) > 1 else None)
    curls = surface_curls(work.face_vertices, work.normals)
    qp = quadrature_points(work.face_vertices, regular_order)
    m11, m10 = mass_blocks(regular_order)
    slots = slot_tables(work)
    pairs = pending.result()
    return PreparedMesh(work, curls, qp, pairs, slots, m11, m10, fperm, vperm, mesh, symmetry,
                        np.asarray(signs, dtype=np.int32), p1_orbit_weights(work, symmetry),
                        plane_masks(pairs, work.n_faces, len(signs)), int(regular_order), int(singular_order))
