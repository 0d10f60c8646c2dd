"""FP64 CPU oracle for the exterior Helmholtz BEM path (Burton-Miller, Galerkin P1/DP0).

TEST INFRASTRUCTURE ONLY. This module is the checker, never the product: only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import it. The shipped solver lives in ``boundary-lab_amd/`` and runs on HIP.

It restates, in float64 NumPy, the BEAT CPU algorithm of the reference:

* mesh / spaces          ``src/blab/solvers/julia_local/src/BeatEngineCore.jl:122-130,193-256,314-315``
* quadrature rules       ``BeatEngineCore.jl:517-631`` (triangle rule, 1-D Gauss, Duffy rules)
* Duffy remaps           ``BeatEngineCore.jl:633-655,751-771,820-873``
* adjacency              ``BeatEngineCore.jl:441-479``
* curls                  ``BeatEngineCore.jl:685-702``
* singular cache         ``BeatEngineCore.jl:875-915``
* identity matrices      ``BeatEngineCore.jl:731-749,952-979``
* regular pair assembly  ``src/blab/solvers/julia_local/src/BeatEngineCpuAssembly.jl:278-409``
* singular pair assembly ``BeatEngineCpuAssembly.jl:151-218,447-486``
* Burton-Miller system   ``src/blab/solvers/julia_local/src/BeatEngineCpuSolve.jl:45-71``
* field evaluation       ``BeatEngineCore.jl:1240-1280``, ``BeatEngineCpuField.jl:1-45``
* request driver         ``src/blab/solvers/julia_local/solver.jl:252-631,680-1020``
* plane symmetry x / xy  ``BeatEngineCore.jl:317-439,917-945,952-979,1245-1280``,
                         ``BeatEngineCpuAssembly.jl:61-100,411-445,488-527,571-633,713-886``

Pinning (see tests/test_oracle_*.py): rule sums and monomial exactness, Duffy
weight sums and tensor-rule agreement on smooth integrands, curl identities,
the closed-form P1 mass matrix, brute-force adaptive integration of singular
pairs, the analytic pulsating-sphere solution, LAPACK residuals, and a direct
double-loop field sum. Every function below is covered by one of those checks.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

FOUR_PI = 4.0 * math.pi

# --------------------------------------------------------------------------------------
# Mesh (BeatEngineCore.jl:122-130, 193-256)
# --------------------------------------------------------------------------------------


@dataclass
class Mesh:
    vertices: np.ndarray  # (V, 3) float64
    faces: np.ndarray  # (F, 3) int64, 0-based
    tags: np.ndarray  # (F,) int64 physical tags
    mesh_ids: np.ndarray  # (F,) int64, 1-based source-mesh id (solver.jl:252-269)

    def __post_init__(self):
        v = self.vertices
        f = self.faces
        self.face_vertices = v[f]  # (F, 3, 3)
        e1 = self.face_vertices[:, 1] - self.face_vertices[:, 0]
        e2 = self.face_vertices[:, 2] - self.face_vertices[:, 0]
        cross = np.cross(e1, e2)
        norm = np.linalg.norm(cross, axis=1)
        self.areas = norm / 2.0
        self.normals = cross / norm[:, None]
        self.centroids = self.face_vertices.mean(axis=1)

    @property
    def n_vertices(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.faces.shape[0])


def load_gmsh22(path: str, scale: float = 1.0, translation=(0.0, 0.0, 0.0)):
    """Gmsh 2.2 ASCII reader: keep type-2 triangles, tag = 4th token (BeatEngineCore.jl:193-234)."""
    with open(path, "r", encoding="utf-8") as handle:
        lines = [line.rstrip("\n") for line in handle]
    stripped = [line.strip() for line in lines]
    ns, ne = stripped.index("$Nodes"), stripped.index("$EndNodes")
    es, ee = stripped.index("$Elements"), stripped.index("$EndElements")
    index = {}
    verts = []
    for line in lines[ns + 2 : ne]:
        parts = line.split()
        if len(parts) < 4:
            continue
        verts.append([float(parts[1]) * scale, float(parts[2]) * scale, float(parts[3]) * scale])
        index[int(parts[0])] = len(verts) - 1
    faces, tags = [], []
    for line in lines[es + 2 : ee]:
        parts = line.split()
        if len(parts) < 8 or int(parts[1]) != 2:
            continue
        ids = [index.get(int(parts[-3]), -1), index.get(int(parts[-2]), -1), index.get(int(parts[-1]), -1)]
        if min(ids) < 0:
            continue
        faces.append(ids)
        tags.append(int(parts[3]))
    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3) + np.asarray(translation, dtype=np.float64)
    return v, np.asarray(faces, dtype=np.int64).reshape(-1, 3), np.asarray(tags, dtype=np.int64)


# --------------------------------------------------------------------------------------
# Quadrature (BeatEngineCore.jl:517-655)
# --------------------------------------------------------------------------------------


def triangle_rule(order: int = 4):
    """Reference-triangle rule; weights sum to 1/2 (BeatEngineCore.jl:517-549)."""
    if order <= 1:
        return np.array([[1.0 / 3.0, 1.0 / 3.0]]), np.array([0.5])
    if order == 4:
        pts = np.array(
            [
                [0.4459484909159651, 0.4459484909159651],
                [0.0915762135097710, 0.0915762135097700],
                [0.1081030181680700, 0.4459484909159651],
                [0.4459484909159651, 0.1081030181680700],
                [0.8168475729804590, 0.0915762135097700],
                [0.0915762135097710, 0.8168475729804580],
            ]
        )
        w = 0.5 * np.array(
            [
                0.2233815896780110,
                0.1099517436553220,
                0.2233815896780110,
                0.2233815896780110,
                0.1099517436553220,
                0.1099517436553220,
            ]
        )
        return pts, w
    return (
        np.array([[1.0 / 6.0, 1.0 / 6.0], [2.0 / 3.0, 1.0 / 6.0], [1.0 / 6.0, 2.0 / 3.0]]),
        np.array([1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0]),
    )


def gauss_rule_1d(order: int):
    """Gauss-Legendre on [0, 1] (BeatEngineCore.jl:551-569)."""
    if order == 1:
        return np.array([0.5]), np.array([1.0])
    if order == 2:
        a = 0.5 / math.sqrt(3.0)
        return np.array([0.5 - a, 0.5 + a]), np.array([0.5, 0.5])
    if order == 3:
        a = math.sqrt(3.0 / 5.0) / 2.0
        return np.array([0.5 - a, 0.5, 0.5 + a]), np.array([5.0 / 18.0, 4.0 / 9.0, 5.0 / 18.0])
    if order == 4:
        x1 = math.sqrt(3.0 / 7.0 - 2.0 / 7.0 * math.sqrt(6.0 / 5.0)) / 2.0
        x2 = math.sqrt(3.0 / 7.0 + 2.0 / 7.0 * math.sqrt(6.0 / 5.0)) / 2.0
        w1 = (18.0 + math.sqrt(30.0)) / 72.0
        w2 = (18.0 - math.sqrt(30.0)) / 72.0
        return np.array([0.5 - x2, 0.5 - x1, 0.5 + x1, 0.5 + x2]), np.array([w2, w1, w1, w2])
    raise ValueError("Duffy 1D Gauss order must be between 1 and 4.")


def duffy_rule(order: int, kind: str):
    """Duffy product rules over T x T for coincident / edge / vertex pairs (BeatEngineCore.jl:571-631).

    Returns (test_points (n,2), trial_points (n,2), weights (n,)). Weights sum to 1/4.
    """
    x, w = gauss_rule_1d(order)
    tp, tw = [], []
    for i in range(len(x)):
        for j in range(len(x)):
            tp.append((x[j], x[i]))
            tw.append(w[i] * w[j])
    test, trial, weights = [], [], []

    def add(tx, ty, sx, sy, wt):
        test.append((tx - ty, ty))
        trial.append((sx - sy, sy))
        weights.append(wt)

    for a in range(len(tp)):
        xsi, eta1 = tp[a]
        for b in range(len(tp)):
            eta2, eta3 = tp[b]
            eta12 = eta1 * eta2
            eta123 = eta12 * eta3
            base = tw[a] * tw[b]
            if kind == "coincident":
                wt = base * xsi**3 * eta1**2 * eta2
                add(xsi, xsi * (1 - eta1 + eta12), xsi * (1 - eta123), xsi * (1 - eta1), wt)
                add(xsi * (1 - eta123), xsi * (1 - eta1), xsi, xsi * (1 - eta1 + eta12), wt)
                add(xsi, xsi * (eta1 - eta12 + eta123), xsi * (1 - eta12), xsi * (eta1 - eta12), wt)
                add(xsi * (1 - eta12), xsi * (eta1 - eta12), xsi, xsi * (eta1 - eta12 + eta123), wt)
                add(xsi * (1 - eta123), xsi * (eta1 - eta123), xsi, xsi * (eta1 - eta12), wt)
                add(xsi, xsi * (eta1 - eta12), xsi * (1 - eta123), xsi * (eta1 - eta123), wt)
            elif kind == "edge":
                wt = base * xsi**3 * eta1**2
                add(xsi, xsi * eta1 * eta3, xsi * (1 - eta12), xsi * eta1 * (1 - eta2), wt)
                add(xsi, xsi * eta1, xsi * (1 - eta123), xsi * eta1 * eta2 * (1 - eta3), wt * eta2)
                add(xsi * (1 - eta12), xsi * eta1 * (1 - eta2), xsi, xsi * eta123, wt * eta2)
                add(xsi * (1 - eta123), xsi * eta12 * (1 - eta3), xsi, xsi * eta1, wt * eta2)
                add(xsi * (1 - eta123), xsi * eta1 * (1 - eta2 * eta3), xsi, xsi * eta12, wt * eta2)
            elif kind == "vertex":
                wt = base * xsi**3 * eta2
                add(xsi, xsi * eta1, xsi * eta2, xsi * eta2 * eta3, wt)
                add(xsi * eta2, xsi * eta2 * eta3, xsi, xsi * eta1, wt)
            else:
                raise ValueError(kind)
    return np.asarray(test), np.asarray(trial), np.asarray(weights)


_REF_VERTS = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])


def remap_shared_vertex(points, vid):
    """Map the Duffy singular corner to local vertex ``vid`` (0-based) (BeatEngineCore.jl:633-642)."""
    p = np.asarray(points, dtype=np.float64)
    if vid == 0:
        return p.copy()
    if vid == 1:
        return np.stack([1.0 - p[:, 0] - p[:, 1], p[:, 1]], axis=1)
    return np.stack([p[:, 0], 1.0 - p[:, 0] - p[:, 1]], axis=1)


def remap_shared_edge(points, va, vb):
    """Map the Duffy shared edge to local vertices (va, vb), 0-based (BeatEngineCore.jl:644-655)."""
    p = np.asarray(points, dtype=np.float64)
    rem = 3 - va - vb
    v0, v1, v2 = _REF_VERTS[va], _REF_VERTS[vb], _REF_VERTS[rem]
    return v0[None, :] + p[:, :1] * (v1 - v0)[None, :] + p[:, 1:2] * (v2 - v0)[None, :]


def adjacency_info(face_a, face_b):
    """Shared-vertex classification with orientation (BeatEngineCore.jl:453-479), 0-based locals."""
    sa, sb = [], []
    for i in range(3):
        for j in range(3):
            if face_a[i] == face_b[j]:
                sa.append(i)
                sb.append(j)
    if len(sa) == 3:
        return "coincident", (0, 1, 2), (0, 1, 2)
    if len(sa) == 2:
        if sb[1] < sb[0]:
            sa = [sa[1], sa[0]]
            sb = [sb[1], sb[0]]
        return "edge", (sa[0], sa[1]), (sb[0], sb[1])
    if len(sa) == 1:
        return "vertex", (sa[0],), (sb[0],)
    return "regular", (), ()


def orientation_key(kind, tv, sv):
    """BeatEngineCore.jl:820-844 (0-based vertex ids, -1 for unused)."""
    if kind == "coincident":
        return (1, -1, -1, -1, -1)
    if kind == "edge":
        return (2, tv[0], tv[1], sv[0], sv[1])
    return (3, tv[0], -1, sv[0], -1)


def remapped_rule(base, key):
    """BeatEngineCore.jl:800-818,846-873."""
    test, trial, w = base
    kind = key[0]
    if kind == 1:
        return test.copy(), trial.copy(), w.copy()
    if kind == 2:
        return remap_shared_edge(test, key[1], key[2]), remap_shared_edge(trial, key[3], key[4]), w.copy()
    return remap_shared_vertex(test, key[1]), remap_shared_vertex(trial, key[3]), w.copy()


def p1_values(points):
    p = np.asarray(points, dtype=np.float64)
    return np.stack([1.0 - p[:, 0] - p[:, 1], p[:, 0], p[:, 1]], axis=1)


def local_to_global(face_vertices, points):
    """(F,3,3) x (Q,2) -> (F,Q,3) (BeatEngineCore.jl:673-676)."""
    b = p1_values(points)  # (Q,3)
    return np.einsum("ql,flc->fqc", b, face_vertices)


def surface_curls(face_vertices, normals):
    """P1 surface curls n x grad(phi_i) via the Gram matrix (BeatEngineCore.jl:685-702). -> (F,3,3)."""
    e1 = face_vertices[:, 1] - face_vertices[:, 0]
    e2 = face_vertices[:, 2] - face_vertices[:, 0]
    g11 = np.einsum("fc,fc->f", e1, e1)
    g12 = np.einsum("fc,fc->f", e1, e2)
    g22 = np.einsum("fc,fc->f", e2, e2)
    det = g11 * g22 - g12 * g12
    inv11, inv12, inv22 = g22 / det, -g12 / det, g11 / det
    # lift = jac * inv(gram): columns l1, l2 (3-vectors)
    l1 = e1 * inv11[:, None] + e2 * inv12[:, None]
    l2 = e1 * inv12[:, None] + e2 * inv22[:, None]
    grads = np.stack([-l1 - l2, l1, l2], axis=1)  # (F,3,3)
    return np.cross(normals[:, None, :], grads)


# --------------------------------------------------------------------------------------
# Singular cache (BeatEngineCore.jl:784-915)
# --------------------------------------------------------------------------------------


@dataclass
class SingularCache:
    test: np.ndarray  # (P,) test element
    trial: np.ndarray  # (P,) trial element
    rule: np.ndarray  # (P,) rule index
    rules: list  # [(test_pts, trial_pts, weights)]
    keys: list


def adjacent_lists(faces, n_vertices):
    """Vertex-sharing trial candidates per test element, in reference order (BeatEngineCore.jl:784-811)."""
    incident = [[] for _ in range(n_vertices)]
    for e, f in enumerate(faces):
        for v in f:
            incident[int(v)].append(e)
    out = []
    for e, f in enumerate(faces):
        seen = []
        for v in f:
            for s in incident[int(v)]:
                if s not in seen:
                    seen.append(s)
        out.append(seen)
    return out


def build_singular_cache(mesh: Mesh, order: int = 4) -> SingularCache:
    base = {
        1: duffy_rule(order, "coincident"),
        2: duffy_rule(order, "edge"),
        3: duffy_rule(order, "vertex"),
    }
    rules, keys, index = [], [], {}
    test, trial, rule = [], [], []
    faces = mesh.faces
    for t, cands in enumerate(adjacent_lists(faces, mesh.n_vertices)):
        for s in cands:
            kind, tv, sv = adjacency_info(faces[t], faces[s])
            if kind == "regular":
                continue
            key = orientation_key(kind, tv, sv)
            if key not in index:
                index[key] = len(rules)
                rules.append(remapped_rule(base[key[0]], key))
                keys.append(key)
            test.append(t)
            trial.append(s)
            rule.append(index[key])
    return SingularCache(np.asarray(test), np.asarray(trial), np.asarray(rule), rules, keys)


def adjacency_matrix_rows(faces, rows):
    """Boolean (len(rows), F): trial shares any vertex with test (BeatEngineCore.jl:441-451)."""
    f = faces
    t = f[rows]
    return (t[:, :, None, None] == f[None, None, :, :]).any(axis=(1, 3))


# --------------------------------------------------------------------------------------
# Operators (BeatEngineCpuAssembly.jl:151-218, 278-409, 636-886)
# --------------------------------------------------------------------------------------


@dataclass
class Operators:
    single_layer: np.ndarray  # S  (V, F)
    double_layer: np.ndarray  # D  (V, V)
    adjoint_double_layer: np.ndarray  # D* (V, F)
    hypersingular: np.ndarray  # H  (V, V)


def _pair_blocks(x, y, nt, ns, w, bt, bs, k):
    """Integrand contraction for a batch of element pairs.

    x (P,Qa,3) test points, y (P,Qb,3) trial points, nt/ns (P,3) normals,
    w (P,Qa,Qb) full weights (rule weights x jacobians), bt (Qa,3) / bs (Qb,3) P1 values.
    Returns S (P,3), Ds (P,3), D (P,3,3), GM (P,3,3) = sum phi_i psi_j G w, Gs (P,) = sum G w.
    Arithmetic follows BeatEngineCpuAssembly.jl:308-328 (G, grad_scale, test/trial dots).
    """
    dx = y[:, None, :, 0] - x[:, :, None, 0]
    dy = y[:, None, :, 1] - x[:, :, None, 1]
    dz = y[:, None, :, 2] - x[:, :, None, 2]
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    ok = r > 0.0
    rs = np.where(ok, r, 1.0)
    inv = 1.0 / rs
    green = np.exp(1j * k * rs) * (inv / FOUR_PI)
    green = np.where(ok, green, 0.0)
    grad = green * (-inv + 1j * k)
    trial_dot = (dx * ns[:, 0, None, None] + dy * ns[:, 1, None, None] + dz * ns[:, 2, None, None]) * inv
    test_dot = -(dx * nt[:, 0, None, None] + dy * nt[:, 1, None, None] + dz * nt[:, 2, None, None]) * inv
    gw = green * w
    dw = grad * trial_dot * w
    aw = grad * test_dot * w
    S = gw.sum(axis=2) @ bt
    Ds = aw.sum(axis=2) @ bt
    D = np.matmul(np.matmul(bt.T, dw), bs)
    GM = np.matmul(np.matmul(bt.T, gw), bs)
    Gs = gw.sum(axis=(1, 2))
    return S, Ds, D, GM, Gs


def assemble_operators(mesh: Mesh, k: float, rule=None, singular_order: int = 4, chunk: int = 0,
                       singular_cache: SingularCache | None = None, with_singular: bool = True) -> Operators:
    """All four Galerkin operators S, D, D*, H in complex128 (BeatEngineCpuAssembly.jl:636-886)."""
    pts, wts = rule if rule is not None else triangle_rule(4)
    F, V = mesh.n_faces, mesh.n_vertices
    faces = mesh.faces
    fv = mesh.face_vertices
    X = local_to_global(fv, pts)  # (F,Q,3)
    B = p1_values(pts)  # (Q,3)
    Q = len(wts)
    curls = surface_curls(fv, mesh.normals)
    S = np.zeros((V, F), complex)
    Ds = np.zeros((V, F), complex)
    D = np.zeros((V, V), complex)
    H = np.zeros((V, V), complex)
    W2 = np.outer(wts, wts)  # (Q,Q)
    if chunk <= 0:
        chunk = max(1, int(2.0e6 // max(1, F * Q * Q)))
    cols_all = np.arange(F)
    for t0 in range(0, F, chunk):
        rows = np.arange(t0, min(F, t0 + chunk))
        adj = adjacency_matrix_rows(faces, rows)  # (T,F)
        ti, si = np.nonzero(~adj)
        if ti.size == 0:
            continue
        tt = rows[ti]
        jac = 4.0 * mesh.areas[tt] * mesh.areas[si]
        w = W2[None, :, :] * jac[:, None, None]
        Sb, Dsb, Db, GMb, Gsb = _pair_blocks(X[tt], X[si], mesh.normals[tt], mesh.normals[si], w, B, B, k)
        nn = np.einsum("pc,pc->p", mesh.normals[tt], mesh.normals[si])
        cp = np.einsum("pic,pjc->pij", curls[tt], curls[si])
        Hb = cp * Gsb[:, None, None] - (k * k) * nn[:, None, None] * GMb
        _scatter(S, Ds, D, H, faces, tt, si, Sb, Dsb, Db, Hb)
    del cols_all
    if with_singular:
        cache = singular_cache if singular_cache is not None else build_singular_cache(mesh, singular_order)
        add_singular(mesh, k, cache, curls, S, Ds, D, H)
    return Operators(S, D, Ds, H)


def _scatter(S, Ds, D, H, faces, tt, si, Sb, Dsb, Db, Hb):
    for i in range(3):
        r = faces[tt, i]
        np.add.at(S, (r, si), Sb[:, i])
        np.add.at(Ds, (r, si), Dsb[:, i])
        for j in range(3):
            c = faces[si, j]
            np.add.at(D, (r, c), Db[:, i, j])
            np.add.at(H, (r, c), Hb[:, i, j])


def singular_pair_blocks(mesh: Mesh, k: float, cache: SingularCache, curls=None, trial_geometry=None):
    """Per-pair Duffy blocks (S, Ds, D, H) for every vertex-sharing pair (BeatEngineCpuAssembly.jl:151-218).
    ``trial_geometry`` = (face_vertices, normals, curls) of reflected trial elements for image pairs."""
    if curls is None:
        curls = surface_curls(mesh.face_vertices, mesh.normals)
    fv_s, n_s, c_s = trial_geometry if trial_geometry is not None else (mesh.face_vertices, mesh.normals, curls)
    P = len(cache.test)
    Sb = np.zeros((P, 3), complex)
    Dsb = np.zeros((P, 3), complex)
    Db = np.zeros((P, 3, 3), complex)
    Hb = np.zeros((P, 3, 3), complex)
    fv = mesh.face_vertices
    for ri, (tp, sp, w) in enumerate(cache.rules):
        sel = np.nonzero(cache.rule == ri)[0]
        if sel.size == 0:
            continue
        bt, bs = p1_values(tp), p1_values(sp)
        for c0 in range(0, sel.size, 256):
            idx = sel[c0 : c0 + 256]
            t, s = cache.test[idx], cache.trial[idx]
            x = np.einsum("ql,plc->pqc", bt, fv[t])  # (P,n,3)
            y = np.einsum("ql,plc->pqc", bs, fv_s[s])
            jac = (2.0 * mesh.areas[t]) * (2.0 * mesh.areas[s])
            # Duffy points are paired one-to-one: evaluate on the diagonal only.
            rv = y - x
            r = np.sqrt((rv * rv).sum(axis=2))
            ok = r > 0.0
            rs = np.where(ok, r, 1.0)
            inv = 1.0 / rs
            green = np.where(ok, np.exp(1j * k * rs) * (inv / FOUR_PI), 0.0)
            grad = green * (-inv + 1j * k)
            nt, ns = mesh.normals[t], n_s[s]
            trial_dot = (rv * ns[:, None, :]).sum(axis=2) * inv
            test_dot = -(rv * nt[:, None, :]).sum(axis=2) * inv
            ww = w[None, :] * jac[:, None]
            gw = green * ww
            dw = grad * trial_dot * ww
            Sb[idx] = gw @ bt
            Dsb[idx] = (grad * test_dot * ww) @ bt
            Db[idx] = np.matmul((dw[:, :, None] * bt[None]).transpose(0, 2, 1), bs)
            GM = np.matmul((gw[:, :, None] * bt[None]).transpose(0, 2, 1), bs)
            Gs = gw.sum(axis=1)
            nn = (nt * ns).sum(axis=1)
            cp = np.einsum("pic,pjc->pij", curls[t], c_s[s])
            Hb[idx] = cp * Gs[:, None, None] - (k * k) * nn[:, None, None] * GM
    return Sb, Dsb, Db, Hb


def add_singular(mesh, k, cache, curls, S, Ds, D, H):
    Sb, Dsb, Db, Hb = singular_pair_blocks(mesh, k, cache, curls)
    _scatter(S, Ds, D, H, mesh.faces, cache.test, cache.trial, Sb, Dsb, Db, Hb)


def identity_matrices(mesh: Mesh, rule=None):
    """M11 (V,V) and M10 (V,F) with the regular rule (BeatEngineCore.jl:731-749,952-979)."""
    pts, wts = rule if rule is not None else triangle_rule(4)
    B = p1_values(pts)
    V, F = mesh.n_vertices, mesh.n_faces
    loc11 = np.einsum("qi,qj,q->ij", B, B, wts)  # reference element block / (2A)
    loc10 = np.einsum("qi,q->i", B, wts)
    M11 = np.zeros((V, V))
    M10 = np.zeros((V, F))
    jac = 2.0 * mesh.areas
    for i in range(3):
        np.add.at(M10, (mesh.faces[:, i], np.arange(F)), loc10[i] * jac)
        for j in range(3):
            np.add.at(M11, (mesh.faces[:, i], mesh.faces[:, j]), loc11[i, j] * jac)
    return M11, M10


def burton_miller(ops: Operators, M11, M10, k: float):
    """lhs = 1/2 M11 - D + (i/k) H ; rhs_op = -S - (i/k)(D* + 1/2 M10) (BeatEngineCpuSolve.jl:45-53)."""
    c = 1j / k
    lhs = 0.5 * M11 - ops.double_layer + c * ops.hypersingular
    rhs_op = -ops.single_layer - c * (ops.adjoint_double_layer + 0.5 * M10)
    return lhs, rhs_op


def solve_dense(lhs, rhs):
    """LAPACK getrf/getrs through scipy (the reference's cgetrf/cgetrs role, BeatEngineCpuSolve.jl:55-71)."""
    import scipy.linalg as sla

    lu, piv = sla.lu_factor(lhs)
    return sla.lu_solve((lu, piv), rhs)


# --------------------------------------------------------------------------------------
# Field (BeatEngineCore.jl:1240-1280, BeatEngineCpuField.jl:1-45)
# --------------------------------------------------------------------------------------


def field_sources(mesh: Mesh, rule=None):
    pts, wts = rule if rule is not None else triangle_rule(4)
    X = local_to_global(mesh.face_vertices, pts)  # (F,Q,3)
    F, Q = X.shape[0], X.shape[1]
    return {
        "points": X.reshape(F * Q, 3),
        "normals": np.repeat(mesh.normals, Q, axis=0),
        "weights": (wts[None, :] * (2.0 * mesh.areas)[:, None]).reshape(F * Q),
        "faces": np.repeat(mesh.faces, Q, axis=0),
        "elements": np.repeat(np.arange(F), Q),
        "basis": np.tile(p1_values(pts), (F, 1)),
    }


def evaluate_field(points, mesh: Mesh, pressure, q_neumann, k: float, sources=None, chunk: int = 64):
    """u(x) = sum_m [dG/dn_y p_w - G q_w] (BeatEngineCpuField.jl:1-45). pressure (V,) or (V,C)."""
    src = sources if sources is not None else field_sources(mesh)
    p = np.asarray(pressure)
    q = np.asarray(q_neumann)
    single = p.ndim == 1
    if single:
        p = p[:, None]
        q = q[:, None]
    pw = np.einsum("ml,mlc->mc", src["basis"], p[src["faces"]]) * src["weights"][:, None]
    qw = q[src["elements"]] * src["weights"][:, None]
    pts = np.asarray(points, dtype=np.float64)
    out = np.zeros((len(pts), p.shape[1]), complex)
    for c0 in range(0, len(pts), chunk):
        x = pts[c0 : c0 + chunk]
        rv = src["points"][None, :, :] - x[:, None, :]
        r = np.sqrt(np.einsum("omc,omc->om", rv, rv))
        ok = r > 0
        rs = np.where(ok, r, 1.0)
        green = np.where(ok, np.exp(1j * k * rs) / (FOUR_PI * rs), 0.0)
        proj = np.einsum("omc,mc->om", rv, src["normals"]) / rs
        dl = green * (1j * k - 1.0 / rs) * proj
        out[c0 : c0 + chunk] = dl @ pw - green @ qw
    return out[:, 0] if single else out


# --------------------------------------------------------------------------------------
# Plane symmetry x / xy (BeatEngineCore.jl:317-439,917-945,952-979,1245-1280;
# BeatEngineCpuAssembly.jl:61-100,411-445,488-527,571-633,713-886; solver.jl:607-608,718-719)
# --------------------------------------------------------------------------------------


def symmetry_transforms(mode: str, include_identity: bool = True):
    """[(name, signs (3,), determinant)] in the reference order (BeatEngineCore.jl:323-339)."""
    mode = str(mode).strip().lower()
    if mode not in ("off", "x", "xy"):
        raise ValueError(f"Unsupported symmetry mode: {mode}. Expected off, x, or xy.")
    out = [("identity", np.array([1.0, 1.0, 1.0]), 1.0)] if include_identity else []
    if mode == "x":
        out.append(("x", np.array([-1.0, 1.0, 1.0]), -1.0))
    elif mode == "xy":
        out += [("x", np.array([-1.0, 1.0, 1.0]), -1.0), ("y", np.array([1.0, -1.0, 1.0]), -1.0),
                ("xy", np.array([-1.0, -1.0, 1.0]), 1.0)]
    return out


def _active_axes(mode):
    return {"off": (), "x": (0,), "xy": (0, 1)}[str(mode).strip().lower()]


def symmetry_plane_tolerance(vertices, absolute_floor=1e-9, relative_tolerance=1e-6):
    """BeatEngineCore.jl:358-375: max(floor, |bbox diagonal| * rel)."""
    v = np.asarray(vertices, float)
    scale = float(np.linalg.norm(v.max(0) - v.min(0))) if len(v) else 0.0
    return max(absolute_floor, scale * relative_tolerance)


def snap_symmetry_planes(vertices, mode, tol=None):
    """Coordinates within tol of an active plane become exactly 0 (BeatEngineCore.jl:377-403)."""
    v = np.array(vertices, float)
    tol = symmetry_plane_tolerance(v) if tol is None else tol
    for a in _active_axes(mode):
        v[np.abs(v[:, a]) <= tol, a] = 0.0
    return v


def validate_fundamental_domain(vertices, mode, tol=None):
    """Error if a vertex lies on the negative side of an active plane (BeatEngineCore.jl:405-423)."""
    v = np.asarray(vertices, float)
    tol = symmetry_plane_tolerance(v) if tol is None else tol
    for a in _active_axes(mode):
        bad = np.nonzero(v[:, a] < -tol)[0]
        if bad.size:
            name = "XYZ"[a]
            raise ValueError(f"Mesh is not in the positive {name} fundamental domain for {str(mode).upper()} "
                             f"symmetry. Vertex {bad[0] + 1} has {name.lower()}={v[bad[0], a]} m.")


def p1_orbit_weights(vertices, mode, tol=None):
    """2 per active plane the vertex lies on (BeatEngineCore.jl:425-439)."""
    v = np.asarray(vertices, float)
    tol = symmetry_plane_tolerance(v) if tol is None else tol
    w = np.ones(len(v))
    for a in _active_axes(mode):
        w[np.abs(v[:, a]) <= tol] *= 2.0
    return w


def geometric_adjacency_info(test_vertices, trial_vertices, tol):
    """Shared vertices by position (BeatEngineCore.jl:481-507), 0-based locals, same ordering rules as
    ``adjacency_info``."""
    sa, sb = [], []
    for i in range(3):
        for j in range(3):
            if np.linalg.norm(test_vertices[i] - trial_vertices[j]) <= tol:
                sa.append(i)
                sb.append(j)
    if len(sa) == 3:
        return "coincident", (sa[0], sa[1], sa[2]), (sb[0], sb[1], sb[2])
    if len(sa) == 2:
        if sb[1] < sb[0]:
            sa = [sa[1], sa[0]]
            sb = [sb[1], sb[0]]
        return "edge", (sa[0], sa[1]), (sb[0], sb[1])
    if len(sa) == 1:
        return "vertex", (sa[0],), (sb[0],)
    return "regular", (), ()


def reflected_geometry(mesh: Mesh, signs, det, curls=None):
    """Trial data of the image elements: vertices, normals and curls (BeatEngineCore.jl:341-356)."""
    if curls is None:
        curls = surface_curls(mesh.face_vertices, mesh.normals)
    return mesh.face_vertices * signs, mesh.normals * signs, curls * (det * signs)


def image_singular_cache(mesh: Mesh, order: int, signs, tol: float = 1e-8) -> SingularCache:
    """Pairs (t, s) whose image R s touches t geometrically (BeatEngineCore.jl:917-945,
    BeatEngineCpuAssembly.jl:488-527): candidates through rounded vertex keys, then the geometric
    classification and remapped Duffy rule."""
    base = {1: duffy_rule(order, "coincident"), 2: duffy_rule(order, "edge"), 3: duffy_rule(order, "vertex")}
    fv = mesh.face_vertices
    by_key = {}
    for t in range(mesh.n_faces):
        for p in fv[t]:
            by_key.setdefault(tuple(np.rint(p / tol).astype(np.int64)), []).append(t)
    cands, seen = [], set()
    for s in range(mesh.n_faces):
        for p in fv[s] * signs:
            for t in by_key.get(tuple(np.rint(p / tol).astype(np.int64)), []):
                if (t, s) not in seen:
                    seen.add((t, s))
                    cands.append((t, s))
    rules, keys, tt, ss, rr = [], [], [], [], []
    for t, s in cands:
        kind, tv, sv = geometric_adjacency_info(fv[t], fv[s] * signs, tol)
        if kind == "regular":
            continue
        key = orientation_key(kind, tv, sv)
        if key not in keys:
            keys.append(key)
            rules.append(remapped_rule(base[key[0]], key))
        tt.append(t)
        ss.append(s)
        rr.append(keys.index(key))
    return SingularCache(np.asarray(tt, int), np.asarray(ss, int), np.asarray(rr, int), rules, keys)


def assemble_operators_symmetric(mesh: Mesh, k: float, mode: str, singular_order: int = 4, rule=None,
                                 tol: float | None = None) -> Operators:
    """S, D, D*, H of the symmetric extension: the plain operators, plus for every image transform R all pairs
    (t, R s) with the regular rule except the geometrically touching ones, which take their Duffy rule; then
    the P1 orbit row weights (BeatEngineCpuAssembly.jl:636-886)."""
    ops = assemble_operators(mesh, k, rule=rule, singular_order=singular_order)
    S, D, Ds, H = ops.single_layer, ops.double_layer, ops.adjoint_double_layer, ops.hypersingular
    pts, wts = rule if rule is not None else triangle_rule(4)
    F = mesh.n_faces
    faces = mesh.faces
    curls = surface_curls(mesh.face_vertices, mesh.normals)
    X = local_to_global(mesh.face_vertices, pts)
    B = p1_values(pts)
    W2 = np.outer(wts, wts)
    for _, signs, det in symmetry_transforms(mode, include_identity=False):
        fv_r, n_r, c_r = reflected_geometry(mesh, signs, det, curls)
        X_r = X * signs
        img = image_singular_cache(mesh, singular_order, signs)
        touching = np.zeros((F, F), bool)
        touching[img.test, img.trial] = True
        chunk = max(1, int(2.0e6 // max(1, F * len(wts) ** 2)))
        for t0 in range(0, F, chunk):
            rows = np.arange(t0, min(F, t0 + chunk))
            ti, si = np.nonzero(~touching[rows])
            tt = rows[ti]
            jac = 4.0 * mesh.areas[tt] * mesh.areas[si]
            w = W2[None, :, :] * jac[:, None, None]
            Sb, Dsb, Db, GMb, Gsb = _pair_blocks(X[tt], X_r[si], mesh.normals[tt], n_r[si], w, B, B, k)
            nn = np.einsum("pc,pc->p", mesh.normals[tt], n_r[si])
            cp = np.einsum("pic,pjc->pij", curls[tt], c_r[si])
            Hb = cp * Gsb[:, None, None] - (k * k) * nn[:, None, None] * GMb
            _scatter(S, Ds, D, H, faces, tt, si, Sb, Dsb, Db, Hb)
        if len(img.test):
            Sb, Dsb, Db, Hb = singular_pair_blocks(mesh, k, img, curls, trial_geometry=(fv_r, n_r, c_r))
            _scatter(S, Ds, D, H, faces, img.test, img.trial, Sb, Dsb, Db, Hb)
    wgt = p1_orbit_weights(mesh.vertices, mode, tol)[:, None]
    return Operators(S * wgt, D * wgt, Ds * wgt, H * wgt)


def identity_matrices_symmetric(mesh: Mesh, mode: str, rule=None, tol=None):
    """Identity blocks with the orbit row weights (BeatEngineCore.jl:952-979)."""
    M11, M10 = identity_matrices(mesh, rule)
    wgt = p1_orbit_weights(mesh.vertices, mode, tol)[:, None]
    return M11 * wgt, M10 * wgt


def field_sources_symmetric(mesh: Mesh, mode: str, rule=None):
    """Sources of every transform, element-major then transform (BeatEngineCore.jl:1245-1280)."""
    pts, wts = rule if rule is not None else triangle_rule(4)
    tfs = symmetry_transforms(mode)
    X = local_to_global(mesh.face_vertices, pts)  # (F,Q,3)
    F, Q, R = X.shape[0], X.shape[1], len(tfs)
    signs = np.stack([t[1] for t in tfs])  # (R,3)
    pts_all = X[:, None, :, :] * signs[None, :, None, :]  # (F,R,Q,3)
    nrm_all = np.repeat((mesh.normals[:, None, :] * signs[None, :, :])[:, :, None, :], Q, axis=2)
    return {
        "points": pts_all.reshape(F * R * Q, 3),
        "normals": nrm_all.reshape(F * R * Q, 3),
        "weights": np.broadcast_to((wts[None, :] * (2.0 * mesh.areas)[:, None])[:, None, :], (F, R, Q)).reshape(-1),
        "faces": np.repeat(mesh.faces, R * Q, axis=0),
        "elements": np.repeat(np.arange(F), R * Q),
        "basis": np.tile(p1_values(pts), (F * R, 1)),
    }


def solve_frequency_symmetric(mesh: Mesh, freq: float, q_unit, mode: str, sound_speed=343.0, rho=1.21,
                              points=None):
    """One frequency of the plain method on the fundamental domain of an x / xy symmetric model."""
    omega = 2.0 * math.pi * freq
    k = omega / sound_speed
    ops = assemble_operators_symmetric(mesh, k, mode)
    M11, M10 = identity_matrices_symmetric(mesh, mode)
    lhs, rhs_op = burton_miller(ops, M11, M10, k)
    q = 1j * rho * omega * np.asarray(q_unit, complex).reshape(mesh.n_faces, -1)
    rhs = rhs_op @ q
    p = solve_dense(lhs, rhs)
    out = {"pressure": p, "q": q, "lhs": lhs, "rhs": rhs, "k": k}
    if points is not None:
        out["field"] = evaluate_field(points, mesh, p, q, k, sources=field_sources_symmetric(mesh, mode))
    return out


# --------------------------------------------------------------------------------------
# Request driver pieces (solver.jl)
# --------------------------------------------------------------------------------------


def polar_observation_points(config):
    """Polar arcs, angles built in Float32 as the reference does (solver.jl:320-348)."""
    step = float(config.get("step_size", 5.0))
    amin = float(config.get("min_angle", -180.0))
    amax = float(config.get("max_angle", 180.0))
    n = int(math.floor((amax - amin) / step + 1e-9)) + 1
    angles = np.array([amin + i * step for i in range(n)], dtype=np.float64).astype(np.float32)
    if angles.size == 0 or angles[-1] < np.float32(amax):
        angles = np.append(angles, np.float32(amax))
    angles = np.clip(angles, np.float32(amin), np.float32(amax)).astype(np.float32)
    d = np.float32(config.get("distance", 2.0))
    off = np.float32(config.get("axial_offset", 0.0))
    ang = np.float32(np.pi) * angles / np.float32(180.0)
    s, c = np.sin(ang).astype(np.float32), np.cos(ang).astype(np.float32)
    zero = np.zeros_like(s)
    horizontal = np.stack([d * s, zero, d * c + off], axis=1).astype(np.float32)
    vertical = np.stack([zero, d * s, d * c + off], axis=1).astype(np.float32)
    on_axis = int(np.argmin(np.abs(angles.astype(np.float64))))
    return angles, horizontal, vertical, on_axis


def fibonacci_observation(config):
    """Fibonacci sphere in the reference ordering (solver.jl:350-388)."""
    if not bool(config.get("spherical_sampling_enabled", False)):
        return None
    n = int(config.get("spherical_sampling_points", 6000))
    d = np.float32(config.get("distance", 2.0))
    off = np.float32(config.get("axial_offset", 0.0))
    golden = np.float32(math.pi * (3.0 - math.sqrt(5.0)))
    i = np.arange(n)
    z = (1.0 - (2.0 * i + 1.0) / n).astype(np.float32)
    xy = np.sqrt(np.maximum(np.float32(1.0) - z * z, np.float32(0.0))).astype(np.float32)
    az = (i.astype(np.float32) * golden).astype(np.float32)
    x = (d * xy * np.cos(az)).astype(np.float32)
    y = (d * xy * np.sin(az)).astype(np.float32)
    zz = (d * z + off).astype(np.float32)
    theta = np.arccos(np.clip(z, -1, 1)).astype(np.float32)
    phi = np.mod(np.arctan2(y, x), np.float32(2 * np.pi)).astype(np.float32)
    return np.stack([x, y, zz], axis=1), {"r_distance_m": np.full(n, d, np.float32), "theta_polar_rad": theta,
                                           "phi_azimuth_rad": phi}


def solve_frequency(mesh: Mesh, freq: float, q_unit, sound_speed=343.0, rho=1.21, singular_cache=None,
                    identity=None, points=None, regular_order: int = 4, singular_order: int = 4):
    """One frequency of the plain method in float64: operators, BM system, dense LU, field. The regular rule
    (triangle_rule(regular_order)) serves the regular pairs, the identity blocks and the field sources, as in
    the reference (solver.jl:729-735); singular pairs take Duffy rules of singular_order.

    q_unit: (F, C) Neumann drives without the i rho omega factor (columns = channels).
    Returns dict(pressure (V,C), q (F,C), field (N,C) if points given, lhs, rhs, rhs_op).
    """
    omega = 2.0 * math.pi * freq
    k = omega / sound_speed
    rule = triangle_rule(regular_order)
    ops = assemble_operators(mesh, k, rule=rule, singular_order=singular_order, singular_cache=singular_cache)
    M11, M10 = identity if identity is not None else identity_matrices(mesh, rule)
    lhs, rhs_op = burton_miller(ops, M11, M10, k)
    q = 1j * rho * omega * np.asarray(q_unit, complex).reshape(mesh.n_faces, -1)
    rhs = rhs_op @ q
    p = solve_dense(lhs, rhs)
    out = {"pressure": p, "q": q, "lhs": lhs, "rhs": rhs, "rhs_op": rhs_op, "k": k}
    if points is not None:
        out["field"] = evaluate_field(points, mesh, p, q, k, sources=field_sources(mesh, rule))
    return out


def pulsating_sphere_pressure(r, a, k, rho, omega, v=1.0):
    """Exact radiated pressure of a pulsating sphere, exp(-i w t), outgoing G (closed form)."""
    return 1j * rho * omega * v * a * a * np.exp(1j * k * (r - a)) / ((1j * k * a - 1.0) * r)
