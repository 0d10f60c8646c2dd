"""Seeded synthetic boundary meshes for the solver's workloads.

* ``icosphere``           pulsating-sphere case: icosahedron subdivided n times, projected to radius a.
* ``horn_in_box``         axisymmetric exponential horn flush-mounted in a closed rectangular box, throat
                          disk as the radiator.
* ``loudspeaker_box``     box with a horn and circular piston cones as separate radiator tags.

All generators return ``(vertices (V,3) float64 [m], faces (F,3) int64 0-based, tags (F,) int64)``.
Surfaces are closed, consistently oriented and genus 0; normals point into the fluid, i.e. out of the
solid body, which is the orientation the solver expects (normals from ``cross(v2-v1, v3-v1)``).
"""

from __future__ import annotations

import math

import numpy as np

TAG_WALL = 1
TAG_THROAT = 2


# ----------------------------------------------------------------------------------------------
# helpers
# ----------------------------------------------------------------------------------------------


class _Builder:
    def __init__(self):
        self.v: list = []
        self.f: list = []
        self.t: list = []
        self.key = {}

    def vertex(self, p, key=None):
        if key is not None and key in self.key:
            return self.key[key]
        self.v.append(tuple(float(c) for c in p))
        idx = len(self.v) - 1
        if key is not None:
            self.key[key] = idx
        return idx

    def tri(self, a, b, c, tag, outward):
        """Add triangle, flipped if needed so that its normal has positive dot with ``outward``."""
        pa, pb, pc = (np.asarray(self.v[i]) for i in (a, b, c))
        n = np.cross(pb - pa, pc - pa)
        if float(np.dot(n, outward)) < 0.0:
            b, c = c, b
        self.f.append((a, b, c))
        self.t.append(tag)

    def arrays(self):
        return (
            np.asarray(self.v, dtype=np.float64),
            np.asarray(self.f, dtype=np.int64),
            np.asarray(self.t, dtype=np.int64),
        )


def _edge_points(p0, p1, n):
    """n segments -> n+1 points from p0 to p1 inclusive."""
    p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
    return [p0 + (p1 - p0) * (i / n) for i in range(n + 1)]


def _ring_strip(b, inner, outer, tag, outward_fn):
    """Triangulate between two closed rings (lists of vertex ids); sizes equal or inner = outer/2."""
    m_in, m_out = len(inner), len(outer)
    if m_in == m_out:
        for k in range(m_out):
            k1 = (k + 1) % m_out
            a, bb, c, d = outer[k], outer[k1], inner[k1], inner[k]
            b.tri(a, bb, c, tag, outward_fn(a, bb, c))
            b.tri(a, c, d, tag, outward_fn(a, c, d))
    elif 2 * m_in == m_out:
        for k in range(m_in):
            f0, f1, f2 = outer[2 * k], outer[2 * k + 1], outer[(2 * k + 2) % m_out]
            c0, c1 = inner[k], inner[(k + 1) % m_in]
            b.tri(f0, f1, c0, tag, outward_fn(f0, f1, c0))
            b.tri(f1, c1, c0, tag, outward_fn(f1, c1, c0))
            b.tri(f1, f2, c1, tag, outward_fn(f1, f2, c1))
    else:
        raise ValueError("ring sizes must match or halve")


def _disk_rings(b, rim_ids, center, radius_fn_pts, h, tag, normal, min_ring=6):
    """Fill a planar region bounded by ring ``rim_ids`` toward ``center`` with halving rings."""
    center = np.asarray(center, float)
    rim = [np.asarray(b.v[i]) for i in rim_ids]
    ring_ids = list(rim_ids)
    ring_pts = rim
    dist = max(float(np.linalg.norm(p - center)) for p in rim)
    n_rings = max(1, int(round(dist / h)))
    outward = lambda *ids: normal  # noqa: E731
    for r in range(1, n_rings):
        s = 1.0 - r / n_rings  # fraction of the way from center to rim
        m = len(ring_ids)
        spacing = 2.0 * math.pi * dist * s / m
        take_half = m % 2 == 0 and m // 2 >= min_ring and spacing < 0.55 * h
        src = ring_pts[::2] if take_half else ring_pts
        new_pts = [center + (p - center) * (s / (1.0 - (r - 1) / n_rings)) for p in src]
        new_ids = [b.vertex(p) for p in new_pts]
        _ring_strip(b, new_ids, ring_ids, tag, outward)
        ring_ids, ring_pts = new_ids, new_pts
    c = b.vertex(center)
    m = len(ring_ids)
    for k in range(m):
        a, bb = ring_ids[k], ring_ids[(k + 1) % m]
        b.tri(a, bb, c, tag, normal)


# ----------------------------------------------------------------------------------------------
# icosphere
# ----------------------------------------------------------------------------------------------


def icosphere(subdivisions: int = 3, radius: float = 0.1, tag: int = TAG_THROAT):
    """Icosahedron subdivided ``subdivisions`` times: 20*4^n faces, outward normals."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [
        (-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
        (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
        (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1),
    ]
    faces = [
        (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
        (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
        (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
        (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
    ]
    v = [np.asarray(p, float) / np.linalg.norm(p) for p in verts]
    for _ in range(subdivisions):
        cache = {}
        new_faces = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                cache[key] = len(v) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    V = np.asarray(v) * radius
    F = np.asarray(faces, dtype=np.int64)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = np.einsum("ij,ij->i", n, V[F].mean(axis=1)) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F, np.full(len(F), tag, dtype=np.int64)


# ----------------------------------------------------------------------------------------------
# box with holes (horn mouth, pistons)
# ----------------------------------------------------------------------------------------------


def uv_sphere(n_lat: int = 12, n_lon: int = 24, radius: float = 0.1, tag: int = TAG_THROAT):
    """Latitude/longitude sphere with outward normals, mirror-symmetric in x and y. ``n_lon`` is a multiple of 4,
    so the planes x = 0 and y = 0 are unions of mesh edges (meridians); coordinates on them are exactly zero."""
    if n_lon % 4 or n_lat < 2:
        raise ValueError("n_lon must be a multiple of 4 and n_lat >= 2")
    V = [(0.0, 0.0, radius)]
    for i in range(1, n_lat):
        th = math.pi * i / n_lat
        for j in range(n_lon):
            q = j % (n_lon // 4)  # exact zeros on the x = 0 / y = 0 meridians
            ph = 2.0 * math.pi * j / n_lon
            c = 0.0 if (q == 0 and (j // (n_lon // 4)) % 2 == 1) else math.cos(ph)
            sn = 0.0 if (q == 0 and (j // (n_lon // 4)) % 2 == 0) else math.sin(ph)
            V.append((radius * math.sin(th) * c, radius * math.sin(th) * sn, radius * math.cos(th)))
    V.append((0.0, 0.0, -radius))
    south = len(V) - 1

    def ring(i, j):
        return 1 + (i - 1) * n_lon + (j % n_lon)

    Fl = []
    for j in range(n_lon):
        Fl.append((0, ring(1, j), ring(1, j + 1)))
    for i in range(1, n_lat - 1):
        for j in range(n_lon):
            a, b, c, d = ring(i, j), ring(i, j + 1), ring(i + 1, j), ring(i + 1, j + 1)
            # quad diagonals alternate by quadrant so that x -> -x and y -> -y map the mesh onto itself
            if (j // (n_lon // 4)) % 2 == 0:
                Fl += [(a, c, d), (a, d, b)]
            else:
                Fl += [(a, c, b), (b, c, d)]
    for j in range(n_lon):
        Fl.append((south, ring(n_lat - 1, j + 1), ring(n_lat - 1, j)))
    Vn, Fn = np.asarray(V, float), np.asarray(Fl, np.int64)
    return Vn, Fn, np.full(len(Fn), tag, dtype=np.int64)


def symmetric_part(V, F, T, mode: str):
    """Faces of a plane-symmetric mesh inside the fundamental domain (x >= 0 for "x"; x, y >= 0 for "xy"),
    with unused vertices dropped (vertex order kept)."""
    axes = {"x": (0,), "xy": (0, 1)}[mode]
    c = V[F].mean(axis=1)
    keep = np.all(c[:, axes] > 0.0, axis=1)
    Fk = F[keep]
    used = np.unique(Fk)
    remap = np.full(len(V), -1, dtype=np.int64)
    remap[used] = np.arange(len(used))
    return V[used], remap[Fk], T[keep]


def _box_with_holes(width, height, depth, bands, h, horn=None, seed=None, jitter=0.0):
    """Closed box surface; the front face (z=0, normal +z) is split into horizontal bands, each with one
    circular hole filled by either a horn (dict) or a flat piston disk.

    bands: list of dicts {y0, y1, cx, cy, radius, kind: 'horn'|'piston', tag}
    """
    b = _Builder()
    x0, x1 = -width / 2.0, width / 2.0
    n_w = max(4, int(round(width / h)))
    depth_n = max(2, int(round(depth / h)))
    perim_right = []  # per band: ids bottom->top on the right edge (excluding the band's top corner)
    perim_left = []  # per band: ids bottom->top on the left edge
    shared_bottom = None
    band_rims = []

    def key_xy(p):
        return ("f", round(p[0], 12), round(p[1], 12), 0.0)

    for bi, band in enumerate(bands):
        y0, y1 = band["y0"], band["y1"]
        n_h = max(2, int(round((y1 - y0) / h)))
        while (n_w + n_h) % 8:
            n_h += 1  # ring size 2*(n_w+n_h) divisible by 16, so rings can halve toward the hole
        # perimeter ccw from (x0, y0)
        bottom = _edge_points((x0, y0, 0.0), (x1, y0, 0.0), n_w)
        right = _edge_points((x1, y0, 0.0), (x1, y1, 0.0), n_h)
        top = _edge_points((x1, y1, 0.0), (x0, y1, 0.0), n_w)
        left = _edge_points((x0, y1, 0.0), (x0, y0, 0.0), n_h)
        pts = bottom[:-1] + right[:-1] + top[:-1] + left[:-1]
        ids = [b.vertex(p, key_xy(p)) for p in pts]
        perim_right.append([b.vertex(p, key_xy(p)) for p in right])
        perim_left.append([b.vertex(p, key_xy(p)) for p in left[::-1]])
        c = np.array([band["cx"], band["cy"], 0.0])
        rad = band["radius"]
        angles = [math.atan2(p[1] - c[1], p[0] - c[0]) for p in pts]
        rim_pts = [c + rad * np.array([math.cos(a), math.sin(a), 0.0]) for a in angles]
        dists = [float(np.linalg.norm(np.asarray(p) - c)) - rad for p in pts]
        n_s = max(1, int(round(max(dists) / h)))
        prev = ids
        sel = list(range(len(pts)))
        dmean = sum(dists) / len(dists)
        for s in range(1, n_s + 1):
            frac = 1.0 - s / n_s
            spacing = 2.0 * math.pi * (rad + frac * dmean) / len(sel)
            if len(sel) % 2 == 0 and len(sel) // 2 >= 16 and spacing < 0.55 * h:
                sel = sel[::2]
            ring = [rim_pts[k] + (np.asarray(pts[k]) - rim_pts[k]) * frac for k in sel]
            ring_ids = [b.vertex(p) for p in ring]
            _ring_strip(b, ring_ids, prev, TAG_WALL, lambda *a: np.array([0.0, 0.0, 1.0]))
            prev = ring_ids
        band_rims.append((prev, c, rad, band))
        shared_bottom = top
    del shared_bottom

    # full front perimeter ccw (viewed from +z)
    first = bands[0]
    last = bands[-1]
    bottom_row = _edge_points((x0, first["y0"], 0.0), (x1, first["y0"], 0.0), n_w)
    top_row = _edge_points((x1, last["y1"], 0.0), (x0, last["y1"], 0.0), n_w)
    perim = [b.vertex(p, key_xy(p)) for p in bottom_row[:-1]]
    for r in perim_right:
        perim += r[:-1]
    perim += [b.vertex(p, key_xy(p)) for p in top_row[:-1]]
    for lft in reversed(perim_left):
        perim += lft[::-1][:-1]
    perim_xy = [np.asarray(b.v[i]) for i in perim]
    center_xy = np.array([0.0, 0.5 * (first["y0"] + last["y1"]), 0.0])

    # side walls
    prev = perim
    for j in range(1, depth_n + 1):
        z = -depth * j / depth_n
        ring = [b.vertex(np.array([p[0], p[1], z])) for p in perim_xy]
        for k in range(len(ring)):
            k1 = (k + 1) % len(ring)
            a, bb, c, d = prev[k], prev[k1], ring[k1], ring[k]
            mid = 0.25 * (np.asarray(b.v[a]) + np.asarray(b.v[bb]) + np.asarray(b.v[c]) + np.asarray(b.v[d]))
            out = mid - np.array([center_xy[0], center_xy[1], mid[2]])
            b.tri(a, bb, c, TAG_WALL, out)
            b.tri(a, c, d, TAG_WALL, out)
        prev = ring
    # back face: tensor grid whose boundary nodes are the last side ring
    xs = [p[0] for p in bottom_row]
    ys = [first["y0"]]
    for r in perim_right:
        ys += [b.v[i][1] for i in r[1:]]
    back_key = lambda ix, iy: ("b", ix, iy)  # noqa: E731
    grid = {}
    ring_pos = {(round(p[0], 12), round(p[1], 12)): rid for p, rid in zip(perim_xy, prev)}
    for iy, yv in enumerate(ys):
        for ix, xv in enumerate(xs):
            rid = ring_pos.get((round(xv, 12), round(yv, 12)))
            grid[ix, iy] = rid if rid is not None else b.vertex(np.array([xv, yv, -depth]), back_key(ix, iy))
    down = np.array([0.0, 0.0, -1.0])
    for iy in range(len(ys) - 1):
        for ix in range(len(xs) - 1):
            a, bb, c, d = grid[ix, iy], grid[ix + 1, iy], grid[ix + 1, iy + 1], grid[ix, iy + 1]
            if (ix + iy) % 2 == 0:
                b.tri(a, bb, c, TAG_WALL, down)
                b.tri(a, c, d, TAG_WALL, down)
            else:
                b.tri(a, bb, d, TAG_WALL, down)
                b.tri(bb, c, d, TAG_WALL, down)

    # holes
    for rim_ids, c, rad, band in band_rims:
        if band["kind"] == "piston":
            _disk_rings(b, rim_ids, c, None, h, band["tag"], np.array([0.0, 0.0, 1.0]))
            continue
        hl = band["horn_length"]
        rt = band["throat_radius"]
        m = math.log(rad / rt) / hl
        # march rings from mouth (z=0) to throat (z=-hl)
        rim_pts = [np.asarray(b.v[i]) for i in rim_ids]
        angles = [math.atan2(p[1] - c[1], p[0] - c[0]) for p in rim_pts]
        zs = [0.0]
        counts = [len(angles)]
        z = 0.0
        mcount = len(angles)
        while z > -hl + 1e-12:
            r = rt * math.exp(m * (hl + z))
            circ = 2.0 * math.pi * r / mcount
            if mcount % 2 == 0 and mcount // 2 >= 8 and circ < 0.55 * h:
                mcount //= 2
                circ *= 2.0
            step = min(h, max(circ, 0.15 * h))
            z -= step / math.sqrt(1.0 + (m * r) ** 2)
            zs.append(z)
            counts.append(mcount)
        scale = hl / -zs[-1]
        zs = [zz * scale for zz in zs]
        ring_ids, ring_angles = list(rim_ids), angles
        for j in range(1, len(zs)):
            zz = zs[j]
            r = rt * math.exp(m * (hl + zz))
            if counts[j] < len(ring_ids):
                ring_angles = ring_angles[::2]
            pts = [np.array([c[0] + r * math.cos(a), c[1] + r * math.sin(a), zz]) for a in ring_angles]
            ids = [b.vertex(p) for p in pts]

            def inward(a, bb, cc, _c=c):
                mid = (np.asarray(b.v[a]) + np.asarray(b.v[bb]) + np.asarray(b.v[cc])) / 3.0
                return np.array([_c[0] - mid[0], _c[1] - mid[1], 0.0])

            _ring_strip(b, ids, ring_ids, TAG_WALL, inward)
            ring_ids = ids
        _disk_rings(b, ring_ids, np.array([c[0], c[1], -hl]), None, h, band["tag"], np.array([0.0, 0.0, 1.0]))

    V, F, T = b.arrays()
    if jitter > 0.0:
        V = _jitter(V, F, jitter, seed)
    return V, F, T


def _jitter(V, F, frac, seed):
    rng = np.random.default_rng(seed)
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    lengths = np.linalg.norm(V[e[:, 0]] - V[e[:, 1]], axis=1)
    acc = np.zeros(len(V))
    cnt = np.zeros(len(V))
    np.add.at(acc, e[:, 0], lengths)
    np.add.at(acc, e[:, 1], lengths)
    np.add.at(cnt, e[:, 0], 1)
    np.add.at(cnt, e[:, 1], 1)
    local = acc / np.maximum(cnt, 1)
    return V + rng.uniform(-frac, frac, size=V.shape) * local[:, None]


def horn_in_box(h: float = 0.0225, seed: int = 1234, jitter: float = 0.05, width=0.4, height=0.6, depth=0.3,
                mouth_radius=0.15, throat_radius=0.0125, horn_length=0.2):
    """Exponential horn (throat 25 mm, mouth 300 mm, length 200 mm) in a closed 0.4x0.6x0.3 m box.

    Throat disk carries tag 2 (radiator); everything else tag 1. ``h`` is the target edge length;
    the default gives 6,000 triangles (3,002 vertices).
    """
    bands = [dict(y0=-height / 2, y1=height / 2, cx=0.0, cy=0.0, radius=mouth_radius, kind="horn",
                  tag=TAG_THROAT, horn_length=horn_length, throat_radius=throat_radius)]
    return _box_with_holes(width, height, depth, bands, h, seed=seed, jitter=jitter)


def loudspeaker_box(h: float = 0.017, seed: int = 4321, jitter: float = 0.05, width=0.5, height=0.9, depth=0.4,
                    mouth_radius=0.15, throat_radius=0.0125, horn_length=0.2, piston_radii=(0.08, 0.11)):
    """Two-way loudspeaker-like box: horn (tag 2) over two piston cones (tags 3, 4)."""
    y0 = -height / 2
    b_low = 2.0 * piston_radii[1] + 0.08
    b_mid = 2.0 * piston_radii[0] + 0.08
    bands = [
        dict(y0=y0, y1=y0 + b_low, cx=0.0, cy=y0 + b_low / 2, radius=piston_radii[1], kind="piston", tag=4),
        dict(y0=y0 + b_low, y1=y0 + b_low + b_mid, cx=0.0, cy=y0 + b_low + b_mid / 2, radius=piston_radii[0],
             kind="piston", tag=3),
        dict(y0=y0 + b_low + b_mid, y1=height / 2, cx=0.0, cy=0.5 * (y0 + b_low + b_mid + height / 2),
             radius=mouth_radius, kind="horn", tag=TAG_THROAT, horn_length=horn_length, throat_radius=throat_radius),
    ]
    return _box_with_holes(width, height, depth, bands, h, seed=seed, jitter=jitter)


def box_with_pistons(h: float = 0.02, seed: int = 99, jitter: float = 0.05, width=0.4, height=0.7, depth=0.3,
                     piston_radii=(0.06, 0.09)):
    """Closed box with flush piston disks (tags 2, 3, ...) — the serving stream's third generator."""
    y0 = -height / 2
    n = len(piston_radii)
    bh = height / n
    bands = [
        dict(y0=y0 + i * bh, y1=y0 + (i + 1) * bh, cx=0.0, cy=y0 + (i + 0.5) * bh, radius=r, kind="piston",
             tag=2 + i)
        for i, r in enumerate(piston_radii)
    ]
    return _box_with_holes(width, height, depth, bands, h, seed=seed, jitter=jitter)


def mesh_for_triangle_count(kind: str, target: int, seed: int = 0):
    """Generate ``kind`` ('sphere'|'horn'|'pistons') near ``target`` triangles (edge length by area scaling)."""
    if kind == "sphere":
        # icospheres have 20 x 4^n triangles: the level nearest the target in log scale, one level lower when
        # that one exceeds the target by more than 30 % (a 15,000 target gives 5,120, not 20,480)
        n = max(1, int(round(math.log(max(target, 20) / 20.0, 4))))
        if n > 1 and 20 * 4 ** n > 1.3 * target:
            n -= 1
        return icosphere(n, 0.1)
    gen = {"horn": horn_in_box, "pistons": box_with_pistons}[kind]
    h = 0.03
    for _ in range(6):
        V, F, T = gen(h=h, seed=seed)
        ratio = len(F) / float(target)
        if abs(ratio - 1.0) < 0.08:
            break
        h *= math.sqrt(ratio)
    return V, F, T


# ----------------------------------------------------------------------------------------------
# checks and IO
# ----------------------------------------------------------------------------------------------


def check_closed_oriented(F: np.ndarray) -> bool:
    """Every directed edge appears once and its reverse once: closed and consistently oriented."""
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    directed = {(int(a), int(b)) for a, b in e}
    if len(directed) != len(e):
        return False
    return all((b, a) in directed for a, b in directed)


def signed_volume(V: np.ndarray, F: np.ndarray) -> float:
    p = V[F]
    return float(np.einsum("ij,ij->i", p[:, 0], np.cross(p[:, 1], p[:, 2])).sum() / 6.0)


def write_gmsh22(path, V, F, T, scale: float = 1.0):
    """Write a Gmsh 2.2 ASCII surface mesh (coordinates divided by ``scale``: 0.001 -> file in mm)."""
    with open(path, "w", encoding="utf-8") as out:
        out.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(V))
        for i, p in enumerate(np.asarray(V) / scale, start=1):
            out.write("%d %.17g %.17g %.17g\n" % (i, p[0], p[1], p[2]))
        out.write("$EndNodes\n$Elements\n%d\n" % len(F))
        for i, (f, t) in enumerate(zip(F, T), start=1):
            out.write("%d 2 2 %d %d %d %d %d\n" % (i, t, t, f[0] + 1, f[1] + 1, f[2] + 1))
        out.write("$EndElements\n")


# ---------------------------------------------------------------------------------------------
# Tetrahedral volumes for the coupled FEM-BEM path
# ---------------------------------------------------------------------------------------------

TAG_VOLUME, TAG_INNER, TAG_OUTER = 1, 2, 3


def prism_layers_to_tets(faces: np.ndarray, n_surface: int, n_layers: int) -> np.ndarray:
    """Tetrahedra of the prisms between consecutive copies of a triangulated surface (vertex v of layer j is
    j * n_surface + v). Each prism (a, b, c) with a < b < c splits into (a, b, c, c'), (a, b, b', c'),
    (a, a', b', c'): the diagonal of every side quad joins the lower-numbered bottom vertex to the top copy of the
    higher-numbered one, so neighbouring prisms agree on it and the mesh is conforming."""
    s = np.sort(np.asarray(faces, dtype=np.int64), axis=1)
    a, b, c = s[:, 0], s[:, 1], s[:, 2]
    out = []
    for j in range(n_layers):
        lo, hi = j * n_surface, (j + 1) * n_surface
        out += [np.stack([a + lo, b + lo, c + lo, c + hi], 1), np.stack([a + lo, b + lo, b + hi, c + hi], 1),
                np.stack([a + lo, a + hi, b + hi, c + hi], 1)]
    return np.concatenate(out)


def spherical_shell(subdivisions: int = 2, inner: float = 0.1, outer: float = 0.15, layers: int = 2):
    """Air between two concentric icospheres as a tetrahedral volume (tag TAG_VOLUME), its inner sphere faces
    (TAG_INNER) and outer sphere faces (TAG_OUTER) as boundary triangles, and the outer sphere as a BEM surface
    with outward normals (tag TAG_OUTER, same vertex positions). Radii of the layers grow linearly.
    Returns (VolumeMesh, (V, F, T) of the BEM surface)."""
    from .fem import VolumeMesh

    V, F, _ = icosphere(subdivisions, 1.0)
    nv = len(V)
    radii = np.linspace(inner, outer, layers + 1)
    verts = np.concatenate([V * r for r in radii])
    tets = prism_layers_to_tets(F, nv, layers)
    bfaces = np.concatenate([F, F + layers * nv])
    btags = np.concatenate([np.full(len(F), TAG_INNER), np.full(len(F), TAG_OUTER)])
    names = {(3, TAG_VOLUME): "Volume", (2, TAG_INNER): "Radiator", (2, TAG_OUTER): "Interface"}
    vm = VolumeMesh(verts, tets, np.full(len(tets), TAG_VOLUME), bfaces, btags, names)
    return vm, (V * outer, F.copy(), np.full(len(F), TAG_OUTER))


def structured_cube(n: int = 4, size: float = 1.0):
    """The cube [0, size]^3 cut into n^3 hexahedra of six tetrahedra each (Kuhn split along the main diagonal,
    conforming), boundary faces tagged 1..6 by side (x-, x+, y-, y+, z-, z+), volume tag 1."""
    from .fem import VolumeMesh

    g = np.linspace(0.0, size, n + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    verts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)

    def vid(i, j, k):
        return (i * (n + 1) + j) * (n + 1) + k

    tets = []
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    for i in range(n):
        for j in range(n):
            for k in range(n):
                base = np.array([i, j, k])
                for p in perms:
                    path = [base.copy()]
                    cur = base.copy()
                    for ax in p:
                        cur = cur.copy()
                        cur[ax] += 1
                        path.append(cur)
                    tets.append([vid(*q) for q in path])
    tets = np.array(tets, dtype=np.int64)
    # boundary faces: tetrahedral faces on the cube's sides
    f = np.concatenate([tets[:, [0, 1, 2]], tets[:, [0, 1, 3]], tets[:, [0, 2, 3]], tets[:, [1, 2, 3]]])
    key = np.sort(f, 1)
    uniq, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    bf = uniq[cnt == 1]
    c = verts[bf].mean(1)
    tag = np.zeros(len(bf), np.int64)
    eps = 1e-12 * size
    for t, (ax, val) in enumerate([(0, 0.0), (0, size), (1, 0.0), (1, size), (2, 0.0), (2, size)]):
        tag[np.abs(c[:, ax] - val) <= eps] = t + 1
    return VolumeMesh(verts, tets, np.ones(len(tets), np.int64), bf, tag, {(3, 1): "Volume"})
