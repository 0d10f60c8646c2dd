"""Host mesh preparation (product) against the oracle's independent restatement, plus table invariants."""

import os

import numpy as np
import pytest

from beat_hip import geometry as G
from beat_hip import synthetic as syn
from oracle import bem_oracle as O

REF_MESHES = "/root/reference/src/blab/solvers/julia_local/test_meshes"


@pytest.fixture(scope="module")
def sphere():
    V, F, T = syn.icosphere(2, 0.1)
    return G.BoundaryMesh(V, F, T)


@pytest.mark.parametrize("gen", ["sphere", "horn", "pistons"])
def test_generators_closed_oriented(gen):
    V, F, T = {"sphere": lambda: syn.icosphere(2, 0.1), "horn": lambda: syn.horn_in_box(h=0.05),
               "pistons": lambda: syn.box_with_pistons(h=0.05)}[gen]()
    assert syn.check_closed_oriented(F)
    assert syn.signed_volume(V, F) > 0  # normals point out of the body, into the fluid
    assert len(V) - len(F) // 2 == 2  # Euler: V - E + F = 2 with E = 3F/2 -> genus 0
    assert 2 in set(T.tolist())


def test_default_workload_sizes():
    V, F, T = syn.horn_in_box()
    assert len(F) == 6000 and len(V) == 3002
    V, F, T = syn.icosphere(3, 0.1)
    assert len(F) == 1280 and len(V) == 642


def test_singular_pairs_match_oracle(sphere):
    prep = G.prepare(sphere, reorder=False)
    om = O.Mesh(sphere.vertices, sphere.faces, sphere.tags, sphere.mesh_ids)
    cache = O.build_singular_cache(om, 4)
    a = sorted(zip(cache.test.tolist(), cache.trial.tolist()))
    b = sorted(zip(prep.pairs.test.tolist(), prep.pairs.trial.tolist()))
    assert a == b
    ko = {(t, s): cache.keys[r] for t, s, r in zip(cache.test, cache.trial, cache.rule)}
    for t, s, r in zip(prep.pairs.test, prep.pairs.trial, prep.pairs.rule):
        key = prep.pairs.keys[r]
        assert ko[(t, s)] == key
        tp, sp, w = cache.rules[cache.keys.index(key)]
        off, n = prep.pairs.rule_off[r], prep.pairs.rule_len[r]
        got = prep.pairs.rule_pts[off : off + n]
        assert np.allclose(got[:, :2], tp, atol=1e-15) and np.allclose(got[:, 2:4], sp, atol=1e-15)
        assert np.allclose(got[:, 4], w, atol=1e-18)
    # pairs are grouped by test element (CSR)
    assert np.all(np.diff(prep.pairs.test) >= 0)
    assert prep.pairs.ptr[-1] == len(prep.pairs.test)


def test_curls_and_rule_match_oracle(sphere):
    prep = G.prepare(sphere, reorder=False)
    om = O.Mesh(sphere.vertices, sphere.faces, sphere.tags, sphere.mesh_ids)
    assert np.abs(prep.curls - O.surface_curls(om.face_vertices, om.normals)).max() < 1e-9
    pts, w = O.triangle_rule(4)
    assert np.array_equal(pts, G.RULE_POINTS) and np.array_equal(w, G.RULE_WEIGHTS)
    m11, m10 = G.mass_blocks()
    assert np.allclose(m11, (np.ones((3, 3)) + np.eye(3)) / 24.0, atol=1e-15)
    assert np.allclose(m10, 1.0 / 6.0, atol=1e-15)


def test_slot_tables_cover_every_incidence_once(sphere):
    prep = G.prepare(sphere)
    s = prep.slots
    sphere = prep.mesh  # tables index the locality-ordered mesh
    F, V = sphere.n_faces, sphere.n_vertices
    seen = np.zeros((F, 3), dtype=int)
    first = np.zeros(V, dtype=int)
    for c in range(len(s.slot_count)):
        verts = s.slot_vertex[c, : s.slot_count[c]] >> 1
        assert len(set(verts.tolist())) == len(verts)  # distinct vertices within a chunk: no lane conflicts
        for k in range(s.slot_count[c]):
            code_v = int(s.slot_vertex[c, k])
            vk, again = code_v >> 1, code_v & 1
            # a vertex without an LDS slot is stored by its first chunk and accumulated by later ones
            in_lds = (int(s.slot_pend[c, k]) & 0xFFFF) != G.PEND_NONE
            assert again == (first[vk] > 0 and not in_lds)
            first[vk] += 1
            for e in range(s.slot_start[c, k], s.slot_start[c, k + 1]):
                code = int(s.slot_inc[c, e])
                lane, m = code >> 2, code & 3
                elem = s.order[c * 64 + lane]
                assert elem >= 0
                assert sphere.faces[elem, m] == vk
                seen[elem, m] += 1
    assert np.all(seen == 1)
    assert np.all(first >= 1)
    # row CSR: strip rows 3t+l with faces[t, l] == i
    for i in range(V):
        rows = s.vstar_row[s.vstar_ptr[i] : s.vstar_ptr[i + 1]]
        assert np.all(sphere.faces[rows // 3, rows % 3] == i)
    assert s.vstar_ptr[-1] == 3 * F


@pytest.mark.parametrize("pend_slots", [G.PEND_SLOTS, 3, 0])
def test_strip_store_and_gather_reproduce_dense_assembly(sphere, pend_slots):
    # Emulate the regular kernel's strip stores with random element blocks: a multi-chunk vertex keeps its
    # partial sums in an LDS slot until its last chunk stores them; without a free slot it is stored by its first
    # chunk and read-modify-written later. Then the row gather. The result must equal the direct scatter, and no
    # LDS slot may be shared by two live vertices.
    prep = G.prepare(sphere)
    m = prep.mesh
    s = G.slot_tables(m, pend_slots=pend_slots)
    F, V = m.n_faces, m.n_vertices
    rng = np.random.default_rng(3)
    blocks = rng.standard_normal((F, F, 3, 3))  # [t, s, l, m]
    dense = np.zeros((V, V))
    for t in range(F):
        for l in range(3):
            np.add.at(dense[m.faces[t, l]], m.faces[:, :], blocks[t, :, l, :])
    strip = np.full((3 * F, V), np.nan)
    lds = {}
    owner = {}
    used_lds = 0
    for c in range(len(s.slot_count)):
        live = set()
        for k in range(s.slot_count[c]):
            vk, again = int(s.slot_vertex[c, k]) >> 1, int(s.slot_vertex[c, k]) & 1
            pc = int(s.slot_pend[c, k])
            pidx, from_lds, to_strip = pc & 0xFFFF, bool(pc & G.PEND_FROM_LDS), bool(pc & G.PEND_TO_STRIP)
            val = np.zeros((F, 3))
            for e in range(s.slot_start[c, k], s.slot_start[c, k + 1]):
                code = int(s.slot_inc[c, e])
                val += blocks[:, s.order[c * 64 + (code >> 2)], :, code & 3]
            val = val.reshape(3 * F)
            if pidx != G.PEND_NONE:
                assert pidx < pend_slots and pidx not in live
                live.add(pidx)
                used_lds += 1
                if from_lds:
                    assert owner[pidx] == vk
                    val = val + lds[pidx]
                else:
                    assert not again
                owner[pidx] = vk
                if to_strip:
                    assert np.all(np.isnan(strip[:, vk]))  # stored exactly once
                    strip[:, vk] = val
                else:
                    lds[pidx] = val
            else:
                assert to_strip
                strip[:, vk] = strip[:, vk] + val if again else val
    assert (used_lds > 0) == (pend_slots > 0)
    A = np.zeros((V, V))
    for i in range(V):
        A[i] = strip[s.vstar_row[s.vstar_ptr[i] : s.vstar_ptr[i + 1]]].sum(axis=0)
    assert np.allclose(A, dense, rtol=1e-12, atol=1e-12)


def test_gmsh_roundtrip(tmp_path, sphere):
    path = tmp_path / "s.msh"
    syn.write_gmsh22(path, sphere.vertices, sphere.faces, sphere.tags, scale=0.001)
    V, F, T = G.read_gmsh22(path, 0.001)
    Vo, Fo, To = O.load_gmsh22(str(path), 0.001)
    assert np.allclose(V, sphere.vertices, atol=1e-15) and np.array_equal(F, sphere.faces)
    assert np.array_equal(V, Vo) and np.array_equal(F, Fo) and np.array_equal(T, To)


@pytest.mark.skipif(not os.path.isdir(REF_MESHES), reason="reference meshes not present")
def test_reference_meshes_load_and_prepare():
    # read in place (GPL fixtures are not copied); counts from the reference's own test (runtests.jl:50-63)
    V, F, T = G.read_gmsh22(os.path.join(REF_MESHES, "sample.msh"), 0.001)
    assert len(F) == 2776 and len(V) == 1390
    assert np.sum(T == 2) == 128 and np.sum(T == 1) == 2648
    V2, F2, T2 = G.read_gmsh22(os.path.join(REF_MESHES, "two_tetrahedra.msh"), 1.0)
    assert len(F2) == 8 and len(V2) == 8
    prep = G.prepare(G.BoundaryMesh(V, F, T))
    om = O.Mesh(V, F, T, np.ones(len(F), int))
    cache = O.build_singular_cache(om, 4)
    assert len(cache.test) == len(prep.pairs.test)


def test_locality_order_is_a_consistent_permutation(sphere):
    prep = G.prepare(sphere)
    assert sorted(prep.fperm.tolist()) == list(range(sphere.n_faces))
    assert sorted(prep.vperm.tolist()) == list(range(sphere.n_vertices))
    # internal face i is original face fperm[i] with vertices renamed by vperm
    assert np.array_equal(prep.vperm[prep.mesh.faces], sphere.faces[prep.fperm])
    assert np.allclose(prep.mesh.normals, sphere.normals[prep.fperm])
    # the singular pair set maps onto the original one
    ref = G.prepare(sphere, reorder=False).pairs
    a = sorted(zip(prep.fperm[prep.pairs.test].tolist(), prep.fperm[prep.pairs.trial].tolist()))
    b = sorted(zip(ref.test.tolist(), ref.trial.tolist()))
    assert a == b
    x = np.arange(sphere.n_vertices, dtype=float)[prep.vperm]
    assert np.array_equal(prep.unpermute_vertices(x), np.arange(sphere.n_vertices))
    # locality: consecutive internal vertices are close in space on average
    d = np.linalg.norm(np.diff(prep.mesh.vertices, axis=0), axis=1).mean()
    d0 = np.linalg.norm(np.diff(sphere.vertices, axis=0), axis=1).mean()
    assert d <= d0


def test_locality_order_permutes_the_operators():
    """The oracle's operators on the renumbered mesh are the original ones with rows/columns taken by vperm."""
    V, F, T = syn.icosphere(1, 0.1)
    prep = G.prepare(G.BoundaryMesh(V, F, T))
    assert not np.array_equal(prep.vperm, np.arange(len(V)))
    q = np.arange(1.0, len(F) + 1.0)[:, None]
    outs = []
    for m, qq in ((prep.original, q), (prep.mesh, q[prep.fperm])):
        om = O.Mesh(m.vertices, m.faces, m.tags, m.mesh_ids)
        outs.append(O.solve_frequency(om, 1500.0, qq, 343.0, 1.205))
    orig, internal = outs
    vp = prep.vperm
    assert np.abs(internal["lhs"] - orig["lhs"][np.ix_(vp, vp)]).max() < 1e-12 * np.abs(orig["lhs"]).max()
    assert np.abs(internal["rhs"] - orig["rhs"][vp]).max() < 1e-12 * np.abs(orig["rhs"]).max()
    assert np.allclose(prep.unpermute_vertices(internal["pressure"]), orig["pressure"], rtol=1e-10, atol=0)


@pytest.mark.parametrize("mode", ["x", "xy"])
def test_image_pairs_match_oracle(mode):
    """Symmetric preparation: image pairs (trial key s + r F), their orientation keys and Duffy rules equal the
    oracle's image caches; masks cover every image pair; orbit weights as the oracle."""
    V, F, T = syn.uv_sphere(6, 12)
    v, f, t = syn.symmetric_part(V, F, T, mode)
    mesh = G.BoundaryMesh(v, f, t)
    prep = G.prepare(mesh, reorder=False, symmetry=mode)
    om = O.Mesh(v, f, t, np.ones(len(f), int))
    nF = len(f)
    pr = prep.pairs
    got = {}
    for tt, key, r in zip(pr.test, pr.trial, pr.rule):
        got[(int(tt), int(key))] = pr.keys[r]
    expect = {}
    base = O.build_singular_cache(om, 4)
    for tt, ss, r in zip(base.test, base.trial, base.rule):
        expect[(int(tt), int(ss))] = base.keys[r]
    tfs = O.symmetry_transforms(mode)
    for ri, (_, signs, _) in enumerate(tfs[1:], start=1):
        img = O.image_singular_cache(om, 4, signs)
        for tt, ss, r in zip(img.test, img.trial, img.rule):
            expect[(int(tt), int(ss) + ri * nF)] = img.keys[r]
    assert got == expect
    assert any(key >= nF for (_, key) in got)
    # rule points of image pairs equal the oracle's remapped rules
    for (tt, key), k in got.items():
        if key >= nF:
            ri = key // nF
            img = O.image_singular_cache(om, 4, tfs[ri][1])
            tp, sp, w = img.rules[img.keys.index(k)]
            r = pr.keys.index(k)
            pts = pr.rule_pts[pr.rule_off[r]: pr.rule_off[r] + pr.rule_len[r]]
            assert np.allclose(pts[:, :2], tp, atol=1e-15) and np.allclose(pts[:, 2:4], sp, atol=1e-15)
            break
    img_pairs = [(tt, key) for (tt, key) in got if key >= nF]
    for tt, key in img_pairs:
        bit = 1 << (key // nF)
        assert prep.plane_mask[tt] & bit and prep.plane_mask[key % nF] & bit
    assert np.array_equal(prep.row_weights, O.p1_orbit_weights(v, mode))
    assert np.array_equal(prep.transforms, np.array([s for _, s, _ in tfs], dtype=np.int32))


def test_symmetry_snap_and_validation():
    V, F, T = syn.uv_sphere(6, 12)
    v, f, t = syn.symmetric_part(V, F, T, "x")
    v = v.copy()
    v[v[:, 0] == 0.0, 0] = 3e-11  # within the plane tolerance
    m = G.snap_symmetry_planes(G.BoundaryMesh(v, f, t), "x")
    assert np.sum(m.vertices[:, 0] == 0.0) == np.sum(v[:, 0] == 3e-11)
    G.validate_symmetry_domain(m, "x")
    bad = G.BoundaryMesh(v * np.array([-1.0, 1.0, 1.0]), f[:, ::-1], t)
    with pytest.raises(ValueError, match="positive X fundamental domain"):
        G.validate_symmetry_domain(bad, "x")
    assert G.symmetry_reduction_factor("xy") == 4 and G.symmetry_reduction_factor("off") == 1
    with pytest.raises(ValueError):
        G.normalized_symmetry_mode("z")


@pytest.mark.parametrize("pend_slots", [G.PEND_SLOTS, 2])
def test_paired_strips_reproduce_dense_assembly(pend_slots):
    # Emulate the paired regular assembly: the wave of test element t visits chunks from chunk_start[t]; a trial
    # element s after t that shares no vertex with t gives the (t, s) block to t's row strip and the transposed
    # (s, t) block to t's column strip; an element sharing a vertex gives its own block to the row strip only;
    # earlier non-adjacent elements give nothing. Partial sums follow the LDS / read-modify-write rules with the
    # previous-chunk test; the gather masks columns by the vertex's last chunk. Must equal the dense scatter.
    V, F, T = syn.icosphere(2, 0.1)  # 320 faces: 5 chunks
    m = G.prepare(G.BoundaryMesh(V, F, T)).mesh
    s = G.slot_tables(m, pend_slots=pend_slots)
    F, V = m.n_faces, m.n_vertices
    rng = np.random.default_rng(5)
    B = rng.standard_normal((F, F, 3, 3))  # block (t, s)[l][m] for every ordered pair
    dense = np.zeros((V, V))
    for t in range(F):
        for l in range(3):
            np.add.at(dense[m.faces[t, l]], m.faces[:, :], B[t, :, l, :])
    pos = s.elem_pos
    fv = [set(f) for f in m.faces.tolist()]
    RS = np.full((3 * F, V), np.nan)
    CS = np.full((3 * F, V), np.nan)
    nch = len(s.slot_count)
    for t in range(F):
        c0 = int(s.chunk_start[t])
        assert c0 <= pos[t] // 64
        lds = {}
        for c in range(c0, nch):
            blk = np.zeros((64, 3, 3))
            blkp = np.zeros((64, 3, 3))
            for lane in range(64):
                sidx = s.order[c * 64 + lane]
                if sidx < 0:
                    continue
                if fv[t] & fv[sidx]:
                    blk[lane] = B[t, sidx]
                    assert pos[sidx] // 64 >= c0
                elif pos[sidx] > pos[t]:
                    blk[lane] = B[t, sidx]
                    blkp[lane] = B[sidx, t].T
            for k in range(s.slot_count[c]):
                vk, again = int(s.slot_vertex[c, k]) >> 1, bool(int(s.slot_vertex[c, k]) & 1)
                pc, prev = int(s.slot_pend[c, k]), int(s.slot_prev[c, k])
                pidx = pc & 0xFFFF
                vals = []
                for bb in (blk, blkp):
                    a = np.zeros(3)
                    for e in range(s.slot_start[c, k], s.slot_start[c, k + 1]):
                        code = int(s.slot_inc[c, e])
                        a += bb[code >> 2, :, code & 3]
                    vals.append(a)
                for w, (strip, a) in enumerate(zip((RS, CS), vals)):
                    # the column strip only from t's own chunk on (no later element sits in an earlier chunk)
                    cw = c0 if w == 0 else int(pos[t] // 64)
                    if c < cw:
                        continue
                    from_lds = bool(pc & G.PEND_FROM_LDS) and prev >= cw
                    again_w = bool(int(s.slot_vertex[c, k]) & 1) and prev >= cw
                    rows = 3 * t + np.arange(3)
                    if pidx != G.PEND_NONE:
                        if from_lds:
                            a = a + lds[(pidx, w)]
                        if pc & G.PEND_TO_STRIP:
                            strip[rows, vk] = a
                        else:
                            lds[(pidx, w)] = a
                    else:
                        strip[rows, vk] = strip[rows, vk] + a if again_w else a
    A = np.zeros((V, V))
    cs_row = np.repeat(s.chunk_start, 3)
    cs_col = np.repeat(s.elem_pos // 64, 3)
    for i in range(V):
        for r in s.vstar_row[s.vstar_ptr[i] : s.vstar_ptr[i + 1]]:
            ok = s.vertex_last >= cs_row[r]
            A[i, ok] += RS[r, ok]
            ok = s.vertex_last >= cs_col[r]
            A[ok, i] += CS[r, ok]
    assert np.allclose(A, dense, rtol=1e-12, atol=1e-12)


def test_gmsh_block_reader_matches_line_reader(tmp_path):
    # the vectorised readers (on the whole text, on the list of lines) and the line loop give the same mesh; files
    # they cannot take as uniform blocks (mixed element lines, sparse node ids with unknown references) fall back to
    # the loop with the same result
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn

    V, F, T = syn.horn_in_box(h=0.05)
    p = tmp_path / "h.msh"
    syn.write_gmsh22(p, V, F, T, scale=0.001)
    fast = G.read_gmsh22(p, 0.001, (0.1, 0.0, -0.2))
    assert G._gmsh22_blocks is not None
    mixed = tmp_path / "mixed.msh"
    mixed.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n5\n10 0 0 0\n20 1 0 0\n30 0 1 0\n40 0 0 1\n"
                     "50 9 9 9\n$EndNodes\n$Elements\n5\n1 15 2 0 0 50\n2 1 2 7 7 10 20\n3 2 2 3 3 10 30 20\n"
                     "4 2 3 4 4 9 10 20 40\n5 2 2 5 5 20 30 40\n$EndElements\n")
    vm, fm, tm = G.read_gmsh22(mixed)
    assert fm.tolist() == [[0, 2, 1], [0, 1, 3], [1, 2, 3]] and tm.tolist() == [3, 4, 5]
    assert np.allclose(vm[4], [9, 9, 9])
    # Windows line ends and blanks around the section markers: the same mesh
    crlf = tmp_path / "crlf.msh"
    crlf.write_bytes(p.read_text().replace("$Nodes\n", "  $Nodes \n").replace("\n", "\r\n").encode())
    fast_crlf = G.read_gmsh22(crlf, 0.001, (0.1, 0.0, -0.2))
    orig_text, orig_blocks = G._gmsh22_text, G._gmsh22_blocks
    try:
        G._gmsh22_text = lambda *a: None  # the line-list block reader
        blocks = G.read_gmsh22(p, 0.001, (0.1, 0.0, -0.2))
        G._gmsh22_blocks = lambda *a: None  # the line loop
        loop = G.read_gmsh22(p, 0.001, (0.1, 0.0, -0.2))
        loop_mixed = G.read_gmsh22(mixed)
        loop_crlf = G.read_gmsh22(crlf, 0.001, (0.1, 0.0, -0.2))
    finally:
        G._gmsh22_text, G._gmsh22_blocks = orig_text, orig_blocks
    for a, b, c, d, e in zip(fast, blocks, loop, fast_crlf, loop_crlf):
        assert np.array_equal(a, b) and np.array_equal(a, c) and np.array_equal(a, d) and np.array_equal(a, e)
    for a, b in zip((vm, fm, tm), loop_mixed):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("case", ["horn", "sphere", "fan", "pistons", "ragged", "x-sym"])
def test_native_tables_equal_numpy_tables(case):
    # the library's host builders (bem_host_vertex_pairs, bem_host_slot_tables) produce exactly the tables of the
    # NumPy restatement: slot tables, pair lists, Duffy rule indices, rule points and orientation keys
    from beat_hip import geometry as G
    from beat_hip import synthetic as syn

    V, F, T = {"horn": lambda: syn.horn_in_box(), "sphere": lambda: syn.icosphere(3, 0.1),
               "fan": lambda: syn.uv_sphere(8, 80, 0.1), "pistons": lambda: syn.box_with_pistons(h=0.05),
               "ragged": lambda: syn.horn_in_box(h=0.05), "x-sym": lambda: syn.uv_sphere(10, 20, 0.1)}[case]()
    mesh = G.BoundaryMesh(V, F, T)
    sym = "off"
    if case == "x-sym":  # the half x >= 0 of a mirror-symmetric sphere
        keep = (mesh.face_vertices[:, :, 0] >= -1e-12).all(axis=1)
        used = np.unique(mesh.faces[keep])
        remap = -np.ones(len(V), int)
        remap[used] = np.arange(len(used))
        mesh = G.snap_symmetry_planes(G.BoundaryMesh(V[used], remap[mesh.faces[keep]], T[keep]), "x")
        sym = "x"
    a = G.prepare(mesh, symmetry=sym)
    fp, vp = a.fperm, a.vperm
    w = a.mesh
    b_slots = G.slot_tables(w, native=False)
    for f in b_slots.__dataclass_fields__:
        assert np.array_equal(np.asarray(getattr(a.slots, f)), np.asarray(getattr(b_slots, f))), f
    signs = G.SYMMETRY_SIGNS[sym]
    b_pairs = G.singular_pairs(w, 4, transforms=signs if len(signs) > 1 else None, native=False)
    for f in ("test", "trial", "rule", "ptr", "rule_off", "rule_len", "rule_pts"):
        assert np.array_equal(getattr(a.pairs, f), getattr(b_pairs, f)), f
    assert a.pairs.keys == b_pairs.keys


def test_native_table_builders_reject_bad_input():
    import ctypes as C

    from beat_hip import _lib

    lib = _lib.load()
    faces = np.array([[0, 1, 2], [0, 2, 5]], np.int32)
    out = [np.zeros(64, np.int32) for _ in range(4)]
    n, nc = C.c_longlong(), C.c_int()
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    assert lib.bem_host_vertex_pairs(2, 4, ip(faces), 64, *[ip(o) for o in out], C.byref(nc), C.byref(n)) == \
        _lib.BEM_E_ARG  # vertex 5 >= V
    faces[1, 2] = 3
    assert lib.bem_host_vertex_pairs(2, 4, ip(faces), 2, *[ip(o) for o in out], C.byref(nc), C.byref(n)) == \
        _lib.BEM_E_ARG  # capacity too small (4 pairs)
    assert lib.bem_host_vertex_pairs(2, 4, ip(faces), 64, *[ip(o) for o in out], C.byref(nc), C.byref(n)) == 0
    # two self pairs (coincident) and the shared edge in both directions, whose orientations differ:
    # (0, 1): test corners 0, 2 meet trial corners 0, 1 -> 200201; (1, 0): test 0, 1 meet trial 0, 2 -> 200102
    assert n.value == 4 and nc.value == 3
    assert out[3][:3].tolist() == [100000, 200102, 200201]
