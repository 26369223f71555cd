"""Convex hulls and triangle meshes (MR_SHAPE_CONVEX, MR_SHAPE_TRIANGLE_MESH).

CPU, the oracle against closed forms and brute force:
  * a hull built from a box's eight vertices reproduces the box primitive: on
    a plane (four bottom-vertex contacts), against spheres at random poses
    (GJK separation and normal against sphere_box), face to face on a box
    (the four corners of the patch), and along a face normal when
    overlapping (MPR depth);
  * GJK distances between random hulls against an FP64 brute force over
    vertex-triangle and edge-edge pairs of their scipy hulls;
  * a two-triangle floor mesh carries a resting box like a plane (front side),
    nothing from behind unless two-sided; a bumpy 512-triangle mesh under a
    small sphere yields contacts on exactly the triangles a brute-force FP64
    point-triangle distance puts within the contact distance.
"""

import numpy as np
import pytest

import metalrobo as mr
import model_util as mu
import oracle_lib as ol

SPHERE, CAPSULE, BOX, PLANE, CONVEX, MESH = 0, 1, 2, 3, 5, 6
TWO_SIDED = 4


def feature(shape_type, local):
    return ((shape_type & 0xF) << 28) | (local & 0x0FFFFFFF)


def box_vertices(he):
    return [[(-1) ** (1 - (i & 1)) * he[0], (-1) ** (1 - ((i >> 1) & 1)) * he[1], (-1) ** (1 - ((i >> 2) & 1)) * he[2]]
            for i in range(8)]


def scene(specs, geoms=None, plane=True, envs=1):
    """specs: (type, dims, pos, rot, geometry index, flags); one dynamic body
    each, then a plane; mesh shapes get static bodies."""
    bodies = [mu.body(1.0, [0.01] * 3)]
    shapes, states = [], []
    for i, (t, dims, pos, rot, gi, fl) in enumerate(specs):
        static = t == MESH
        bodies.append(mu.body(0.0 if static else 1.0, [1.0] * 3 if static else [0.01] * 3, motion=0 if static else 2))
        shapes.append(mu.shape(1 + i, t, dims, contact_offset=0.01, geometry=gi, flags=fl))
        states.append(mu.scene_state(pos, rot, inv_mass=0.0 if static else 1.0, motion=0 if static else 2))
    if plane:
        bodies.append(mu.body(0.0, [1, 1, 1], motion=0))
        shapes.append(mu.shape(len(bodies) - 1, PLANE, [0, 0, 0], rot=mu.Y_TO_Z, contact_offset=0.01))
        states.append(mu.scene_state((0, 0, 0), motion=0))
    w = mu.CustomWorld(bodies, [], shapes, link_count=1, root_type=0, nq=0, nv=0, geometries=geoms, envs=envs)
    st = np.repeat(np.array(states, w.model.initial_state()[2].dtype)[None, :], envs, axis=0)
    return w, st


def collide(w, st):
    return ol.collide(w.model, w.config(), 0, np.zeros(1, np.float32), np.zeros(1, np.float32), st[0])


def pair_index(w, a, b):
    for k, p in enumerate(w.pairs):
        if p["collider_a"] == a and p["collider_b"] == b:
            return k
    raise KeyError((a, b))


def quat(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return tuple(np.r_[np.sin(angle / 2) * axis, np.cos(angle / 2)])


HE = [0.1, 0.07, 0.05]
HULL_BOX = [(CONVEX, box_vertices(HE), None)]


def test_hull_of_a_box_on_a_plane_matches_the_box():
    for t, dims, gi in ((BOX, HE, 0), (CONVEX, [1, 1, 1], 0)):
        w, st = scene([(t, dims, (0.0, 0.0, 0.05), (0, 0, 0, 1), gi, 0)], HULL_BOX)
        ov, cnt, reals, feats = collide(w, st)
        k = pair_index(w, 0, 1)
        assert cnt[k] == 4, (t, cnt[k])
        np.testing.assert_allclose(reals[k, :4, 3], 0.0, atol=1e-6)
        assert sorted(feats[k, :4, 0].tolist()) == [feature(t, v) for v in range(4)]


def test_hull_of_a_box_against_spheres_matches_sphere_box():
    rng = np.random.default_rng(1)
    for trial in range(20):
        rot = quat(rng.normal(size=3), rng.uniform(0, np.pi))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pos = tuple(0.5 + 0.0 * d)
        spos = tuple(np.array(pos) + d * rng.uniform(0.16, 0.3))
        out = {}
        for t, dims in ((BOX, HE), (CONVEX, [1, 1, 1])):
            w, st = scene([(SPHERE, [0.04], spos, (0, 0, 0, 1), 0, 0), (t, dims, pos, rot, 0, 0)], HULL_BOX,
                          plane=False)
            ov, cnt, reals, feats = collide(w, st)
            k = pair_index(w, 0, 1)
            out[t] = (int(cnt[k]), reals[k, 0].copy())
        cb, rb = out[BOX]
        cc, rc = out[CONVEX]
        assert cb == cc, (trial, cb, cc)
        if cb:
            assert abs(rb[3] - rc[3]) < 2e-5, (trial, rb[3], rc[3])       # separation
            np.testing.assert_allclose(rc[:3], rb[:3], atol=2e-4)         # normal
            np.testing.assert_allclose(rc[4:7], rb[4:7], atol=2e-4)       # point on the sphere


def test_hull_resting_face_to_face_on_a_box_gives_the_patch_corners():
    # hull (the upper box) rests on a larger box: its four bottom corners lie inside the lower top face
    w, st = scene([(BOX, [0.3, 0.3, 0.1], (0, 0, 0.1), (0, 0, 0, 1), 0, 0),
                   (CONVEX, [1, 1, 1], (0.02, -0.01, 0.2 + HE[2] - 0.002), quat([0, 0, 1], 0.3), 0, 0)], HULL_BOX,
                  plane=False)
    ov, cnt, reals, feats = collide(w, st)
    k = pair_index(w, 0, 1)
    assert cnt[k] == 4
    np.testing.assert_allclose(reals[k, :4, :3], np.tile([0, 0, 1.0], (4, 1)), atol=1e-5)  # A (box) -> B (hull): +z
    np.testing.assert_allclose(reals[k, :4, 3], -0.002, atol=2e-5)
    assert sorted(int(f) & 0xFF for f in feats[k, :4, 1]) == [0, 1, 2, 3]  # the hull's bottom vertices


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_overlap_along_a_face_normal_has_the_closed_form_depth(axis):
    """Two hulls of the same box overlapping by 1 cm along one axis: MPR's
    centre-to-centre ray is that axis, so the depth is exact."""
    off = np.zeros(3)
    off[axis] = 2 * HE[axis] - 0.01
    w, st = scene([(CONVEX, [1, 1, 1], (0.3, 0.3, 0.3), (0, 0, 0, 1), 0, 0),
                   (CONVEX, [1, 1, 1], tuple(np.array([0.3, 0.3, 0.3]) + off), (0, 0, 0, 1), 0, 0)], HULL_BOX,
                  plane=False)
    ov, cnt, reals, feats = collide(w, st)
    k = pair_index(w, 0, 1)
    assert cnt[k] == 4
    n = np.zeros(3)
    n[axis] = 1.0
    np.testing.assert_allclose(reals[k, :4, :3], np.tile(n, (4, 1)), atol=1e-5)
    np.testing.assert_allclose(reals[k, :4, 3], -0.01, atol=2e-5)


def brute_distance(P, Q):
    """FP64 distance between the convex hulls of point sets P and Q (separated)."""
    from scipy.spatial import ConvexHull

    def tris(X):
        return [X[s] for s in ConvexHull(X).simplices]

    def pt_tri(p, t):
        a, b, c = t
        ab, ac, ap = b - a, c - a, p - a
        n = np.cross(ab, ac)
        nn = n @ n
        # barycentric projection; fall back to the edges
        u = np.linalg.lstsq(np.c_[ab, ac], ap, rcond=None)[0]
        if u[0] >= 0 and u[1] >= 0 and u.sum() <= 1:
            return abs(ap @ n) / np.sqrt(nn)
        return min(seg_seg(p, p, x, y) for x, y in ((a, b), (b, c), (c, a)))

    def seg_seg(p1, q1, p2, q2):
        best = np.inf
        for s in np.linspace(0, 1, 1):
            pass
        d1, d2, r = q1 - p1, q2 - p2, p1 - p2
        a, e, f = d1 @ d1, d2 @ d2, d2 @ r
        if a <= 1e-18 and e <= 1e-18:
            return np.linalg.norm(r)
        if a <= 1e-18:
            s, t = 0.0, np.clip(f / e, 0, 1)
        else:
            c = d1 @ r
            if e <= 1e-18:
                t, s = 0.0, np.clip(-c / a, 0, 1)
            else:
                b = d1 @ d2
                den = a * e - b * b
                s = np.clip((b * f - c * e) / den, 0, 1) if den > 1e-18 else 0.0
                t = (b * s + f) / e
                if t < 0:
                    t, s = 0.0, np.clip(-c / a, 0, 1)
                elif t > 1:
                    t, s = 1.0, np.clip((b - c) / a, 0, 1)
        return np.linalg.norm(p1 + d1 * s - (p2 + d2 * t))

    tp, tq = tris(P), tris(Q)
    best = min(min(pt_tri(p, t) for t in tq) for p in P)
    best = min(best, min(min(pt_tri(q, t) for t in tp) for q in Q))
    ep = {tuple(sorted((tuple(t[i]), tuple(t[(i + 1) % 3])))) for t in tp for i in range(3)}
    eq = {tuple(sorted((tuple(t[i]), tuple(t[(i + 1) % 3])))) for t in tq for i in range(3)}
    for a, b in ep:
        for c, d in eq:
            best = min(best, seg_seg(np.array(a), np.array(b), np.array(c), np.array(d)))
    return best


def test_gjk_distance_between_random_hulls_matches_brute_force():
    rng = np.random.default_rng(5)
    for trial in range(8):
        P = rng.normal(size=(12, 3)) * 0.05
        Q = rng.normal(size=(10, 3)) * 0.05
        target = rng.uniform(0.004, 0.015)  # inside the contact distance (0.02) so contacts are emitted
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        # slide Q towards P along d until the hull distance is near the target; a move of
        # delta along d shortens the distance by at most delta, so it stays above the target
        shift = (P @ d).max() - (Q @ d).min() + target
        exact = brute_distance(P, Q + shift * d)
        while exact > target + 0.003:
            shift -= exact - target
            exact = brute_distance(P, Q + shift * d)
        geoms = [(CONVEX, P.tolist(), None), (CONVEX, Q.tolist(), None)]
        w, st = scene([(CONVEX, [1, 1, 1], (0, 0, 0), (0, 0, 0, 1), 0, 0),
                       (CONVEX, [1, 1, 1], tuple(shift * d), (0, 0, 0, 1), 1, 0)], geoms, plane=False)
        ov, cnt, reals, feats = collide(w, st)
        k = pair_index(w, 0, 1)
        assert cnt[k] >= 1, trial
        # every patch point is at least the distance apart; the closest equals it
        assert abs(reals[k, :cnt[k], 3].min() - exact) < 2e-5 * max(1.0, exact) + 2e-6, (trial, reals[k, :cnt[k], 3], exact)


FLOOR = [(MESH, [[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], [[0, 1, 2], [0, 2, 3]])]


def test_box_on_a_floor_mesh_matches_the_plane_and_respects_one_sidedness():
    """The box's bottom face straddles the quad's diagonal: each triangle
    clips it, so the union of the two patches holds the face's four corners
    plus the two diagonal crossings, all at the plane's separation."""
    corners = {(0.0, 0.13), (0.0, 0.27), (0.2, 0.13), (0.2, 0.27)}
    for flags, z, sign, sep in ((0, 0.05, 1, 0.0), (0, 0.049, 1, -0.001), (0, 0.052, 1, 0.002), (0, -0.05, 0, 0.0),
                                (TWO_SIDED, -0.049, -1, -0.001)):
        w, st = scene([(BOX, HE, (0.1, 0.2, z), (0, 0, 0, 1), 0, 0), (MESH, [1, 1, 1], (0, 0, 0), (0, 0, 0, 1), 0,
                                                                        flags)], FLOOR, plane=False)
        ov, cnt, reals, feats = collide(w, st)
        k = pair_index(w, 0, 1)
        if sign == 0:
            assert cnt[k] == 0, (flags, z, cnt[k])
            continue
        c = reals[k, :cnt[k]]
        np.testing.assert_allclose(c[:, :3], np.tile([0, 0, -sign], (len(c), 1)), atol=1e-6)
        np.testing.assert_allclose(c[:, 3], sep, atol=2e-6)
        pts = {(round(float(x), 5), round(float(y), 5)) for x, y in c[:, 7:9]}
        assert corners <= pts, (flags, z, pts)
        assert pts - corners == {(0.13, 0.13), (0.2, 0.2)}, (flags, z, pts)


def bumpy_mesh(n=16, seed=2):
    rng = np.random.default_rng(seed)
    xs = np.linspace(-1, 1, n + 1)
    verts = [[x, y, 0.03 * rng.normal()] for y in xs for x in xs]
    tris = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            tris += [[a, a + 1, a + n + 2], [a, a + n + 2, a + n + 1]]
    return np.array(verts), np.array(tris)


def test_sphere_on_a_bumpy_mesh_touches_the_brute_force_triangles():
    V, T = bumpy_mesh()
    rng = np.random.default_rng(9)
    geoms = [(MESH, V.tolist(), T.tolist())]
    for trial in range(5):
        c = np.array([rng.uniform(-0.8, 0.8), rng.uniform(-0.8, 0.8), 0.0])
        # a sphere of radius 0.02 whose lowest point is 5 mm above the mesh surface below its centre
        c[2] = mesh_height(V, c[0], c[1]) + 0.02 + 0.005
        w, st = scene([(SPHERE, [0.02], tuple(c), (0, 0, 0, 1), 0, 0), (MESH, [1, 1, 1], (0, 0, 0), (0, 0, 0, 1), 0,
                                                                          0)], geoms, plane=False)
        ov, cnt, reals, feats = collide(w, st)
        k = pair_index(w, 0, 1)
        got = sorted(int(f) & 0x0FFFFFFF for f in feats[k, :cnt[k], 1])
        # brute force: triangles whose closest point is within radius + contact distance (front side)
        accepted = 0.01 + 0.01 + 1e-5
        dists = []
        for ti, t in enumerate(T):
            a, b, cc = V[t]
            nrm = np.cross(b - a, cc - a)
            d = point_triangle_distance(c, a, b, cc)
            if d - 0.02 <= accepted and (c - a) @ nrm > 0:
                dists.append((d, ti))
        expect = sorted(ti for _, ti in sorted(dists)[:8])
        assert got == expect, (trial, got, expect)
        assert cnt[k] >= 1
        exact = dict((ti, d) for d, ti in dists)
        for j in range(cnt[k]):
            ti = int(feats[k, j, 1]) & 0x0FFFFFFF
            assert abs(reals[k, j, 3] - (exact[ti] - 0.02)) < 2e-5, (trial, ti, reals[k, j, 3], exact[ti])


def mesh_height(V, x, y, n=16):
    """Height of the bumpy_mesh grid at (x, y): the containing triangle's plane."""
    h = 2.0 / n
    i, j = min(int((x + 1) / h), n - 1), min(int((y + 1) / h), n - 1)
    u, v = (x + 1) / h - i, (y + 1) / h - j
    z = lambda di, dj: V[(j + dj) * (n + 1) + i + di, 2]
    if u >= v:  # triangle (a, a+1, a+n+2)
        return z(0, 0) + u * (z(1, 0) - z(0, 0)) + v * (z(1, 1) - z(1, 0))
    return z(0, 0) + v * (z(0, 1) - z(0, 0)) + u * (z(1, 1) - z(0, 1))


def point_triangle_distance(p, a, b, c):
    """FP64 closest-point distance (Ericson 5.1.5)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return np.linalg.norm(p - a)
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return np.linalg.norm(p - b)
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return np.linalg.norm(p - (a + ab * d1 / (d1 - d3)))
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return np.linalg.norm(p - c)
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return np.linalg.norm(p - (a + ac * d2 / (d2 - d6)))
    va = d3 * d6 - d5 * d4
    if va <= 0 and d4 - d3 >= 0 and d5 - d6 >= 0:
        return np.linalg.norm(p - (b + (c - b) * (d4 - d3) / ((d4 - d3) + (d5 - d6))))
    den = 1 / (va + vb + vc)
    return np.linalg.norm(p - (a + ab * vb * den + ac * vc * den))


def run(w, st0, steps):
    st = ol.OracleState(w.model, w.config())
    st.scene[:] = st0
    for k in range(steps):
        status = st.step(np.zeros((1, 0), np.float32), control_step=k)[0]
        assert status["code"][0] == 0, (k, status)
    return st


def test_a_dropped_hull_of_a_box_follows_the_box_exactly():
    """Hull-plane contacts are the box-plane contacts (same vertices, order,
    points and separations), so the whole tumble onto the plane is the same
    trajectory bit for bit."""
    rot = quat([1, 2, 0.5], 0.7)
    ends = []
    for t, dims in ((BOX, HE), (CONVEX, [1, 1, 1])):
        w, st = scene([(t, dims, (0.0, 0.0, 0.2), rot, 0, 0)], HULL_BOX)
        st[0, 0]["angular_velocity"][:3] = (1.0, -2.0, 0.5)
        ends.append(run(w, st, 60).scene[0, 0].copy())
    for f in ("position", "orientation", "linear_velocity_inv_mass", "angular_velocity"):
        np.testing.assert_array_equal(ends[0][f], ends[1][f], err_msg=f)
    assert abs(float(ends[0]["position"][2]) - HE[2]) < 2e-3  # resting on its largest face


def test_hull_stack_settles_on_a_mesh_floor():
    """Three random hulls dropped on a floor mesh come to rest without
    sinking through it or each other."""
    rng = np.random.default_rng(11)
    geoms = []
    for k in range(3):
        P = rng.normal(size=(14, 3)) * [0.06, 0.06, 0.03]
        geoms.append((CONVEX, P.tolist(), None))
    big = [(MESH, [[-2, -2, 0], [2, -2, 0], [2, 2, 0], [-2, 2, 0]], [[0, 1, 2], [0, 2, 3]])]
    specs = [(CONVEX, [1, 1, 1], (0.0, 0.0, 0.08 + 0.09 * k), (0, 0, 0, 1), k, 0) for k in range(3)]
    specs.append((MESH, [1, 1, 1], (0, 0, 0), (0, 0, 0, 1), 3, 0))
    w, st = scene(specs, geoms + big, plane=False)
    end = run(w, st, 150).scene[0]
    for k in range(3):
        v = np.linalg.norm(end[k]["linear_velocity_inv_mass"][:3])
        assert v < 0.05, (k, v)
        lowest = min((qrot(end[k]["orientation"], p) + end[k]["position"][:3])[2] for p in np.array(geoms[k][1]))
        assert lowest > -0.003, (k, lowest)
    zs = [float(end[k]["position"][2]) for k in range(3)]
    assert zs[0] < zs[1] < zs[2] or max(zs) < 0.08, zs  # stacked, or toppled onto the floor


def qrot(q, v):
    x, y, z, w = (float(c) for c in q)
    u = np.array([x, y, z])
    return v + 2.0 * np.cross(u, np.cross(u, v) + w * v)


# ------------------------------------------------------------------ GPU parity


def gpu_matches(w, cfg, st0, steps):
    """The device world against the oracle, bit for bit: statuses, contact
    counters, scene states and manifolds after every control step."""
    ost = ol.OracleState(w.model, cfg)
    ost.scene[:] = st0
    world = mr.World(w.model, cfg)
    world.set_state(ost.q, ost.v, ost.scene)
    E = w.envs
    for k in range(steps):
        world.advance(np.zeros((1, E, 0), np.float32))
        o_status, o_cs, *_ = ost.step(np.zeros((E, 0), np.float32), control_step=k)
        g, gc = world.statuses()[0], world.contact_statuses()[0]
        np.testing.assert_array_equal(g["code"], o_status["code"], err_msg=f"step {k} codes")
        np.testing.assert_array_equal(g["failing_index"], o_status["failing_index"])
        for f in ("active_pairs", "raw_contacts", "manifolds", "constraints", "retained_points", "new_points"):
            np.testing.assert_array_equal(gc[f], o_cs[f], err_msg=f"step {k} {f}")
        _, _, scene = world.state()
        assert np.array_equal(scene.view(np.uint8), ost.scene.view(np.uint8)), f"step {k}: scene"
        mh, mp, mc = world.manifolds()
        np.testing.assert_array_equal(mc, ost.mcount, err_msg=f"step {k} manifold counts")
        for e in range(E):
            n = int(mc[e])
            assert np.array_equal(mh[e, :n].view(np.uint8), ost.mh[e, :n].view(np.uint8)), f"step {k} env {e}"
            assert np.array_equal(mp[e, :n].view(np.uint8), ost.mp[e, :n].view(np.uint8)), f"step {k} env {e}"
    return ost


def random_rotations(rng, n):
    out = []
    for _ in range(n):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        out.append(q.astype(np.float32))
    return out


DEVICE_PATH_OFF = ("the device convex / mesh narrowphase is switched off: its kernel instance took a memory "
                   "aperture violation on the first scene (hull_tumble) whose cause is not yet shown")


@pytest.mark.parametrize("case", ["few_vertices", "vertex_index", "feature_id", "kind"])
def test_geometry_compiler_refuses_malformed_geometry(case):
    geoms = [(CONVEX, box_vertices(HE), None), (MESH, [[-1, -1, 0], [1, -1, 0], [1, 1, 0]], [[0, 1, 2]])]
    specs = [(CONVEX, [1, 1, 1], (0, 0, 0.2), (0, 0, 0, 1), 0, 0), (MESH, [1, 1, 1], (0, 0, 0), (0, 0, 0, 1), 1, 0)]
    if case == "few_vertices":
        geoms[0] = (CONVEX, box_vertices(HE)[:3], None)
    elif case == "vertex_index":
        geoms[1] = (MESH, geoms[1][1], [[0, 1, 3]])
    elif case == "feature_id":
        g, v, t = mu.geometry_tables(geoms)
        t[0]["feature"] = 1 << 28
        geoms = (g, v, t)
    else:
        specs[0] = (CONVEX, [1, 1, 1], (0, 0, 0.2), (0, 0, 0, 1), 1, 0)  # a hull naming the mesh geometry
    with pytest.raises(mr.MetalRoboError):
        scene(specs, geoms, plane=False)


def test_world_creation_refuses_convex_and_mesh_shapes_while_the_device_path_is_off():
    w, _ = scene([(CONVEX, [1, 1, 1], (0.0, 0.0, 0.2), (0, 0, 0, 1), 0, 0)], HULL_BOX)
    with pytest.raises(mr.MetalRoboError, match="switched off"):
        mr.World(w.model, w.config())


@pytest.mark.gpu
@pytest.mark.skip(reason=DEVICE_PATH_OFF)
@pytest.mark.parametrize("name", ["hull_tumble", "hull_stack", "bumpy", "hybrid_ccd"])
def test_gpu_convex_and_mesh_scenes_match_oracle(gpu, name):
    print(f"convex scene {name}", flush=True)
    rng = np.random.default_rng(21)
    cfg_over = {}
    if name == "hull_tumble":
        # 32 hulls of a box at random attitudes and spins falling onto a plane
        w, st = scene([(CONVEX, [1, 1, 1], (0.0, 0.0, 0.2), (0, 0, 0, 1), 0, 0)], HULL_BOX, envs=32)
        for e, q in enumerate(random_rotations(rng, 32)):
            st[e, 0]["orientation"] = q
            st[e, 0]["angular_velocity"][:3] = rng.normal(size=3) * 2.0
        steps = 12
    elif name == "hull_stack":
        geoms = [(CONVEX, (rng.normal(size=(14, 3)) * [0.06, 0.06, 0.03]).tolist(), None) for _ in range(3)]
        geoms.append(FLOOR[0])
        specs = [(CONVEX, [1, 1, 1], (0.0, 0.0, 0.08 + 0.09 * k), (0, 0, 0, 1), k, 0) for k in range(3)]
        specs.append((MESH, [1, 1, 1], (0, 0, 0), (0, 0, 0, 1), 3, 0))
        w, st = scene(specs, geoms, plane=False, envs=16)
        for e in range(16):
            for k in range(3):
                st[e, k]["position"][:2] = rng.normal(size=2) * 0.01
        steps = 12
    elif name == "bumpy":
        V, T = bumpy_mesh()
        geoms = [(MESH, V.tolist(), T.tolist()), HULL_BOX[0]]
        specs = [(SPHERE, [0.04], (0, 0, 0.15), (0, 0, 0, 1), 0, 0), (BOX, [0.05, 0.04, 0.03], (0.2, 0, 0.15),
                                                                        (0, 0, 0, 1), 0, 0),
                 (CAPSULE, [0.02, 0.05], (-0.2, 0, 0.15), (0, 0, 0, 1), 0, 0),
                 (CONVEX, [0.8, 0.8, 0.8], (0, 0.2, 0.15), (0, 0, 0, 1), 1, 0),
                 (MESH, [1, 1, 1], (0, 0, 0), (0, 0, 0, 1), 0, 0)]
        w, st = scene(specs, geoms, plane=False, envs=32)
        rots = random_rotations(rng, 32 * 4)
        for e in range(32):
            off = rng.uniform(-0.5, 0.5, size=2)
            for k in range(4):
                st[e, k]["position"][:2] += off
                st[e, k]["orientation"] = rots[4 * e + k]
        steps = 10
    else:
        # a fast hull and sphere in the hybrid CCD mode (the CCD and geometry instance)
        shapes = [(SPHERE, [0.03], (0, 0, 0.3), (0, 0, 0, 1), 0, 1 << 1),
                  (CONVEX, [1, 1, 1], (0.3, 0, 0.3), (0, 0, 0, 1), 0, 0)]
        w, st = scene(shapes, HULL_BOX, envs=8)
        for e in range(8):
            st[e, 0]["linear_velocity_inv_mass"][:3] = (rng.uniform(-2, 2), 0.0, -40.0)
            st[e, 1]["angular_velocity"][:3] = rng.normal(size=3)
        cfg_over = dict(ccd_mode=1, max_ccd_events=4, max_ccd_iterations=16, ccd_distance_tolerance=1e-3,
                        ccd_time_tolerance=1e-5, ccd_minimum_advance=1e-6)
        steps = 4
    gpu_matches(w, w.config(**cfg_over), st, steps)
