"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_fixtures.py
from the FP64 oracle): the oracle reproduces them, and their contents obey
checks that do not share the oracle's code -- scipy rotations (FK), the
joint-limit rule, brute force and surface sampling (meshes), the affine and
sphere closed forms (SDF), central differences (penalty gradient) and an
exhaustive scan of samples (edges).  CPU only.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import fixtures_io as F
import oracle
from klampt_amd import synthetic as S

BAND = 1e-6


@pytest.fixture(scope="module")
def c1():
    z = F.load("c1_small.npz")
    sc = F.unpack_scene(z)
    ow, orb = oracle.world_from_scene(sc)
    return z, sc, ow, orb


def _limits_ok(sc, Q):
    q = Q.astype(np.float64)
    return np.all((q >= sc.robot.qmin) & (q <= sc.robot.qmax), axis=1)


def _scipy_frames(spec, q):
    """Link world frames (R, t) by scipy rotations, T_i = T_p T0_i Rot(w_i q_i)."""
    out = []
    for i in range(spec.n):
        R0, t0 = spec.T0[i][:9].reshape(3, 3).T, spec.T0[i][9:]
        J = Rotation.from_rotvec(spec.axes[i] * float(q[i])).as_matrix()
        if spec.parents[i] < 0:
            out.append((R0 @ J, t0.copy()))
        else:
            Rp, tp = out[spec.parents[i]]
            out.append((Rp @ R0 @ J, Rp @ t0 + tp))
    return out


def _aabb(mesh, R, t):
    lo, hi = mesh.verts.min(0), mesh.verts.max(0)
    c, e = (lo + hi) / 2, (hi - lo) / 2
    wc, we = R @ c + t, np.abs(R) @ e
    return wc - we, wc + we


def _klampt_first_pair(sc, q, collides):
    """First colliding pair of SingleRobotCSpace::CheckCollisionFree for one
    configuration, restated here in Python from the reference
    (Cpp/Planning/PlannerSettings.cpp:241-331, RobotCSpace.cpp:794-823):
    environment set reduced against the links' AABB union, then the links
    against the environment's, each with swap-erase; links outermost; then
    the self pairs i < j over the unreduced links.  `collides` maps a world
    ID pair to its narrow-phase answer.  Rigid objects only (every link /
    object pair enabled); self pairs are the robot's default set."""
    nobj = len(sc.objects)
    rid = nobj
    links = []
    for i, (R, t) in enumerate(_scipy_frames(sc.robot, q)):
        if sc.robot.link_meshes[i] is not None:
            links.append((rid + 1 + i,) + _aabb(sc.robot.link_meshes[i], R, t))
    env = []
    for o, (m, x) in enumerate(sc.objects):
        env.append((o,) + _aabb(m, x[:9].reshape(3, 3).T, x[9:]))

    def meets(a, b):
        return bool(np.all(a[1] <= b[2]) and np.all(b[1] <= a[2]))

    def reduce(keep, against):
        u = (None, np.min([a[1] for a in against], 0), np.max([a[2] for a in against], 0)) if against else None
        i = 0
        while i < len(keep):
            if u is None or not meets(keep[i], u):
                keep[i] = keep[-1]
                keep.pop()
            else:
                i += 1

    L1, L2 = list(links), list(env)
    reduce(L2, L1)
    reduce(L1, L2)
    for a in L1:
        for b in L2:
            if meets(a, b) and collides.get((a[0], b[0]), False):
                return (a[0], b[0])
    pairs = {tuple(map(int, p)) for p in sc.robot.default_self_pairs()}
    for ia in range(len(links)):
        for ib in range(ia + 1, len(links)):
            a, b = links[ia], links[ib]
            if (a[0] - rid - 1, b[0] - rid - 1) not in pairs:
                continue
            if meets(a, b) and collides.get((a[0], b[0]), False):
                return (a[0], b[0])
    return (-1, -1)


def test_c1_oracle_reproduces_fixture(c1):
    z, sc, ow, orb = c1
    Q = z["Q"]
    feas, env, slf, fp = ow.feasible_batch(orb, Q)
    assert np.array_equal(feas, z["feasible"])
    assert np.array_equal(env, z["env"]) and np.array_equal(slf, z["self"])
    assert np.array_equal(fp, z["first_pair"])
    assert np.array_equal(ow.pair_list(orb), z["pair_list"])
    assert np.allclose(ow.distance_batch(orb, Q), z["dist"], atol=1e-12, rtol=0)
    # brute force over all triangle pairs gives the same booleans
    fb, _, _, _ = ow.feasible_batch(orb, Q[:16], brute=True)
    assert np.array_equal(fb, z["feasible"][:16])


def test_c1_fk_fixture_matches_scipy(c1):
    z, sc, _, _ = c1
    spec = sc.robot
    for Qc, T in zip(z["Q"][:16].astype(np.float64), z["fk"]):
        Rw, tw = np.eye(3), np.zeros(3)
        for i in range(spec.n):
            R0, t0 = spec.T0[i][:9].reshape(3, 3).T, spec.T0[i][9:]
            tw = Rw @ t0 + tw
            Rw = Rw @ R0 @ Rotation.from_rotvec(spec.axes[i] * Qc[i]).as_matrix()
            assert np.allclose(T[i][:9].reshape(3, 3).T, Rw, atol=1e-12)
            assert np.allclose(T[i][9:], tw, atol=1e-12)


def test_c1_fixture_is_consistent(c1):
    """feasible = limits ok and no checked pair intersects; env/self flags
    follow the intersecting pairs, the first pair is one of them (environment
    before self); distance is 0 exactly when the signed separation is <= 0."""
    z, sc, _, _ = c1
    Q, sep, dist, pl = z["Q"], z["sep"], z["dist"], z["pair_list"]
    lim = _limits_ok(sc, Q)
    assert not lim[67] and not lim[68] and lim[65] and lim[66]      # float32(pi) > pi; the next float below is in
    assert np.all(z["feasible"][~lim] == 0)
    robot_ids = set(range(17, 24))
    is_self = np.array([a in robot_ids and b in robot_ids for a, b in pl])
    nb = ~z["in_band"]
    for c in np.nonzero(lim & nb)[0]:
        hit = sep[c] < 0
        assert z["feasible"][c] == (0 if hit.any() else 1)
        assert z["env"][c] == int(hit[~is_self].any())
        assert z["self"][c] == int(hit[is_self].any())
        if hit.any():
            # the reported pair intersects, and the environment is checked
            # before self collision
            k = [i for i in range(len(pl)) if tuple(pl[i]) == tuple(z["first_pair"][c])]
            assert len(k) == 1 and hit[k[0]]
            assert is_self[k[0]] == (not hit[~is_self].any())
        else:
            assert tuple(z["first_pair"][c]) == (-1, -1)
        # and it is exactly the pair Klampt's loops reach first
        assert _klampt_first_pair(sc, Q[c], dict(zip(map(tuple, pl.tolist()), hit))) == tuple(z["first_pair"][c])
    assert np.all((dist[nb] == 0) == (sep[nb] <= 0))
    assert np.all(dist >= 0)


@pytest.fixture(scope="module")
def tiny():
    z = F.load("tiny_meshes.npz")
    cube, tet = F._unpack_meshes("mesh", z)
    return z, cube, tet


def _world_tris(mesh, x12):
    R, t = x12[:9].reshape(3, 3).T, x12[9:]
    return (mesh.verts[mesh.tris] @ R.T) + t


def _on_triangle(p, tri, tol=1e-9):
    a, b, c = tri
    n = np.cross(b - a, c - a)
    if abs(np.dot(p - a, n)) > tol * np.linalg.norm(n):
        return False
    for u, v in ((a, b), (b, c), (c, a)):
        if np.dot(np.cross(v - u, p - u), n) < -tol * np.linalg.norm(n) * np.linalg.norm(v - u):
            return False
    return True


def test_tiny_meshes_fixture(tiny):
    z, cube, tet = tiny
    ow = oracle.OracleWorld()
    ga, gb = ow.add_mesh(cube), ow.add_mesh(tet)
    rng = np.random.default_rng(0)
    for k in range(len(z["collide"])):
        xa, xb = z["xa"][k], z["xb"][k]
        # BVH traversal equals the brute force recorded in the fixture
        assert ow.mesh_collide(ga, xa, gb, xb) == bool(z["collide"][k])
        d, cp1, cp2, el = ow.mesh_distance(ga, xa, gb, xb)
        assert abs(d - z["dist"][k]) < 1e-12
        assert (z["dist"][k] == 0) == bool(z["collide"][k])
        assert (z["sep"][k] < 0) == bool(z["collide"][k])
        Ta, Tb = _world_tris(cube, xa), _world_tris(tet, xb)
        if not z["collide"][k]:
            # closest points lie on the reported triangles, at the distance
            assert abs(np.linalg.norm(z["cp1"][k] - z["cp2"][k]) - z["dist"][k]) < 1e-12
            assert _on_triangle(z["cp1"][k], Ta[z["elems"][k][0]])
            assert _on_triangle(z["cp2"][k], Tb[z["elems"][k][1]])
            # no pair of surface samples is closer
            def samples(T, n=400):
                i = rng.integers(0, len(T), n)
                u, v = rng.random(n), rng.random(n)
                m = u + v > 1
                u[m], v[m] = 1 - u[m], 1 - v[m]
                return T[i, 0] + u[:, None] * (T[i, 1] - T[i, 0]) + v[:, None] * (T[i, 2] - T[i, 0])
            Pa, Pb = samples(Ta), samples(Tb)
            dmin = np.sqrt(((Pa[:, None] - Pb[None]) ** 2).sum(-1)).min()
            assert dmin >= z["dist"][k] - 1e-12


def test_sdf_fixture_closed_forms():
    z = F.load("sdf_cases.npz")
    for name in ("plane", "sphere"):
        g = S.SDFGrid(z[f"{name}_bmin"], z[f"{name}_bmax"], z[f"{name}_values"].astype(np.float64))
        d, gr = oracle.sdf_eval(g, z[f"{name}_pts"].astype(np.float64))
        assert np.allclose(d, z[f"{name}_d"], atol=1e-12, rtol=0)
        assert np.allclose(gr, z[f"{name}_grad"], atol=1e-12, rtol=0)
    # plane: trilinear interpolation reproduces an affine field, and its
    # gradient, between the first and last cell centres
    lo, hi, vals = z["plane_bmin"], z["plane_bmax"], z["plane_values"]
    h = (hi - lo) / np.array(vals.shape)
    P = z["plane_pts"].astype(np.float64)
    inner = np.all((P >= lo + h / 2) & (P <= hi - h / 2), axis=1)
    assert inner.sum() > 100
    n, c = z["plane_normal"], float(z["plane_offset"])
    assert np.abs(z["plane_d"][inner] - (P[inner] @ n - c)).max() < 1e-6   # float32 grid values
    assert np.abs(z["plane_grad"][inner] - n).max() < 1e-5
    # sphere |x - c| - r: the Hessian norm is 1/rho, so the trilinear error
    # is at most h^2 / (4 rho_min), rho_min = rho - sqrt(3) h over the cell
    # (a half-cell index shift would err by about h/2)
    lo, hi, vals = z["sphere_bmin"], z["sphere_bmax"], z["sphere_values"]
    h = ((hi - lo) / np.array(vals.shape)).max()
    P = z["sphere_pts"].astype(np.float64)
    inner = np.all((P >= lo + h / 2) & (P <= hi - h / 2), axis=1)
    rho = np.linalg.norm(P - np.array([0.05, -0.02, 0.03]), axis=1)
    far = inner & (rho > math.sqrt(3) * h + 0.02)
    assert far.sum() > 50
    bound = h * h / (4 * (rho[far] - math.sqrt(3) * h)) + 1e-6
    assert np.all(np.abs(z["sphere_d"][far] - (rho[far] - 0.25)) <= bound)


def test_sdf_penalty_fixture_matches_central_differences():
    z = F.load("sdf_penalty_small.npz")
    sc = F.unpack_scene(z)
    g, gfd = z["grad"], z["grad_fd"]
    rel = np.abs(g - gfd) / np.abs(gfd).max(axis=1, keepdims=True)
    assert rel.max() < 1e-5
    ow = oracle.OracleWorld()
    orb = ow.add_robot(sc.robot)
    grid = S.SDFGrid(z["bmin"], z["bmax"], z["values"].astype(np.float64))
    loss, grad = ow.sdf_penalty(orb, grid, z["Q"].astype(np.float64), z["pts"], z["ppl"], dsafe=float(z["dsafe"]))
    assert np.allclose(loss, z["loss"], rtol=1e-12, atol=0)
    assert np.allclose(grad, z["grad"], rtol=1e-12, atol=1e-12)
    # loss restated in numpy from FK (scipy) and the affine field
    spec, n, c = sc.robot, z["plane_normal"], float(z["plane_offset"])
    for qi, q in enumerate(z["Q"].astype(np.float64)[:4]):
        Rw, tw, L, off = np.eye(3), np.zeros(3), 0.0, 0
        for i in range(spec.n):
            R0, t0 = spec.T0[i][:9].reshape(3, 3).T, spec.T0[i][9:]
            tw = Rw @ t0 + tw
            Rw = Rw @ R0 @ Rotation.from_rotvec(spec.axes[i] * q[i]).as_matrix()
            X = z["pts"][off:off + z["ppl"][i]] @ Rw.T + tw
            off += z["ppl"][i]
            L += (np.maximum(0, float(z["dsafe"]) - (X @ n - c)) ** 2).sum()
        assert abs(L - z["loss"][qi]) < 1e-5 * max(1.0, L)


def test_c1_edges_fixture_equals_exhaustive_scan():
    z = F.load("c1_edges.npz")
    sc = F.unpack_scene(z)
    ow, orb = oracle.world_from_scene(sc)
    A, B = z["A"].astype(np.float64), z["B"].astype(np.float64)
    for tag, norm_inf in (("linf", True), ("l2", False)):
        eps = float(z[f"{tag}_eps"])
        fu, ns, _ = ow.edge_batch(orb, A, B, eps=eps, norm_inf=norm_inf)
        assert np.array_equal(fu, z[f"{tag}_first_u"]) and np.array_equal(ns, z[f"{tag}_nsamples"])
        for s in range(len(A)):
            d = np.abs(B[s] - A[s]).max() if norm_inf else np.linalg.norm(B[s] - A[s])
            n = int(ns[s])
            assert n == math.ceil(d / eps)
            order = oracle.bisection_order(n)
            u = order / n
            Qs = A[s][None] * (1 - u[:, None]) + B[s][None] * u[:, None]
            feas = ow.feasible_batch(orb, Qs)[0] if len(Qs) else np.zeros(0, np.uint8)
            bad = np.nonzero(feas == 0)[0]
            want = -1.0 if len(bad) == 0 else u[bad[0]]
            assert fu[s] == want
