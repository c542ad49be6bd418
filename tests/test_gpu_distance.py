"""GPU parity of TriangleMesh distances against the FP64 oracle.

Bar: FP32 distances within 1e-5 m of FP64; 0 exactly when the oracle reports
an intersection (up to the 1e-6 m contact band); closest points consistent
with the reported distance; withinDistance equal to (distance <= tol)
outside the band.
"""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S

pytestmark = pytest.mark.gpu
TOL = 1e-5


def test_c1_pair_distances(ctx):
    sc = S.boxes_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    assert np.array_equal(w.pair_list(0), ow.pair_list(orb))
    Q = S.uniform_configs(sc.robot, 256, seed=4)
    d = w.distance_batch(Q)
    d64 = ow.distance_batch(orb, Q)
    err = np.abs(d - d64)
    print(f"C1 distances: {d.size} pairs, max |err| {err.max():.2e} m, zeros {int((d64 == 0).sum())}")
    assert err.max() < TOL


def test_c2_soup_pair_distances(ctx):
    sc = S.soup_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    Q = S.uniform_configs(sc.robot, 24, seed=5)
    d = w.distance_batch(Q)
    d64 = ow.distance_batch(orb, Q)
    err = np.abs(d - d64)
    print(f"C2 distances: {d.shape}, max |err| {err.max():.2e} m")
    assert err.max() < TOL


def test_pair_distance_closest_points_and_within(ctx):
    rng = np.random.default_rng(12)
    cap = S.capsule_mesh(0.05, 0.3, axis=2)
    box = S.box_mesh([0.1, 0.05, 0.08], edge=0.02)
    ga, gb = engine.Geom.trimesh(ctx, cap.verts, cap.tris), engine.Geom.trimesh(ctx, box.verts, box.tris)
    ow = oracle.OracleWorld()
    oa, ob = ow.add_mesh(cap), ow.add_mesh(box)
    for _ in range(48):
        Ra, ta = S.random_rotation(rng), rng.uniform(-0.3, 0.3, 3)
        Rb, tb = S.random_rotation(rng), rng.uniform(-0.3, 0.3, 3)
        xa, xb = np.r_[Ra.T.ravel(), ta], np.r_[Rb.T.ravel(), tb]
        ga.set_transform(xa[:9], xa[9:])
        gb.set_transform(xb[:9], xb[9:])
        res = ga.distance(gb)
        d64 = ow.mesh_distance(oa, xa, ob, xb)[0]
        assert abs(res.d - d64) < TOL
        if d64 > 1e-6:
            assert res.hasClosestPoints and res.hasGradients
            cp1, cp2 = np.array(res.cp1), np.array(res.cp2)
            assert abs(np.linalg.norm(cp1 - cp2) - res.d) < 1e-9
            assert np.allclose(np.array(res.grad1), -np.array(res.grad2))
            # the reported elements realise the distance
            t1, t2 = cap.tris[res.elem1], box.tris[res.elem2]
            Pa = cap.verts[t1] @ Ra.T + ta
            Pb = box.verts[t2] @ Rb.T + tb
            assert abs(oracle.tritri_dist(Pa.ravel(), Pb.ravel()) - d64) < 1e-9
        for tol in (0.01, 0.05):
            if abs(d64 - tol) > 1e-6:
                assert ga.within_distance(gb, tol) == (d64 <= tol)
