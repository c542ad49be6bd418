"""The HIP path, through the C ABI, on the golden fixtures of tests/golden.

Bars: booleans and first pairs bit-exact outside the 1e-6 m contact band;
FP32 distances within 1e-5 m; FK poses within 1e-5 (rotation entries and
metres); SDF values within 1e-5 m and gradients within 1e-4; penalty loss
within 1e-4 relative and gradients within 1e-4 of the largest component per
configuration; edge first-u bit-exact (no fixture segment has a sample in the
contact band).
"""
import numpy as np
import pytest

import fixtures_io as F
from klampt_amd import engine

pytestmark = pytest.mark.gpu


def _world(ctx, sc):
    return engine.world_from_scene(ctx, sc)


def test_c1_golden_feasibility_distance_fk(ctx):
    z = F.load("c1_small.npz")
    sc = F.unpack_scene(z)
    w, r = _world(ctx, sc)
    Q = z["Q"]
    feas, env, slf, fp = w.feasible_batch(Q, flags=True, first_pair=True)
    nb = ~z["in_band"].astype(bool)
    print(f"c1 golden: {len(Q)} configs, {int((~nb).sum())} in band")
    for got, want, what in ((feas, z["feasible"], "feasible"), (env, z["env"], "env"), (slf, z["self"], "self")):
        assert np.array_equal(got[nb], want[nb]), what
    assert np.array_equal(fp[nb], z["first_pair"][nb])
    assert np.array_equal(w.pair_list(0), z["pair_list"])
    d = w.distance_batch(Q)
    assert np.abs(d - z["dist"]).max() < 1e-5
    T = r.fk_batch(Q[:16])
    assert np.abs(T - z["fk"]).max() < 1e-5


def test_tiny_meshes_golden(ctx):
    z = F.load("tiny_meshes.npz")
    cube, tet = F._unpack_meshes("mesh", z)
    ga = engine.Geom.trimesh(ctx, cube.verts, cube.tris)
    gb = engine.Geom.trimesh(ctx, tet.verts, tet.tris)
    for k in range(len(z["collide"])):
        xa, xb = z["xa"][k], z["xb"][k]
        ga.set_transform(xa[:9], xa[9:])
        gb.set_transform(xb[:9], xb[9:])
        if abs(z["sep"][k]) >= 1e-6:
            assert ga.collides(gb) == bool(z["collide"][k]), k
        r = ga.distance(gb)
        assert abs(r.d - z["dist"][k]) < 1e-5, k
        if r.d > 0:
            assert abs(np.linalg.norm(np.array(r.cp1) - np.array(r.cp2)) - r.d) < 1e-5


def test_sdf_golden(ctx):
    z = F.load("sdf_cases.npz")
    for name in ("plane", "sphere"):
        g = engine.Geom.sdf(ctx, z[f"{name}_bmin"], z[f"{name}_bmax"], z[f"{name}_values"])
        d, gr = g.sdf_eval(z[f"{name}_pts"])
        assert np.abs(d - z[f"{name}_d"]).max() < 1e-5, name
        assert np.abs(gr - z[f"{name}_grad"]).max() < 1e-4, name


def test_sdf_penalty_golden(ctx):
    z = F.load("sdf_penalty_small.npz")
    sc = F.unpack_scene(z)
    r = engine.Robot.from_spec(ctx, sc.robot)
    sdf = engine.Geom.sdf(ctx, z["bmin"], z["bmax"], z["values"])
    dsafe = float(z["dsafe"])
    loss = r.sdf_penalty(sdf, z["Q"], z["pts"], z["ppl"], dsafe=dsafe)
    dQ = r.sdf_penalty_grad(sdf, z["Q"], z["pts"], z["ppl"], dsafe=dsafe)
    assert np.all(np.abs(loss - z["loss"]) <= 1e-4 * np.maximum(1.0, z["loss"]))
    scale = np.abs(z["grad"]).max(axis=1, keepdims=True)
    assert (np.abs(dQ - z["grad"]) / scale).max() < 1e-4


def test_c1_edges_golden(ctx):
    z = F.load("c1_edges.npz")
    sc = F.unpack_scene(z)
    w, _ = _world(ctx, sc)
    for tag, norm_inf in (("linf", True), ("l2", False)):
        assert not z[f"{tag}_band"].any()
        u = w.edge_batch(z["A"], z["B"], float(z[f"{tag}_eps"]), norm_inf)
        want = z[f"{tag}_first_u"].astype(np.float32)
        assert np.array_equal(u, want), tag
