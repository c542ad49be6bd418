"""The engine's OBB trees (kg_geom_get_bvh) and the oracle's counted
IsFeasible on them: every triangle in exactly one leaf, every node's box
holding its triangles, and on the 200k-triangle soup the counted sequential
check gives the GPU's answers (these counts define the collision path's
algorithmic bytes)."""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S

pytestmark = pytest.mark.gpu


def _check_tree(tree, mesh):
    nn = tree.shape[0]
    child = tree[:, 15].view(np.int32)
    leaves = child < 0
    assert nn == 2 * mesh.tris.shape[0] - 1
    tri = ~child[leaves]
    assert np.array_equal(np.sort(tri), np.arange(mesh.tris.shape[0]))
    inner = np.nonzero(~leaves)[0]
    kids = np.concatenate([child[inner], child[inner] + 1])
    assert np.array_equal(np.sort(kids), np.arange(1, nn))
    # leaf boxes contain their triangle's vertices (FP32 boxes, exact bounds)
    V = mesh.verts[mesh.tris[tri]]                     # (L, 3, 3)
    c = tree[leaves, 0:3].astype(np.float64)
    e = tree[leaves, 3:6].astype(np.float64)
    R = tree[leaves, 6:15].astype(np.float64).reshape(-1, 3, 3)  # R[k] = axis k (column-major)
    loc = np.einsum("lkj,lvj->lvk", R, V - c[:, None, :])
    assert np.all(np.abs(loc) <= e[:, None, :] * (1 + 1e-5) + 1e-6)


def test_soup_trees_and_counted_oracle_equal_gpu(ctx):
    sc = S.soup_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    geoms = list(w._keep) + [g for g in r.link_geoms if g is not None]
    meshes = list(sc.terrains) + [m for m in sc.robot.link_meshes if m is not None]
    assert len(geoms) == len(ow.meshes) == len(meshes)
    for g, og, m in zip(geoms, ow.meshes, meshes):
        t = g.bvh()
        _check_tree(t, m)
        ow.set_obb_tree(og, t)
    Q = S.uniform_configs(sc.robot, 4096, seed=31)
    feas = w.feasible_batch(Q)[0]
    fc, cnt = ow.feasible_counted(orb, Q.astype(np.float64))
    assert np.array_equal(fc, feas)
    tot = cnt.sum(axis=0)
    print("counted on the engine trees (env box, env nodes, env tri, self box, self tri):", tot.tolist())
    assert tot[1] <= tot[0] and tot[2] > 0
