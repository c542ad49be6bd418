"""Writes the golden fixtures of this directory from the FP64 oracle.

The reference implementation of this path cannot be built or imported here
(its collision core, KrisLibrary/PQP, is not in the source tree), and the
reference's own tests hold no golden vectors for it.  These fixtures are
therefore produced by the oracle, whose functions are pinned separately
against brute force, closed forms and invariants (tests/test_oracle.py,
tests/test_golden.py).  Each fixture holds its full input, so it stays valid
if the synthetic generators change.

    python tests/golden/make_fixtures.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import oracle  # noqa: E402
from klampt_amd import synthetic as S  # noqa: E402
import fixtures_io as F  # noqa: E402

BAND = 1e-6


def c1_small():
    """7-DOF arm (7 x 512-tri capsules) among 16 rotated boxes: 64 uniform
    configurations plus joint-limit edge cases."""
    sc = S.boxes_scene(0)
    ow, orb = oracle.world_from_scene(sc)
    Q = S.uniform_configs(sc.robot, 64, seed=11)
    pi_in = np.nextafter(np.float32(math.pi), np.float32(0))   # largest float32 < pi
    pi_out = np.float32(math.pi)                                 # float32(pi) > pi: outside
    special = np.zeros((5, 7), np.float32)
    special[1, :] = pi_in
    special[2, :] = -pi_in
    special[3, 3] = pi_out
    special[4, 0] = -pi_out
    Q = np.concatenate([Q, special]).astype(np.float32)
    feas, env, slf, fp = ow.feasible_batch(orb, Q)
    sep = ow.distance_batch(orb, Q, signed=True)
    dist = ow.distance_batch(orb, Q)
    fk = np.stack([ow.fk(orb, q.astype(np.float64)) for q in Q[:16]])
    out = {}
    F.pack_scene(sc, out)
    out.update(Q=Q, feasible=feas, env=env, self=slf, first_pair=fp, in_band=(np.abs(sep).min(axis=1) < BAND),
               pair_list=ow.pair_list(orb), dist=dist, sep=sep, fk=fk)
    np.savez_compressed(F.path("c1_small.npz"), **out)
    print(f"c1_small: {len(Q)} configs, feasible {int(feas.sum())}, in band {int(out['in_band'].sum())}")


def tiny_meshes():
    """Unit cube (12 triangles) against a tetrahedron (4 triangles) at 48
    poses: brute-force collide, distance, closest points and elements."""
    rng = np.random.default_rng(5)
    cube = S.box_mesh(np.array([0.5, 0.5, 0.5]))
    tet = S.Mesh(S.f32r(np.array([[0, 0, 0], [0.8, 0, 0], [0, 0.7, 0], [0, 0, 0.9]], np.float64)),
                 np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.int32))
    ow = oracle.OracleWorld()
    ga, gb = ow.add_mesh(cube), ow.add_mesh(tet)
    xa, xb, col, d, cp1, cp2, el, sep = [], [], [], [], [], [], [], []
    for k in range(48):
        Xa = S.f32r(S._xf(S.random_rotation(rng), rng.uniform(-0.2, 0.2, 3)))
        Xb = S.f32r(S._xf(S.random_rotation(rng), rng.uniform(-1.3, 1.3, 3)))
        xa.append(Xa)
        xb.append(Xb)
        col.append(ow.mesh_collide(ga, Xa, gb, Xb, brute=True))
        dd, p1, p2, e = ow.mesh_distance(ga, Xa, gb, Xb, brute=True)
        d.append(dd)
        cp1.append(p1)
        cp2.append(p2)
        el.append(e)
        sep.append(ow.mesh_separation(ga, Xa, gb, Xb))
    out = {}
    F._pack_meshes("mesh", [cube, tet], out)
    out.update(xa=np.array(xa), xb=np.array(xb), collide=np.array(col, np.uint8), dist=np.array(d),
               cp1=np.array(cp1), cp2=np.array(cp2), elems=np.array(el, np.int32), sep=np.array(sep))
    np.savez_compressed(F.path("tiny_meshes.npz"), **out)
    print(f"tiny_meshes: {len(col)} poses, colliding {int(sum(col))}")


def _grid_centres(bmin, bmax, dims):
    h = (np.asarray(bmax) - np.asarray(bmin)) / np.asarray(dims)
    ii = [(np.arange(dims[a]) + 0.5) * h[a] + bmin[a] for a in range(3)]
    return np.stack(np.meshgrid(*ii, indexing="ij"), -1)


def sdf_cases():
    """Trilinear value and gradient of a cell-centred plane grid and sphere
    grid at interior and out-of-grid points."""
    rng = np.random.default_rng(7)
    out = {}
    bmin, bmax, dims = np.array([-0.6, -0.5, -0.4]), np.array([0.6, 0.5, 0.4]), (12, 10, 8)
    n = np.array([0.3, -0.5, 0.8])
    n /= np.linalg.norm(n)
    plane = (_grid_centres(bmin, bmax, dims) @ n - 0.05).astype(np.float32)
    bmin2, bmax2, dims2 = np.array([-0.4, -0.4, -0.4]), np.array([0.4, 0.4, 0.4]), (16, 16, 16)
    X = _grid_centres(bmin2, bmax2, dims2)
    sphere = (np.linalg.norm(X - np.array([0.05, -0.02, 0.03]), axis=-1) - 0.25).astype(np.float32)
    for name, vals, lo, hi in (("plane", plane, bmin, bmax), ("sphere", sphere, bmin2, bmax2)):
        P = np.concatenate([rng.uniform(lo, hi, (200, 3)), rng.uniform(lo - 0.15, hi + 0.15, (40, 3))])
        P = P.astype(np.float32)
        d, g = oracle.sdf_eval(S.SDFGrid(lo, hi, vals.astype(np.float64)), P.astype(np.float64))
        out.update({f"{name}_bmin": lo, f"{name}_bmax": hi, f"{name}_values": vals, f"{name}_pts": P,
                    f"{name}_d": d, f"{name}_grad": g})
    out["plane_normal"] = n
    out["plane_offset"] = np.float64(0.05)
    np.savez_compressed(F.path("sdf_cases.npz"), **out)
    print("sdf_cases: plane 12x10x8, sphere 16^3, 240 points each")


def sdf_penalty_small():
    """Hinge penalty L = sum max(0, dsafe - d)^2 and dL/dq of the 7-DOF arm
    with 16 surface samples per link on a planar SDF (reproduced exactly by
    trilinear interpolation, so central differences pin the gradient)."""
    spec = S.arm7()
    pts = np.concatenate([S.surface_samples(m, 16, seed=i) for i, m in enumerate(spec.link_meshes)])
    ppl = np.full(7, 16, np.int32)
    bmin, bmax, dims = np.array([-2.4, -2.4, -2.4]), np.array([2.4, 2.4, 2.4]), (20, 20, 20)
    X = _grid_centres(bmin, bmax, dims)
    nrm = np.array([0.2, 0.1, 1.0])
    nrm /= np.linalg.norm(nrm)
    vals = (X @ nrm - 0.3).astype(np.float32)
    Q = (S.uniform_configs(spec, 12, seed=21) * 0.5).astype(np.float32)
    ow = oracle.OracleWorld()
    orb = ow.add_robot(spec)
    g64 = S.SDFGrid(bmin, bmax, vals.astype(np.float64))
    loss, grad = ow.sdf_penalty(orb, g64, Q.astype(np.float64), pts, ppl, dsafe=0.6, grad=True)
    _, gfd = ow.sdf_penalty(orb, g64, Q.astype(np.float64), pts, ppl, dsafe=0.6, grad=True, fd_h=1e-4)
    out = {}
    F.pack_scene(S.Scene(spec), out)
    out.update(pts=pts, ppl=ppl, bmin=bmin, bmax=bmax, values=vals, Q=Q, dsafe=np.float64(0.6), loss=loss,
               grad=grad, grad_fd=gfd, plane_normal=nrm, plane_offset=np.float64(0.3))
    np.savez_compressed(F.path("sdf_penalty_small.npz"), **out)
    print(f"sdf_penalty_small: {len(Q)} configs, active {int((loss > 0).sum())}")


def c1_edges():
    """24 straight segments of the 7-DOF arm among the 16 boxes, L-infinity
    and L2 resolution: first infeasible u in level-order bisection."""
    sc = S.boxes_scene(0)
    ow, orb = oracle.world_from_scene(sc)
    A, B = S.edge_segments(sc.robot, 24, eps=0.01, smin=10, smax=60, seed=9)
    out = {}
    F.pack_scene(sc, out)
    out.update(A=A, B=B)
    for tag, norm_inf, eps in (("linf", True, 0.01), ("l2", False, 0.02)):
        fu, ns, _ = ow.edge_batch(orb, A, B, eps=eps, norm_inf=norm_inf)
        # segments with any sample inside the contact band
        band = np.zeros(len(A), np.uint8)
        for s in range(len(A)):
            n = int(ns[s])
            u = np.arange(1, n) / n
            Qs = (A[s].astype(np.float64)[None] * (1 - u[:, None]) + B[s].astype(np.float64)[None] * u[:, None])
            if len(Qs):
                sep = ow.distance_batch(orb, Qs, signed=True)
                band[s] = np.abs(sep).min() < BAND
        out.update({f"{tag}_eps": np.float64(eps), f"{tag}_first_u": fu, f"{tag}_nsamples": ns, f"{tag}_band": band})
        print(f"c1_edges {tag}: {int((fu >= 0).sum())} of {len(A)} blocked, {int(band.sum())} in band")
    np.savez_compressed(F.path("c1_edges.npz"), **out)


if __name__ == "__main__":
    oracle.build()
    c1_small()
    tiny_meshes()
    sdf_cases()
    sdf_penalty_small()
    c1_edges()
