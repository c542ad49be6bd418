"""KineTrajOpt sweep terms on the GPU (kg_sweep_batch[_dev]) against the FP64
oracle's restatement of find_collision_pair / add_sweep
(oracle.sweep_terms): the arm's link hulls swept over trajectory steps
against the 16 boxes of the C1 scene.

Bar: distances within 1e-9 m (both FP64 GJK, stopped at a relative gap of
1e-13); the same queries below dcheck and overlapping; the normal (unique
for separated convex sets) within 1e-7; the support points, alpha and the
constraint rows within 1e-6 wherever the support point is unique; where a
vertex projection ties within 1e-9 m either vertex may be chosen, and the
GPU's choice must attain the support maximum, with alpha and the rows (the
oracle's finite-difference Jacobian) following from it."""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S

pytestmark = pytest.mark.gpu

DCHECK = 0.25


@pytest.fixture(scope="module")
def scene(ctx):
    sc = S.boxes_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    return sc, w, ow, orb


def _segments(spec, N, seed):
    rng = np.random.default_rng(seed)
    Q0 = S.uniform_configs(spec, N, seed=seed).astype(np.float64)
    step = rng.normal(size=Q0.shape)
    step *= rng.uniform(0.02, 0.3, (N, 1)) / np.abs(step).max(axis=1, keepdims=True)
    return Q0, Q0 + step


def _support_tie(P, X, d, tol=1e-9):
    """True when the largest projection of the placed points on d is reached
    by more than one point within tol."""
    R, t = X[:9].reshape(3, 3).T, X[9:]
    v = (P @ R.T + t) @ d
    s = np.sort(v)
    return s[-1] - s[-2] < tol


def test_sweep_terms_vs_oracle(ctx, scene):
    sc, w, ow, orb = scene
    spec = sc.robot
    n, O = spec.n, len(sc.objects)
    assert w.num_obstacles() == O
    Q0, Q1 = _segments(spec, 96, seed=3)
    print("sweep terms: 96 segments x 7 links x 16 boxes", flush=True)
    d, info, coef = w.sweep_batch(Q0, Q1, DCHECK)
    nterms = ntie = 0
    for k in range(len(Q0)):
        for l in range(n):
            P = spec.link_meshes[l].verts
            for o in range(O):
                B, xb = sc.objects[o][0].verts, sc.objects[o][1]
                want = oracle.sweep_terms(ow, orb, P, l, Q0[k], Q1[k], B, xb, DCHECK)
                g = d[k, l, o]
                if want is None:
                    # at the bound both sides may round either way
                    if np.isfinite(g):
                        dd = oracle.hull_distance(P, ow.fk(orb, Q0[k])[l], B, xb, xa2=ow.fk(orb, Q1[k])[l])[0]
                        assert abs(dd - DCHECK) < 1e-9, (k, l, o, g)
                    continue
                assert abs(g - want[0]) < 1e-9, (k, l, o, g, want[0])
                if len(want) == 1:
                    # overlapping hulls: no terms (a GJK stopping a hair
                    # short of the origin reports a distance within 1e-9)
                    assert g < 1e-9
                    if g == 0:
                        assert not coef[k, l, o].any()
                    continue
                nterms += 1
                _, alpha, nrm, ps, po, p0, p1, c0, c1 = want
                gi = info[k, l, o]
                assert np.allclose(gi[1:4], nrm, atol=1e-7), (k, l, o)
                assert abs(np.linalg.norm(gi[4:7] - gi[7:10]) - g) < 1e-9
                X0, X1 = ow.fk(orb, Q0[k])[l], ow.fk(orb, Q1[k])[l]
                if _support_tie(P, X0, -nrm) or _support_tie(P, X1, -nrm):
                    # either tied vertex may be chosen: the GPU's choices
                    # attain the support maximum, and alpha and the rows
                    # follow from them
                    ntie += 1
                    gp0, gp1 = gi[10:13], gi[13:16]
                    for X, gp in ((X0, gp0), (X1, gp1)):
                        R, t = X[:9].reshape(3, 3).T, X[9:]
                        assert ((P @ R.T + t) @ -nrm).max() <= gp @ -nrm + 1e-9
                    a0, a1 = np.linalg.norm(gp0 - gi[4:7]), np.linalg.norm(gp1 - gi[4:7])
                    assert abs(gi[0] - (a1 / (a0 + a1) if a0 + a1 > 0 else 0.5)) < 1e-9
                    assert np.allclose(coef[k, l, o, 0], oracle.normal_row(ow, orb, l, Q0[k], gp0, nrm), atol=1e-6)
                    assert np.allclose(coef[k, l, o, 1], oracle.normal_row(ow, orb, l, Q1[k], gp1, nrm), atol=1e-6)
                    continue
                assert np.allclose(gi[10:13], p0, atol=1e-6) and np.allclose(gi[13:16], p1, atol=1e-6), (k, l, o)
                assert np.allclose(gi[4:7], ps, atol=1e-6) and np.allclose(gi[7:10], po, atol=1e-6), (k, l, o)
                assert abs(gi[0] - alpha) < 1e-6, (k, l, o, gi[0], alpha)
                assert np.allclose(coef[k, l, o, 0], c0, atol=1e-6), (k, l, o)
                assert np.allclose(coef[k, l, o, 1], c1, atol=1e-6), (k, l, o)
    print(f"sweep terms: {nterms} below dcheck, {ntie} with a support tie")
    assert nterms > 20


def test_point_terms_and_dev_path(ctx, scene):
    """Q0 == Q1: the hull of one placement, alpha 1/2 and equal rows; the
    device-pointer call returns the host call's values; a batch whose
    queries do not fill whole waves (37 segments per link and box) returns
    the same values as the same segments inside a larger batch."""
    import torch
    sc, w, ow, orb = scene
    Q0, Q1 = _segments(sc.robot, 32, seed=4)
    d, info, coef = w.sweep_batch(Q0, Q0, DCHECK)
    hit = np.isfinite(d) & (d > 0)
    assert hit.sum() > 5
    assert np.allclose(info[..., 0][hit], 0.5) and np.allclose(coef[..., 0, :][hit], coef[..., 1, :][hit])
    dd, di, dc = w.sweep_batch_dev(torch.from_numpy(Q0).cuda(), torch.from_numpy(Q1).cuda(), DCHECK)
    h = w.sweep_batch(Q0, Q1, DCHECK)
    assert np.array_equal(dd.cpu().numpy(), h[0])
    assert np.array_equal(di.cpu().numpy(), h[1]) and np.array_equal(dc.cpu().numpy(), h[2])
    r = w.sweep_batch(Q0[:37], Q1[:37], DCHECK)
    for x, y in zip(r, h):
        assert np.array_equal(x, y[:37])
