"""GPU parity of straight-line edge checks against the FP64 oracle.

Bar: first_u (u of the first infeasible sample in level-order bisection, -1
if free) bit-exact (as float32 of the FP64 u) outside the contact band.
"""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("norm_inf", [True, False])
def test_c1_edges_first_u(ctx, norm_inf):
    sc = S.boxes_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    A, B = S.edge_segments(sc.robot, 512, eps=0.01, smin=10, smax=200, seed=6)
    fu = w.edge_batch(A, B, eps=0.01, norm_inf=norm_inf)
    fo, ns, nchk = ow.edge_batch(orb, A, B, eps=0.01, norm_inf=norm_inf)
    bad = np.nonzero(fu != fo.astype(np.float32))[0]
    print(f"C1 edges (Linf={norm_inf}): {len(A)} segs, free {(fo < 0).mean():.2f}, mismatches {len(bad)}")
    for s in bad:  # only contact-band samples may differ
        n = ns[s]
        order = oracle.bisection_order(n)
        a, b = A[s].astype(np.float64), B[s].astype(np.float64)
        Q = np.array([a * (1 - i / n) + b * (i / n) for i in order])
        sep = ow.distance_batch(orb, Q, signed=True)
        assert (np.abs(sep).min(axis=1) < 1e-6).any(), f"segment {s} differs outside the contact band"
    assert len(bad) <= 2


def test_c2_soup_edges_first_u(ctx):
    sc = S.soup_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    A, B = S.edge_segments(sc.robot, 256, eps=0.01, smin=10, smax=500, seed=7)
    fu = w.edge_batch(A, B, eps=0.01, norm_inf=True)
    fo, ns, nchk = ow.edge_batch(orb, A, B, eps=0.01, norm_inf=True)
    assert np.array_equal(fu, fo.astype(np.float32))


def test_edge_edge_cases(ctx):
    sc = S.boxes_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    q = S.uniform_configs(sc.robot, 1, seed=1)
    # zero-length and single-step segments have no interior samples
    assert w.edge_batch(q, q, eps=0.01)[0] == -1
    b = q.copy(); b[0, 0] += 0.005
    assert w.edge_batch(q, b, eps=0.01)[0] == -1
    with pytest.raises(Exception):
        w.edge_batch(q, b, eps=0.0)
