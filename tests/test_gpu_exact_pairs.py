"""Triangle pairs the FP32 test cannot decide are decided one by one by the
FP64 pass's exact test (any-hit and group modes), instead of the whole
configuration being traversed again in FP64.  On configurations placed on
the contact boundary (FP64 bisection with the oracle between a feasible and
an infeasible configuration, to 1e-14 of the segment), the answers must be
bit-identical to sending every configuration through the exact FP64 pass
(force_recheck), in all three result modes, and equal to the oracle outside
the 1e-6 m contact band, as everywhere else."""
import numpy as np
import pytest

import oracle
from klampt_amd import engine, synthetic as S
from tests.conftest import engine_params
from tests.test_gpu_feasible import _check_exact

pytestmark = pytest.mark.gpu


def _contact_configs(ow, orb, spec, n, seed):
    """n configurations on each side of the contact boundary."""
    Q = S.uniform_configs(spec, 8 * n, seed=seed).astype(np.float64)
    f = ow.feasible_batch(orb, Q)[0].astype(bool)
    a, b = Q[f][:n], Q[~f][:n]
    m = min(len(a), len(b))
    a, b = a[:m], b[:m]
    lo, hi = np.zeros(m), np.ones(m)
    for _ in range(48):
        u = 0.5 * (lo + hi)
        fu = ow.feasible_batch(orb, a + u[:, None] * (b - a))[0].astype(bool)
        lo = np.where(fu, u, lo)
        hi = np.where(fu, hi, u)
    return np.concatenate([a + lo[:, None] * (b - a), a + hi[:, None] * (b - a)])


@pytest.mark.parametrize("scene", ["boxes", "soup"])
def test_contact_configurations_pair_path_equals_full_fp64_pass(ctx, scene):
    sc = S.boxes_scene(0) if scene == "boxes" else S.soup_scene(0, target_tris=60_000)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    Q = _contact_configs(ow, orb, sc.robot, 64, seed=71)
    print(f"{scene}: {len(Q)} configurations on the contact boundary", flush=True)
    f_any = w.feasible_batch(Q)[0]
    st = w.last_stats()
    print(f"any-hit mode: {st['exact_pairs']} pairs decided exactly, {st['fp64_rechecks']} configurations "
          f"traversed again", flush=True)
    assert st["exact_pairs"] > 0, "no pair was left undecided: the test does not reach the pair path"
    assert st["fp64_rechecks"] == 0 and st["uncertain"] == 0
    grp = w.feasible_batch(Q, flags=True)
    assert w.last_stats()["fp64_rechecks"] == 0
    ordd = w.feasible_batch(Q, flags=True, first_pair=True)
    assert w.last_stats()["fp64_rechecks"] > 0, "the ordered mode traverses such configurations again"
    with engine_params(ctx, force_recheck=1):
        x_any = w.feasible_batch(Q)[0]
        x_grp = w.feasible_batch(Q, flags=True)
        x_ord = w.feasible_batch(Q, flags=True, first_pair=True)
        assert w.last_stats()["fp64_rechecks"] == len(Q)
    assert np.array_equal(f_any, x_any)
    for g, x in zip(grp[:3], x_grp[:3]):
        assert np.array_equal(g, x)
    for g, x in zip(ordd, x_ord):
        assert np.array_equal(g, x)
    f2, e2, s2, fp2 = ow.feasible_batch(orb, Q)
    nb = _check_exact(ow, orb, Q, np.c_[grp[0], grp[1], grp[2]], np.c_[f2, e2, s2], scene + " groups")
    print(f"{scene}: feasible {f_any.mean():.2f}; in-band differences from the oracle {nb}")
    assert 0 < f_any.mean() < 1


def test_device_batch_and_pair_overflow(ctx):
    """The device-pointer entry point takes the same path; with more
    undecided pairs than the list holds, the rest go through the full FP64
    pass and the answers do not change."""
    import torch
    sc = S.boxes_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    ow, orb = oracle.world_from_scene(sc)
    Q = _contact_configs(ow, orb, sc.robot, 64, seed=72)
    ref = w.feasible_batch(Q)[0]
    dQ = torch.from_numpy(Q).cuda()
    got = w.feasible_batch_dev(dQ).cpu().numpy()
    assert np.array_equal(got, ref)
    # many copies of the boundary configurations: more pairs than the list
    big = np.tile(Q, (40, 1))
    fb = w.feasible_batch(big)[0]
    st = w.last_stats()
    print(f"{len(big)} configurations: {st['exact_pairs']} undecided pairs, {st['fp64_rechecks']} "
          f"configurations traversed again", flush=True)
    assert st["exact_pairs"] > 1024 and st["fp64_rechecks"] > 0
    assert np.array_equal(fb, np.tile(ref, 40))
