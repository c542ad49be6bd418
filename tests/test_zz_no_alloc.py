"""kg_feasible_batch_dev_noalloc (every workspace in the library's fixed
scratch array) gives exactly kg_feasible_batch_dev's outputs: two calls in a
row, the second with a smaller batch, in any-hit mode and with the flags and
first pairs; a batch beyond the scratch is rejected before any launch."""
import numpy as np
import pytest

from klampt_amd import _lib, engine, synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def soup(ctx):
    sc = S.soup_scene(0)
    w, r = engine.world_from_scene(ctx, sc)
    return sc, w


def _outputs(n, flags):
    import torch
    f = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
    if not flags:
        return f, None, None, None
    e = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
    s = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
    p = torch.full((n, 2), 7, dtype=torch.int32, device="cuda")
    return f, e, s, p


@pytest.mark.parametrize("flags", [False, True])
def test_noalloc_equals_dev_two_calls(ctx, soup, flags):
    import torch
    sc, w = soup
    for N, seed in ((16384, 41), (3000, 42)):
        Q = torch.from_numpy(S.uniform_configs(sc.robot, N, seed=seed)).cuda()
        ref = _outputs(N, flags)
        w.feasible_batch_dev(Q, feasible=ref[0], env=ref[1], self_hit=ref[2], first_pair=ref[3])
        got = _outputs(N, flags)
        print(f"noalloc N={N} flags={flags}", flush=True)
        w.feasible_batch_dev_noalloc(Q, feasible=got[0], env=got[1], self_hit=got[2], first_pair=got[3])
        status = ctx.status_dev().cpu().numpy()
        assert status[1] == 0
        for a, b in zip(ref, got):
            if a is not None:
                assert torch.equal(a, b)
        print(f"  feasible fraction {ref[0].float().mean().item():.3f}, equal")


def test_noalloc_rejects_batch_beyond_scratch(ctx, soup):
    import torch
    sc, w = soup
    Q = torch.zeros((200000, sc.robot.n), dtype=torch.float32, device="cuda")
    f = torch.empty(200000, dtype=torch.uint8, device="cuda")
    with pytest.raises(_lib.KlamptGPUError) as ei:
        w.feasible_batch_dev_noalloc(Q, feasible=f)
    assert ei.value.code == -5 and "scratch" in str(ei.value)
