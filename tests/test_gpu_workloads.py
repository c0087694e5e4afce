"""The batched-replica workload (256 channels of 2k-40k beads) and the
learned warm start, exercised on the GPU at bounded sizes."""
import math

import numpy as np
import pytest

from mfa_brush_amd import configs
from mfa_brush_amd.sim import Simulation

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4():
    return configs.c4_batch(256)


def test_c4_batch_pair_counts_momentum_and_single_runs(c4, oracle_mod):
    """256 replicas in one context, mixed precision, 20 graph-replayed steps:
    every replica's interacting-pair count equals the oracle's linked-cell
    count at its final positions, every replica's pair forces sum to ~0,
    two replicas' forces match the oracle (with the same DPD noise) within
    the mixed bar, and the smallest, median and largest replicas give the
    same bits batched and alone."""
    ps = [r[0].replace(skin=0.4) for r in c4]
    sizes = np.array([p.n_part for p in ps])
    print("C4: %d replicas, %d beads, %d .. %d per replica" % (len(ps), sizes.sum(), sizes.min(), sizes.max()))
    n = 20
    with Simulation(ps, precision="mixed") as s:
        for i, (_, t, r0, v) in enumerate(c4):
            s.set_state(r0, v, t, isys=i)
        s.md_steps(n, graph=True)
        fin = [s.get_state(i) for i in range(len(ps))]
        counts = [s.pair_count(i) for i in range(len(ps))]
        fl = s.flags()
        s.fluid_fluid()
        pair_f = [s.get_state(i)["force"] for i in range(len(ps))]
    assert fl["nonfinite"] == 0 and fl["list_overflow"] == 0
    for i, p in enumerate(ps):
        co, _ = oracle_mod.Oracle(p).pair_stats_cells(fin[i]["r0"], c4[i][1])
        assert counts[i] == co.sum() // 2, i
        f = pair_f[i]
        assert np.abs(f.sum(0)).max() < 1e-11 * np.abs(f).max() * math.sqrt(p.n_part), i
    order = np.argsort(sizes)
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    for i in picks[::2]:
        fo, _ = oracle_mod.Oracle(ps[i]).fluid_fluid(fin[i]["r0"], fin[i]["v"], c4[i][1], step=n, noise=1)
        err = np.linalg.norm(pair_f[i] - fo, axis=1).max() / np.linalg.norm(fo, axis=1).max()
        print("C4 replica %d (%d beads): mixed force error %.2e" % (i, ps[i].n_part, err))
        assert err < 1e-5, (i, err)
    for i in picks:
        _, t, r0, v = c4[i]
        with Simulation(ps[i], precision="mixed") as s1:
            s1.set_state(r0, v, t)
            s1.md_steps(n, graph=True)
            one = s1.get_state()
        for k in ("r0", "v", "a", "force"):
            assert np.array_equal(one[k], fin[i][k]), (i, k)


def test_warm_start_beats_plain_start():
    """The learned part: a surrogate trained on the batched simulator's
    profiles (64 small channels) seeds a held-out channel whose first
    block of MD is at least 2x closer (L1) to the equilibrium profiles than
    the first block of the plain straight-chain start and of the reference's
    generator start (ordered gen_brush brush, gen_droplet solvent)."""
    from mfa_brush_amd.warmstart import bounded_tte
    r = bounded_tte(n_blocks=30)
    print(r)
    assert r["nonfinite"] == 0
    assert r["surrogate_l1"] < 0.3, r["surrogate_l1"]
    for other in ("plain", "generated"):
        assert r["first_block_l1_warm"] * 2.0 < r["first_block_l1_" + other], (r["first_block_l1_warm"], other,
                                                                               r["first_block_l1_" + other])
