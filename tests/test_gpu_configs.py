"""GPU tests at the sizes of the benchmark configurations: a slice of the
576-system batch (its 600-nt chain, chains across the length distribution,
eight rungs of the replica ladder with a swap attempt), and a 2000-nt
serving job, each against the FP64 oracle."""
import os
import sys

import numpy as np
import pytest

import oracle
from tis_amd import topology as T
from tis_amd.engine import Engine, MODE_GRAPH
from tis_amd.remd import ReplicaLadder

from systems import force_error

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEED = 20261202


def _c1_slice():
    sys.path.insert(0, ROOT)
    import bench
    systems, coords, labels = bench.build_workload()
    chains = [k for k, lab in enumerate(labels) if lab == "chain"]
    by_len = sorted(chains, key=lambda k: systems[k].nmp)
    pick = sorted({by_len[-1]} | {by_len[int(q)] for q in np.linspace(0, len(by_len) - 2, 29)})
    reps = [k for k, lab in enumerate(labels) if lab == "replica"][::8]
    return systems, coords, pick, reps


def _check_against_oracle(eng, i, key, energies):
    s = eng.systems[i]
    x = eng.coords(i)
    v = eng.get_state(i, 1)
    o = oracle.Oracle(s)
    o.neighbor(x)
    fo, _ = o.force(x, seed=0, sys_id=key, step=0)
    err = force_error(eng.get_state(i, 3), fo).max()
    eo = o.energy(x, v)
    eerr = (np.abs(energies[i] - eo) / np.maximum(np.abs(eo), 1e-6)).max()
    return err, eerr


def test_batched_slice_with_replica_swap():
    """The 600-nt chain, 29 more chains across the lognormal length
    distribution and 8 ladder rungs (290-400 K): 200 hipGraph steps with a
    swap attempt after 100.  Forces of a sample (all replicas at their
    current labels) within 1e-5 of the FP64 oracle, energies within 1e-7,
    and three chains bitwise equal to the same chains run alone."""
    systems, coords, pick, reps = _c1_slice()
    sel = pick + reps
    assert max(systems[k].nmp for k in pick) == 1799
    print(f"[C1 slice] {len(pick)} chains, {len(reps)} replicas, {sum(systems[k].nmp for k in sel)} beads",
          flush=True)
    eng = Engine([systems[k] for k in sel], keys=sel, seed=SEED)
    for i, k in enumerate(sel):
        eng.set_coords(i, coords[k])
    eng.langevin_init(tstep=0.05, rebuild_every=10)
    slots = list(range(len(pick), len(sel)))
    ladder = ReplicaLadder(eng, slots, systems[reps[0]], [systems[k].tempk for k in reps], seed=SEED, key=1)
    eng.langevin_steps(100, MODE_GRAPH)
    res = ladder.attempt()
    eng.langevin_steps(100, MODE_GRAPH)
    print(f"   swap attempt: {[(a, b, round(d, 2), acc) for a, b, d, acc in res]}", flush=True)
    assert ladder.attempts == 1 and len(res) == len(reps) // 2
    xb = [eng.coords(i) for i in range(len(sel))]
    assert all(np.isfinite(x).all() for x in xb)
    eng.neighbor(-1)
    eng.force_all()
    energies = eng.energy_all()
    big = sel.index(max(pick, key=lambda k: systems[k].nmp))
    sample = sorted({big, 0, len(pick) // 2, len(pick) - 2} | set(slots))
    worst = worst_e = 0.0
    for i in sample:
        err, eerr = _check_against_oracle(eng, i, sel[i], energies)
        worst, worst_e = max(worst, err), max(worst_e, eerr)
    print(f"   force err {worst:.3e}, energy err {worst_e:.3e} on {len(sample)} systems", flush=True)
    assert worst <= 1e-5
    assert worst_e <= 1e-7
    eng.close()
    for i in (big, 0, len(pick) // 2):
        k = sel[i]
        alone = Engine([systems[k]], keys=[k], seed=SEED)
        alone.set_coords(0, coords[k])
        alone.langevin_init(tstep=0.05, rebuild_every=10)
        alone.langevin_steps(200, MODE_GRAPH)
        assert np.array_equal(alone.coords(0), xb[i]), k
        alone.close()


def test_serving_size_2000nt_job():
    """The longest serving job (2000 nt, 5,999 beads, beyond the reference's
    3,500-bead limit): lists bit-exact and forces within 1e-5 of the FP64
    oracle at a compact configuration and after 50 Langevin steps from the
    helix."""
    rng = np.random.default_rng(SEED + 9)
    seq, pairs = T.hairpin_design(rng, 2000, 10, 6)
    s = T.build_rna_system([(seq, pairs)], name="job2000")
    assert s.nmp == 5999
    for label, x in (("compact", T.compact_coords(s, rng, box=110, min_dist=2.6)),
                     ("helix", T.helix_coords(s, rng))):
        print(f"[2000 nt] {label}", flush=True)
        eng = Engine([s], keys=[5], seed=SEED)
        eng.set_coords(0, x)
        if label == "helix":
            eng.langevin_init(tstep=0.05, rebuild_every=10)
            eng.langevin_steps(50, MODE_GRAPH)
            x = eng.coords(0)
        eng.neighbor(0)
        f = eng.force_sumup(0)
        o = oracle.Oracle(s)
        o.neighbor(x)
        pe = eng.pairs(0, 0)
        oe = o.lists(0)
        assert len(pe) == len(oe)
        assert np.array_equal(pe[np.lexsort(pe.T[::-1])], oe[np.lexsort(oe.T[::-1])])
        fo, _ = o.force(x, seed=0, sys_id=5, step=0)
        err = force_error(f, fo).max()
        print(f"   exv pairs {len(pe)}, force err {err:.3e}", flush=True)
        assert err <= 1e-5
        eng.close()
