"""GPU parity: HIP path (through the C-ABI) against the FP64 oracle."""
import numpy as np
import pytest

import oracle
from tis_amd.engine import Engine, LIST_ELE, LIST_EXV, LIST_HB, MODE_EAGER, MODE_GRAPH, MODE_PLAIN

from tis_amd import topology as T

from systems import (compact_allpairs, force_error, hairpin_design, hairpin64, parity_cases, unit_cases,
                     valence_crowd)

pytestmark = pytest.mark.gpu
CASES = parity_cases()


def _sorted(p):
    p = np.asarray(p)
    if p.ndim == 1:
        return np.sort(p)
    return p[np.lexsort(p.T[::-1])]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_lists_bit_exact(case):
    name, s, x = case
    print(f"[lists] {name} N={s.nmp}", flush=True)
    eng = Engine([s])
    eng.set_coords(0, x)
    eng.neighbor(0)
    o = oracle.Oracle(s)
    o.neighbor(x)
    ge = eng.pairs(0, LIST_EXV)
    oe = o.lists(0)
    assert np.array_equal(_sorted(ge), _sorted(oe)), (len(ge), len(oe))
    gq, gc = eng.pairs(0, LIST_ELE)
    oq, oc = o.lists(1)
    assert np.array_equal(_sorted(gq), _sorted(oq))
    assert np.array_equal(gc[np.lexsort(gq.T[::-1])], oc[np.lexsort(oq.T[::-1])])
    assert np.array_equal(np.sort(eng.pairs(0, LIST_HB)), np.sort(o.lists(2)))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_force_sumup(case):
    name, s, x = case
    print(f"[force] {name} N={s.nmp}", flush=True)
    eng = Engine([s], keys=[0], seed=0)
    eng.set_coords(0, x)
    eng.neighbor(0)
    f = eng.force_sumup(0)
    o = oracle.Oracle(s)
    o.neighbor(x)
    fo, _ = o.force(x, seed=0, sys_id=0, step=0)
    err = force_error(f, fo)
    print(f"   max rel err {err.max():.3e} p99 {np.percentile(err, 99):.3e}", flush=True)
    assert err.max() <= 1e-5


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_energy_sumup(case):
    name, s, x = case
    print(f"[energy] {name}", flush=True)
    rng = np.random.default_rng(5)
    v = rng.normal(size=x.shape)
    eng = Engine([s], keys=[0], seed=0)
    eng.set_coords(0, x)
    eng.set_state(0, 1, v)
    eng.neighbor(0)
    eng.force_sumup(0)
    e = eng.energy_sumup(0)
    o = oracle.Oracle(s)
    o.neighbor(x)
    o.force(x, seed=0, sys_id=0, step=0)
    eo = o.energy(x, v)
    rel = np.abs(e - eo) / np.maximum(np.abs(eo), 1e-6)
    print("   energy rel", rel.max(), flush=True)
    assert rel.max() <= 1e-7


def test_batch_equals_alone_bitwise():
    """A system gives the same bits alone and inside a mixed batch."""
    systems = [hairpin_design(100 + k, n) for k, n in enumerate([30, 77, 150, 64, 211])]
    batch = Engine([s for s, _ in systems], keys=list(range(len(systems))), seed=3)
    for k, (s, x) in enumerate(systems):
        batch.set_coords(k, x)
    batch.neighbor(-1)
    fb = batch.force_all()
    for k, (s, x) in enumerate(systems):
        alone = Engine([s], keys=[k], seed=3)
        alone.set_coords(0, x)
        alone.neighbor(0)
        fa = alone.force_sumup(0)
        assert np.array_equal(fa, fb[k])


def test_mixed_forms_batch_equals_alone_bitwise():
    """Every parity system (Debye-Hueckel and Coulomb chains, a periodic DH
    system, Ewald boxes) in one batch: the electrostatic chunk-pair tasks of
    all forms share one force pass, in the batch's order, and each system's
    forces equal its forces alone, bitwise."""
    batch = Engine([s for _, s, _ in CASES], keys=list(range(len(CASES))), seed=3)
    for k, (_, s, x) in enumerate(CASES):
        batch.set_coords(k, x)
    batch.neighbor(-1)
    fb = batch.force_all()
    for k, (name, s, x) in enumerate(CASES):
        print(f"[mixed batch] {name}", flush=True)
        alone = Engine([s], keys=[k], seed=3)
        alone.set_coords(0, x)
        alone.neighbor(0)
        fa = alone.force_sumup(0)
        assert np.array_equal(fa, fb[k]), name
        alone.close()
    batch.close()


def test_langevin_matches_oracle_short():
    s, x = hairpin64(0, "helix")
    v = np.zeros_like(x)
    eng = Engine([s], keys=[9], seed=11)
    eng.set_coords(0, x)
    eng.set_state(0, 1, v)
    eng.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
    eng.langevin_steps(30, MODE_EAGER)
    xg = eng.coords(0)
    o = oracle.Oracle(s)
    xo, vo, ao, nr = o.langevin(x, v, np.zeros_like(x), 30, tstep=0.05, rebuild_every=10, seed=11, sys_id=9)
    d = np.abs(xg - xo).max()
    print(f"   30-step max |dx| = {d:.3e} A", flush=True)
    assert d < 1e-4


@pytest.mark.parametrize("seed", [10, 12])
def test_langevin_valence_limited_matches_oracle(seed):
    """A compact chain whose every base pair is an hbond candidate, so that
    the valence-limit deletion of force_dtrna_hbond15.F90:373-510 runs at
    each force step: 60 Langevin steps (list rebuilds at 10-step intervals)
    against the oracle's, and the deletions are counted."""
    s, x = compact_allpairs(seed=seed)
    print(f"[valence trajectory] compact{seed} nmp={s.nmp}", flush=True)
    v = np.zeros_like(x)
    eng = Engine([s], keys=[seed], seed=23)
    eng.set_coords(0, x)
    eng.set_state(0, 1, v)
    c0 = eng.counters()["hb_deleted"]
    eng.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
    eng.langevin_steps(60, MODE_EAGER)
    ndel = eng.counters()["hb_deleted"] - c0
    xg = eng.coords(0)
    eng.close()
    o = oracle.Oracle(s)
    xo, vo, ao, nr = o.langevin(x, v, np.zeros_like(x), 60, tstep=0.05, rebuild_every=10, seed=23, sys_id=seed)
    d = np.abs(xg - xo).max()
    print(f"   {ndel} hbonds deleted; 60-step max |dx| = {d:.3e} A", flush=True)
    assert ndel > 0
    assert np.isfinite(xg).all()
    assert d < 1e-4


def test_valence_long_site_sequence():
    """One G with 72 formed hbonds on each of its sites: the valence pass
    runs its first deletions on site sequences longer than a wave (the
    scratch-sequence path) and deletes 71; the deletion set, forces and
    hbond energy equal the oracle's."""
    s, x = valence_crowd()
    print(f"[valence crowd] nmp={s.nmp} nhb={len(s.hb_mp)}", flush=True)
    eng = Engine([s], keys=[0], seed=0)
    eng.set_coords(0, x)
    eng.neighbor(0)
    c0 = eng.counters()["hb_deleted"]
    f = eng.force_sumup(0)
    ndel = eng.counters()["hb_deleted"] - c0
    e = eng.energy_sumup(0)
    eng.close()
    o = oracle.Oracle(s)
    o.neighbor(x)
    fo, _ = o.force(x, seed=0, sys_id=0, step=0)
    eo = o.energy(x)
    err = force_error(f, fo)
    print(f"   {ndel} deletions, max rel err {err.max():.3e}", flush=True)
    assert ndel == len(s.hb_mp) - 1
    assert err.max() <= 1e-5
    assert abs(e[16] - eo[16]) <= 1e-7 * max(abs(eo[16]), 1e-6)


@pytest.mark.parametrize("rebuild_every", [10, 0])
@pytest.mark.parametrize("mode", [MODE_GRAPH, MODE_EAGER, MODE_PLAIN])
def test_list_overflow_resumes_bitwise(mode, rebuild_every):
    """A run whose list capacity is set to the largest list of the start
    configuration: a later rebuild needs more entries, the run stops at that
    step, grows the capacity and resumes.  The trajectory equals, bit for bit,
    that of a run with ample capacity from the start (so no step ran on
    truncated lists), and the resumes are counted."""
    # compact chains expanded by 25 % about their centre: the stretched bonds
    # pull them back together, so later rebuilds hold more neighbours
    systems = []
    for k, n in enumerate([40, 64, 90]):
        s, x = compact_allpairs(seed=41 + k, n=n)
        c = x.mean(axis=0)
        systems.append((s, c + 1.25 * (x - c)))
    systems += [hairpin_design(400, 30)]
    res = []
    for tight in (False, True):
        print(f"[overflow resume] mode={mode} rebuild_every={rebuild_every} tight={tight}", flush=True)
        eng = Engine([s for s, _ in systems], keys=list(range(len(systems))), seed=7)
        for k, (s, x) in enumerate(systems):
            eng.set_coords(k, x)
        eng.langevin_init(tstep=0.05, rebuild_every=rebuild_every)
        if tight:
            need = 0
            for k in range(len(systems)):
                p = eng.pairs(k, LIST_EXV)
                ids = np.concatenate([p[:, 0], p[:, 1]]) if len(p) else np.zeros(0, int)
                need = max(need, int(np.bincount(ids).max()) if len(ids) else 0)
            assert eng.list_capacity(need) == need
        eng.langevin_steps(120, mode)
        c = eng.counters()
        res.append(([eng.coords(k) for k in range(len(systems))], c))
        if tight:
            print(f"   resumes {c['resumes']} capacity {eng.list_capacity()}", flush=True)
            assert c["resumes"] >= 1 and eng.list_capacity() > need
            assert c["nb_overflow"] == 0 and c["run_completed"] == 0
        eng.close()
    for k in range(len(systems)):
        assert np.array_equal(res[0][0][k], res[1][0][k])
    for key in ("steps", "rebuilds", "hb_deleted"):
        assert res[0][1][key] == res[1][1][key], key


def test_modes_bitwise_equal():
    systems = [hairpin_design(200 + k, n) for k, n in enumerate([40, 90, 120])]
    res = []
    for mode in (MODE_EAGER, MODE_GRAPH, MODE_PLAIN):
        eng = Engine([s for s, _ in systems], seed=5)
        for k, (s, x) in enumerate(systems):
            eng.set_coords(k, x)
        eng.langevin_init(tstep=0.05, rebuild_every=10)
        eng.langevin_steps(40, mode)
        res.append([eng.coords(k) for k in range(len(systems))])
    for k in range(len(systems)):
        assert np.array_equal(res[0][k], res[1][k])
        assert np.array_equal(res[0][k], res[2][k])


def _small_margin_systems(margin=1.5):
    """Chains with a 1.5 A list margin, so that the displacement trigger
    fires every few steps."""
    out = []
    for k, n in enumerate([64, 30, 150]):
        rng = np.random.default_rng(300 + k)
        seq, pairs = T.hairpin_design(rng, n, 6, 4)
        s = T.build_rna_system([(seq, pairs)], neigh_margin=margin, name=f"m{k}")
        out.append((s, T.helix_coords(s, rng)))
    return out


@pytest.mark.parametrize("which", [0, 2])
def test_displacement_trigger_matches_oracle(which):
    """Displacement-triggered rebuilds (time_integral.F90:57-93): the GPU
    trajectory and rebuild count against the oracle's trigger loop (a
    64-nt and a 150-nt chain with a small margin)."""
    s, x = _small_margin_systems()[which]
    print(f"[trigger] system {which} nmp={s.nmp}", flush=True)
    v = np.zeros_like(x)
    eng = Engine([s], keys=[4], seed=17)
    eng.set_coords(0, x)
    eng.set_state(0, 1, v)
    c0 = eng.counters()["rebuilds"]
    eng.langevin_init(tstep=0.05, rebuild_every=0, step0=0, init_accel=True)
    eng.langevin_steps(60, MODE_EAGER)
    ng = eng.counters()["rebuilds"] - c0
    xg = eng.coords(0)
    o = oracle.Oracle(s)
    xo, vo, ao, nr = o.langevin(x, v, np.zeros_like(x), 60, tstep=0.05, rebuild_every=0, seed=17, sys_id=4)
    d = np.abs(xg - xo).max()
    print(f"   rebuilds gpu {ng} oracle {nr}; 60-step max |dx| = {d:.3e} A", flush=True)
    assert nr >= 4
    assert ng == nr
    assert d < 1e-4


def test_triggered_rebuild_lists_match_oracle():
    """The lists of a displacement-triggered rebuild (only the fired
    system's tiles working in the gated build) are the oracle's lists at the
    coordinates of that step, pair for pair."""
    s, x = _small_margin_systems()[2]
    eng = Engine([s], keys=[4], seed=17)
    eng.set_coords(0, x)
    eng.langevin_init(tstep=0.05, rebuild_every=0, step0=0, init_accel=True)
    checked = 0
    for step in range(80):
        xb = eng.coords(0)
        r0 = eng.counters()["rebuilds"]
        eng.langevin_steps(1, MODE_EAGER)
        if eng.counters()["rebuilds"] == r0:
            continue
        o = oracle.Oracle(s)
        o.neighbor(xb)
        pe, oe = eng.pairs(0, LIST_EXV), o.lists(0)
        print(f"[triggered lists] step {step + 1}: {len(pe)} exv pairs", flush=True)
        assert len(pe) == len(oe)
        assert np.array_equal(pe[np.lexsort(pe.T[::-1])], oe[np.lexsort(oe.T[::-1])])
        checked += 1
        if checked == 3:
            break
    eng.close()
    assert checked >= 2


def test_displacement_trigger_modes_and_batch_bitwise():
    """With per-system rebuild decisions, graph == eager == plain and each
    system in the batch == the same system alone, bitwise."""
    systems = _small_margin_systems()
    res = []
    for mode in (MODE_EAGER, MODE_GRAPH, MODE_PLAIN):
        eng = Engine([s for s, _ in systems], keys=[0, 1, 2], seed=5)
        for k, (s, x) in enumerate(systems):
            eng.set_coords(k, x)
        eng.langevin_init(tstep=0.05, rebuild_every=0)
        eng.langevin_steps(40, mode)
        res.append([eng.coords(k) for k in range(len(systems))])
        eng.close()
    for k in range(len(systems)):
        assert np.array_equal(res[0][k], res[1][k])
        assert np.array_equal(res[0][k], res[2][k])
    for k, (s, x) in enumerate(systems):
        alone = Engine([s], keys=[k], seed=5)
        alone.set_coords(0, x)
        alone.langevin_init(tstep=0.05, rebuild_every=0)
        alone.langevin_steps(40, MODE_GRAPH)
        assert np.array_equal(alone.coords(0), res[1][k])
        alone.close()


@pytest.mark.parametrize("style", [0, 1])
def test_energy_unit_matches_oracle(style):
    """energy_unit (per unit pair and type, post-processed) through the
    C-ABI against the FP64 oracle, on two chains + ions, ion-rich compact
    chains (one periodic) and an Ewald box."""
    for name, s, x in unit_cases():
        print(f"[energy_unit] {name} nunit={s.nunit} style={style}", flush=True)
        v = np.random.default_rng(5).normal(size=x.shape)
        eng = Engine([s], keys=[0], seed=0)
        eng.set_coords(0, x)
        eng.set_state(0, 1, v)
        eng.neighbor(0)
        eng.force_sumup(0)
        eu = eng.energy_unit(0, style)
        o = oracle.Oracle(s)
        o.neighbor(x)
        o.force(x, seed=0, sys_id=0, step=0)
        euo = o.energy_unit(x, v, style)
        scale = np.abs(euo).max()
        err = np.abs(eu - euo) / np.maximum(np.abs(euo), 1e-6 * scale)
        print(f"   max rel err {err.max():.3e}", flush=True)
        assert err.max() <= 1e-7
        eng.close()

