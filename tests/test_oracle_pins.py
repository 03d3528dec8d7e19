"""CPU tests that pin the FP64 oracle (oracle/tis_oracle.c).

Three kinds of evidence:
  * the reference's own Fortran routines, compiled from its sources into
    oracle/_ref/libtisref.so (skipped where that library is absent);
  * published known-answer vectors (Philox4x32-10);
  * what the mathematics fixes: central-difference gradients of the FP64
    energies, Newton's third law, closed forms of the pair potentials, the
    exact stationary variance of the Langevin scheme for a free bead.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import Oracle, RefLib
from tis_amd import topology as T
from tis_amd.sysdesc import (E_BANGLE, E_BOND, E_ELE, E_EXV, E_HBOND, E_STACK, E_TOTAL,
                             E_TSTACK)

from systems import compact_allpairs, hairpin64, parity_cases, unit_cases

CASES = parity_cases(ewald=False)
needs_ref = pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (reference absent)")


# ----------------------------------------------------------------- Philox
def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32 with 10 rounds."""
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_gaussian_stream_is_standard_normal():
    from scipy import stats
    g = np.array([oracle.gauss3(7, 3, step, bead) for step in range(40) for bead in range(400)]).ravel()
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01
    assert stats.kstest(g, "norm").pvalue > 1e-3
    # distinct (system, step, bead) give distinct numbers
    assert not np.array_equal(oracle.gauss3(7, 3, 1, 5), oracle.gauss3(7, 4, 1, 5))


# ----------------------------------------------------------------- topology
@pytest.mark.parametrize("n", [8, 20, 64])
def test_linear_chain_counts(n):
    """N = 3n-1 beads, 3n-2 bonds, 4n-5 angles, n-3 stacks (5'-S linear chain)."""
    rng = np.random.default_rng(n)
    s = T.build_rna_system([(T.random_sequence(rng, n), [])])
    assert s.nmp == 3 * n - 1
    assert len(s.bd_nat) == 3 * n - 2
    assert len(s.ba_nat) == 4 * n - 5
    assert len(s.st_mp) == n - 3


# ----------------------------------------------------------------- reference
def _sorted(p):
    p = np.asarray(p)
    if p.ndim == 1:
        return np.sort(p)
    return p[np.lexsort(p.T[::-1])]


@needs_ref
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_lists_match_reference(case):
    name, s, x = case
    o = Oracle(s)
    o.neighbor(x)
    R = RefLib()
    R.load(s)
    R.neighbor(x)
    assert np.array_equal(_sorted(o.lists(0)), _sorted(R.lists(0)))
    oq, oc = o.lists(1)
    rq, rc = R.lists(1)
    assert np.array_equal(_sorted(oq), _sorted(rq))
    assert np.array_equal(oc[np.lexsort(oq.T[::-1])], rc[np.lexsort(rq.T[::-1])])
    assert np.array_equal(np.sort(o.lists(2)), np.sort(R.lists(2)))


MT_SEED = 20261203
MT_WORDS = 1 << 18


@needs_ref
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_forces_match_reference(case):
    """Every term against the reference routine.  The hbond term runs the
    valence pass of both on the same Mersenne-Twister stream: the reference
    draws from its replica stream mts(1,0), the oracle from the words of an
    identically seeded stream."""
    name, s, x = case
    o = Oracle(s)
    o.neighbor(x)
    R = RefLib()
    R.load(s)
    R.neighbor(x)
    for t in ["bond", "angle", "stack", "exv", "dh"]:
        fr = R.force_term(t)
        fo = o.term(t, x)
        if isinstance(fo, tuple):
            fo = fo[0]
        scale = max(np.abs(fr).max(), 1e-30)
        assert np.abs(fr - fo).max() <= 1e-12 * scale, t
    R.mt_seed(MT_SEED)
    fr = R.force_term("hbond")
    sr, er = R.hb_state()
    fo, so, eo, _, _ = o.hbond_mt(x, R.mt_raw(MT_SEED, MT_WORDS))
    assert np.array_equal(sr, so.astype(np.int32))
    assert np.abs(er - eo).max(initial=0.0) <= 1e-12 * max(np.abs(er).max(initial=0.0), 1e-30)
    assert np.abs(fr - fo).max() <= 1e-12 * max(np.abs(fr).max(), 1e-30)


def test_mt_stream_known_answer():
    """The reference's mt_stream with the scalar seed 5489 is MT19937
    (Matsumoto & Nishimura 1998): its first outputs are the published
    3499211612, 581869302, 3890346734, 3586334585, 545404204."""
    if not RefLib.available():
        pytest.skip("oracle/_ref not built (reference absent)")
    assert RefLib().mt_raw(5489, 5).tolist() == [3499211612, 581869302, 3890346734, 3586334585, 545404204]


def _valence_cases():
    out = []
    for seed in (10, 12, 41, 42, 43):
        for n in (40, 64):
            s, x = compact_allpairs(seed=seed, n=n)
            out.append((f"compact{seed}_{n}", s, x))
    return out


@needs_ref
@pytest.mark.parametrize("mt_seed", [1, 77, 20261203])
def test_valence_pass_matches_reference(mt_seed):
    """force_dtrna_hbond15.F90:373-510: with the same Mersenne-Twister words
    the oracle deletes exactly the hbonds the reference deletes (site pick,
    shuffle and Metropolis victim pick), over compact chains whose every
    canonical pair is a candidate, so many sites exceed their valence."""
    total = 0
    for name, s, x in _valence_cases():
        o = Oracle(s)
        o.neighbor(x)
        R = RefLib()
        R.load(s)
        R.neighbor(x)
        R.mt_seed(mt_seed)
        fr = R.force_term("hbond")
        sr, er = R.hb_state()
        fo, so, eo, nd, used = o.hbond_mt(x, R.mt_raw(mt_seed, MT_WORDS))
        assert np.array_equal(sr, so.astype(np.int32)), name
        assert np.abs(er - eo).max(initial=0.0) <= 1e-12 * max(np.abs(er).max(initial=0.0), 1e-30), name
        assert np.abs(fr - fo).max() <= 1e-12 * max(np.abs(fr).max(), 1e-30), name
        total += nd
    assert total >= 20


@needs_ref
def test_valence_pass_rejects_wrong_stream():
    """The pin is sharp: a stream shifted by one draw (two words) changes
    the deletion set of at least one case."""
    differ = 0
    for name, s, x in _valence_cases():
        o = Oracle(s)
        o.neighbor(x)
        R = RefLib()
        R.load(s)
        R.neighbor(x)
        R.mt_seed(5)
        R.force_term("hbond")
        sr, _ = R.hb_state()
        _, so, _, nd, _ = o.hbond_mt(x, R.mt_raw(5, MT_WORDS)[2:])
        differ += int(not np.array_equal(sr, so.astype(np.int32)))
    assert differ >= 1


@needs_ref
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_energies_match_reference(case):
    name, s, x = case
    o = Oracle(s)
    o.neighbor(x)
    R = RefLib()
    R.load(s)
    R.neighbor(x)
    o.force_mt(x, R.mt_raw(MT_SEED, MT_WORDS))
    R.force_term("stack")
    R.mt_seed(MT_SEED)
    R.force_term("hbond")
    v = np.random.default_rng(3).normal(size=x.shape)
    e = o.energy(x, v)
    er = R.energy(v)
    for k in (E_BOND, E_BANGLE, E_ELE, E_STACK, E_HBOND, E_EXV, E_TSTACK, E_TOTAL, 1):
        assert abs(e[k] - er[k]) <= 1e-12 * max(abs(er[k]), 1e-6), k


UNIT_CASES = unit_cases(CASES)


@needs_ref
@pytest.mark.parametrize("case", UNIT_CASES, ids=[c[0] for c in UNIT_CASES])
@pytest.mark.parametrize("style", [0, 1])
def test_energy_unit_matches_reference(case, style):
    """energy_unit of energy_sumup (per unit pair and type, after the
    post-processing of i_output_energy_style) against the reference's term
    routines and post-processing."""
    name, s, x = case
    o = Oracle(s)
    o.neighbor(x)
    R = RefLib()
    R.load(s)
    R.neighbor(x)
    o.force_mt(x, R.mt_raw(MT_SEED, MT_WORDS))
    R.force_term("stack")
    R.mt_seed(MT_SEED)
    R.force_term("hbond")
    v = np.random.default_rng(3).normal(size=x.shape)
    eu = o.energy_unit(x, v, style)
    eur = R.energy_unit(v, style)
    assert s.nunit >= 2 and eu.shape == (s.nunit, s.nunit, 25)
    assert np.abs(eu - eur).max() <= 1e-12 * max(np.abs(eur).max(), 1e-6)


@pytest.mark.parametrize("case", UNIT_CASES, ids=[c[0] for c in UNIT_CASES])
def test_energy_unit_diagonal_sums_to_energy(case):
    """With style 0 each off-diagonal pair energy is split between the two
    diagonal entries: the diagonal summed over units gives every term's
    total (pairs are stored with the lower bead's unit first)."""
    name, s, x = case
    o = Oracle(s)
    o.neighbor(x)
    o.force(x)
    v = np.random.default_rng(4).normal(size=x.shape)
    e = o.energy(x, v)
    eu = o.energy_unit(x, v, 0)
    diag = np.einsum("iik->k", eu)
    assert np.abs(np.tril(eu[:, :, E_EXV], -1)).max() == 0.0
    ewald = s.params.ele_form == 2
    if ewald:
        # energy_ele_coulomb_ewld.F90:168-170 adds nothing per unit
        assert np.abs(eu[:, :, E_ELE]).max() == 0.0
        e = e.copy()
        e[E_TOTAL] -= e[E_ELE]
        e[E_ELE] = 0.0
    for k in range(25):
        assert abs(diag[k] - e[k]) <= 1e-12 * max(abs(e[k]), 1.0), k
    # style 1 keeps the raw diagonal and symmetric off-diagonal sums
    eu1 = o.energy_unit(x, v, 1)
    assert np.array_equal(eu1[:, :, E_BOND], eu1[:, :, E_BOND].T)


@needs_ref
@pytest.mark.parametrize("tempk,ionic", [(310.0, 0.150), (290.0, 0.05), (400.0, 1.0)])
def test_temperature_parameters_match_reference(tempk, ionic):
    s, x = hairpin64()
    R = RefLib()
    R.load(s)
    coef, cdist, zp, u0, beta = R.temperature(tempk, ionic)
    c2, d2, z2 = T.ele_params(tempk, ionic)
    assert coef == c2 and cdist == d2 and zp == z2
    ours = np.array([T.stack_u0(h, sv, tm, tempk) for h, sv, tm in s.st_hsTm])
    assert np.array_equal(ours, u0)
    assert beta == T.beta_hb(tempk)


# ----------------------------------------------------------------- mathematics
def _energy_terms(o, x):
    o.neighbor(x)
    o.force(x)
    return o.energy(x)


@pytest.mark.parametrize("which", ["helix", "compact"])
def test_force_is_minus_gradient(which):
    """Central differences of the FP64 energy (the reference's own check,
    simu_checkforce.F90:99-118) against the analytic forces."""
    for seed in range(4, 40):
        s, x = hairpin64(seed=seed, coords="helix")
        if which == "compact":
            x = T.compact_coords(s, np.random.default_rng(seed), box=40, min_dist=2.9)
        o = Oracle(s)
        o.neighbor(x)
        f, ncl = o.force(x)
        if ncl == 0:  # DE_MAX clamps break F = -grad U by design
            break
    assert ncl == 0
    rng = np.random.default_rng(0)
    h = 1e-5
    for i in rng.choice(s.nmp, 12, replace=False):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            ep = o.energy(xp)[E_TOTAL]
            em = o.energy(xm)[E_TOTAL]
            fd = -(ep - em) / (2 * h)
            assert abs(fd - f[i, d]) <= 2e-5 * max(1.0, abs(f[i, d])), (i, d, fd, f[i, d])


def test_hbond_force_is_minus_gradient():
    s, x = compact_allpairs(seed=10)
    o = Oracle(s)
    o.neighbor(x)
    f, st, ene = o.term("hbond", x)
    assert st.sum() > 0
    h = 1e-6
    active = np.nonzero(st)[0]
    beads = np.unique(s.hb_mp[active].ravel())

    def e_hb(xx):
        o.neighbor(x)  # keep the candidate list
        _, st2, en2 = o.term("hbond", xx)
        return en2[active].sum() if np.array_equal(st2[active], st[active]) else np.nan

    for i in beads[:8]:
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            fd = -(e_hb(xp) - e_hb(xm)) / (2 * h)
            if np.isnan(fd):
                continue
            assert abs(fd - f[i, d]) <= 1e-5 * max(1e-3, abs(f[i, d])), (i, d)


@pytest.mark.parametrize("case", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_newton_third_law(case):
    name, s, x = case
    o = Oracle(s)
    o.neighbor(x)
    f, _ = o.force(x)
    assert np.abs(f.sum(axis=0)).max() <= 1e-9 * max(1.0, np.abs(f).max())


def _two_beads(kind, r):
    """Two RNA bases in different units (no bonded terms) at distance r."""
    s = T.build_rna_system([("A", []), ("A", [])])
    keep = [1, 3]  # the two bases
    x = np.zeros((s.nmp, 3))
    x[0] = (-50, 0, 0)
    x[2] = (50, 0, 0)
    x[1] = (0, 0, 0)
    x[3] = (r, 0, 0)
    return s, x, keep


def test_exv_closed_form():
    """mWCA: U = eps[(a/dr)^12 - 2(a/dr)^6 + 1], dr = r + a - D, for r <= D."""
    a, D, eps = 1.5852, 3.2, 0.2
    for r in (2.0, 2.5, 3.0, 3.19):
        s, x, (i, j) = _two_beads("exv", r)
        o = Oracle(s)
        o.neighbor(x)
        e = o.energy(x)[E_EXV]
        dr = r + a - D
        u = eps * ((a / dr) ** 12 - 2 * (a / dr) ** 6 + 1)
        assert abs(e - u) <= 1e-10 * max(1, abs(u))
        f = o.term("exv", x)
        du = eps * (-12 * a ** 12 / dr ** 13 + 12 * a ** 6 / dr ** 7)
        fx = min(-du / r, 50.0) * r  # clamp of |dU/dr|/r at DE_MAX
        assert abs(f[j, 0] - fx) <= 1e-9 * max(1, abs(fx))
        assert abs(f[i, 0] + fx) <= 1e-9 * max(1, abs(fx))
    s, x, (i, j) = _two_beads("exv", 3.3)
    o = Oracle(s)
    o.neighbor(x)
    assert o.energy(x)[E_EXV] == 0.0 and np.all(o.term("exv", x) == 0.0)


def test_dh_closed_form():
    """Debye-Hueckel between two phosphates: F = q q C e^{-r/l}(1/r + 1/l)/r."""
    s = T.build_rna_system([("AA", []), ("AA", [])])
    p = np.nonzero(s.mptype == 2)[0]
    x = np.zeros((s.nmp, 3))
    x[:, 1] = np.arange(s.nmp) * 40.0  # everything far apart
    x[p[1]] = x[p[0]] + (9.0, 0, 0)
    o = Oracle(s)
    o.neighbor(x)
    q = s.charge[p[0]] * s.charge[p[1]] * s.params.ele_coef
    lam = s.params.ele_cdist
    e = o.energy(x)[E_ELE]
    # every other pair is >= 40 A apart; subtract nothing: compare the 9 A pair
    pairs, coef = o.lists(1)
    k = [n for n, (a, b, _) in enumerate(pairs) if {a, b} == {p[0], p[1]}][0]
    f = o.term("dh", x, pairs[k:k + 1], coef[k:k + 1])
    r = 9.0
    fx = q * math.exp(-r / lam) * (1 / r + 1 / lam) / r
    assert abs(f[p[1], 0] - fx) <= 1e-12 * abs(fx)
    assert abs(f[p[0], 0] + fx) <= 1e-12 * abs(fx)
    assert e > 0


def free_beads_system(nunits=150):
    """Units of one nucleotide (S, B) with every interaction switched off."""
    s = T.build_rna_system([("A", [])] * nunits, ele_form=-1)
    s.bd_coef[:] = 0.0
    s.ba_coef[:] = 0.0
    s.exv_eps[:] = 0.0
    x = np.zeros((s.nmp, 3))
    x[:, 0] = np.arange(s.nmp) * 40.0
    return s, x


def test_langevin_free_bead_stationary_variance():
    """With F = 0 the scheme v' = c2 v + c3 c1 (xi_n + xi_n+1) has stationary
    variance 2 (c1 c3)^2 / (1 - c2) exactly (time_integral.F90:169-183)."""
    s, x = free_beads_system()
    o = Oracle(s)
    rl = o.langevin_consts(0.05)
    v = np.zeros_like(x)
    a = np.zeros_like(x)
    xx, v, a, _ = o.langevin(x, v, a, 4000, seed=3)
    snaps = []
    for k in range(6):
        xx, v, a, _ = o.langevin(xx, v, a, 1500, seed=3, step0=4000 + 1500 * k, init_accel=False)
        snaps.append(v.copy())
    vv = np.array(snaps)
    c1, c2, c3 = rl[:, 0], rl[:, 1], rl[:, 2]
    expect = 2 * (c1 * c3) ** 2 / (1 - c2)
    for t in (s.mptype == 4, s.mptype == 3):
        got = vv[:, t, :].var()
        assert abs(got / expect[t][0] - 1) < 0.06, (got, expect[t][0])
