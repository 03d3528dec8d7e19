"""CPU tests of the .ninfo / coarse-grained PDB reader and writer, the
tertiary-stack generator, and the oracle on the reference's bwyv19 fixture
against the golden vectors the reference's own routines produced
(tests/golden/make_bwyv19.py)."""
import hashlib
import os

import numpy as np
import pytest

from oracle import Oracle
from tis_amd import ninfo as NI
from tis_amd import topology as T
from tis_amd.sysdesc import E_STACK, E_TOTAL, E_TSTACK

from systems import GOLDEN, bwyv19, compact_allpairs, tstack_hairpin

REF_NPZ = os.path.join(GOLDEN, "bwyv19_ref.npz")


def test_bwyv19_records():
    rec = NI.read_ninfo(os.path.join(GOLDEN, "bwyv19", "bwyv19.ninfo"))
    assert (len(rec.bond), len(rec.angl), len(rec.bs_dist), len(rec.bs_dihd)) == (84, 111, 27, 54)
    assert (len(rec.hb_dist), len(rec.hb_angl), len(rec.hb_dihd)) == (108, 216, 324)
    assert (len(rec.tbs_dist), len(rec.tbs_angl), len(rec.tbs_dihd)) == (1, 2, 3)
    assert sum(1 for r in rec.hb_dist if r[6] == "T") == 18


def test_bwyv19_system_layout():
    """Bead roles as read_nativeinfo.F90 assigns them: a stack of bases
    (imp1, imp2) takes P1 = imp1-2, S1 = imp1-1, P2 = imp1+1, S2 = imp1+2,
    P3 = imp1+4 (:482-488); a tertiary stack of (imp1, imp2) takes S = imp-1
    and P = imp+1 (:665-670); the +2/-2 types of the tbs-dist record make both
    sides exclusive, and mloop_dtrna.F90:73-136 maps them to the stacks of
    beads (36, 39) and (39, 42) (1-based)."""
    s, x = bwyv19()
    assert s.nmp == 85 + 206 and s.nunit == 2
    assert s.st_mp[0].tolist() == [2, 5, 0, 1, 3, 4, 6]
    assert s.tst_mp[0].tolist() == [35, 41, 34, 40, 36, 42]
    assert s.tst_excl[0].tolist() == [1, 1]
    assert [s.st_mp[k, :2].tolist() for k in s.tst2st[0]] == [[35, 38], [38, 41]]
    assert int(s.hb_tertiary.sum()) == 18
    # hbond sites: every listed atom maps to a site of its own bead
    offs, val = T.hbsite_layout(NI.bead_names(s))
    assert len(val) == len(s.hbsite_valence)
    for h in range(len(s.hb_mp)):
        for q in range(6):
            site = int(s.hb_site[h, q])
            if site < 0:
                continue
            bead = int(s.hb_mp[h, q % 2])
            nxt = min([o for o in offs.values() if o > offs[bead]], default=len(val))
            assert offs[bead] <= site < nxt


def _same_system(a, b, rtol=1e-14):
    for k in ("unit", "mptype", "iclass", "bd_mp", "ba_mp", "st_mp", "hb_mp", "hb_site", "hb_tertiary",
              "hbsite_valence", "tst_mp", "tst2st", "tst_excl"):
        assert np.array_equal(np.asarray(getattr(a, k)), np.asarray(getattr(b, k))), k
    for k in ("bd_nat", "bd_coef", "ba_nat", "ba_coef", "st_nat", "st_coef", "hb_nat", "hb_coef",
              "tst_nat", "tst_coef", "charge", "exv_rad", "exv_eps", "mass", "fric"):
        u, v = np.asarray(getattr(a, k)), np.asarray(getattr(b, k))
        assert u.shape == v.shape, k
        assert np.all(np.abs(u - v) <= rtol * np.maximum(np.abs(u), 1.0)), k


@pytest.mark.parametrize("which", ["bwyv19", "tstack_hairpin", "compact_allpairs"])
def test_ninfo_round_trip(tmp_path, which):
    """write_ninfo then read_ninfo gives back the same system (angles to a
    few ulps through the degree conversion); write_pdb_cg then read_pdb_cg
    gives back the beads, units and coordinates to the format's 1e-3 A."""
    if which == "bwyv19":
        s, x = bwyv19()
        kw = dict(tempk=310.15, ionic_strength=0.057, dh_cutoff_type=1, cutoff_ele=50.0)
    elif which == "tstack_hairpin":
        s, x = tstack_hairpin()
        kw = {}
    else:
        s, x = compact_allpairs(seed=12, ions=True)
        kw = {}
    nf, pf = tmp_path / "s.ninfo", tmp_path / "s.pdb"
    NI.write_ninfo(s, nf)
    NI.write_pdb_cg(s, x, pf)
    beads = NI.read_pdb_cg(str(pf))
    assert beads.names == NI.bead_names(s)
    assert np.array_equal(np.asarray(beads.unit), s.unit)
    assert np.abs(beads.xyz - x).max() <= 5.1e-4
    s2 = NI.system_from_ninfo(NI.read_ninfo(nf), beads, **kw)
    _same_system(s, s2)


def test_pdb_reader_columns():
    """The ion file's coordinates run together ('-205.195-149.235'): they are
    read by column, as the reference's fixed format does."""
    b = NI.read_pdb_cg(os.path.join(GOLDEN, "bwyv19", "bwyv19_force_check.ion.pdb"))
    assert len(b.names) == 206 and set(b.names) == {"Mg"}
    assert b.xyz[0].tolist() == [-205.195, -149.235, 130.726]


def test_tertiary_stack_generator():
    """tstack_hairpin: the map to consecutive stacks (mloop_dtrna.F90:73-136),
    a side with no consecutive stack made non-exclusive, and the st_status
    switch-off of force_dtrna_stack_nlocal.F90:83-92 (d - d0 < 10 only)."""
    import ctypes as C
    import oracle
    s, x = tstack_hairpin()
    assert s.tst2st.tolist() == [[2, 18], [2, 32], [11, 38], [24, -1], [6, 43]]
    assert s.tst_excl.tolist() == [[1, 1], [1, 0], [0, 0], [1, 0], [1, 1]]
    o = Oracle(s)
    st = np.ones(len(s.st_mp), np.uint8)
    f = np.zeros_like(x)
    oracle.lib().oracle_force_tstack(C.byref(o.cs), oracle.fp(x), oracle.fp(x),
                                     st.ctypes.data_as(C.c_char_p), oracle.fp(f))
    assert np.nonzero(st == 0)[0].tolist() == [2, 18, 24]


def test_tstack_force_is_minus_gradient():
    """Central differences of E_TSTACK + E_STACK (st_status re-derived at
    each displaced configuration) against the analytic stack forces."""
    s, x = tstack_hairpin()
    o = Oracle(s)
    o.neighbor(x)
    f = o.term("stack", x)

    def e(xx):
        o.neighbor(xx)
        o.force(xx)
        en = o.energy(xx)
        return en[E_TSTACK] + en[E_STACK]

    h = 1e-6
    beads = np.unique(s.tst_mp.ravel())
    for i in beads[::2]:
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            fd = -(e(xp) - e(xm)) / (2 * h)
            assert abs(fd - f[i, d]) <= 1e-5 * max(1e-2, abs(f[i, d])), (i, d, fd, f[i, d])


def _digest(p):
    p = np.ascontiguousarray(np.asarray(p, dtype=np.int32)[np.lexsort(np.asarray(p).T[::-1])])
    return hashlib.sha256(p.tobytes()).hexdigest()


def test_oracle_matches_reference_golden_bwyv19():
    """The oracle on bwyv19 against the reference's own outputs committed in
    tests/golden/bwyv19_ref.npz: lists bit-exact, every force term 1e-12,
    the valence pass on the reference's Mersenne-Twister words (deletion set
    exact), energies and energy_unit 1e-12."""
    g = np.load(REF_NPZ)
    s, x = bwyv19()
    assert np.array_equal(x, g["xyz"])
    o = Oracle(s)
    o.neighbor(x)
    exv = o.lists(0)
    assert np.array_equal(exv[np.lexsort(exv.T[::-1])], g["exv_pairs"])
    ele, coef = o.lists(1)
    assert len(ele) == int(g["ele_count"]) and _digest(ele) == str(g["ele_digest"])
    assert abs(np.sort(coef).sum() - float(g["ele_coef_sum"])) <= 1e-12 * abs(float(g["ele_coef_sum"]))
    assert np.array_equal(np.sort(o.lists(2)), g["hb_list"])
    for t in ("bond", "angle", "stack", "exv", "dh"):
        fo = o.term(t, x)
        fr = g["f_" + t]
        assert np.abs(fo - fr).max() <= 1e-12 * max(np.abs(fr).max(), 1e-30), t
    fo, so, eo, nd, _ = o.hbond_mt(x, g["mt_words"])
    assert nd > 0 and np.array_equal(so.astype(np.int32), g["hb_status"])
    assert np.abs(fo - g["f_hbond"]).max() <= 1e-12 * np.abs(g["f_hbond"]).max()
    o.force_mt(x, g["mt_words"])
    e = o.energy(x, g["velo"])
    er = g["energy"]
    assert np.all(np.abs(e - er) <= 1e-12 * np.maximum(np.abs(er), 1e-6))
    assert er[E_TSTACK] != 0.0 and er[E_TOTAL] != 0.0
    for style in (0, 1):
        eu = o.energy_unit(x, g["velo"], style)
        eur = g["energy_unit%d" % style]
        assert np.abs(eu - eur).max() <= 1e-12 * np.abs(eur).max()
