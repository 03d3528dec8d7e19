"""Seeded test systems shared by the CPU and GPU tests."""
import os

import numpy as np

from tis_amd import topology as T
from tis_amd.sysdesc import ELE_COULOMB, ELE_EWALD


def hairpin64(seed=0, coords="helix"):
    seq, pairs = T.config0_chain()
    s = T.build_rna_system([(seq, pairs)], name="c0")
    rng = np.random.default_rng(seed)
    x = T.helix_coords(s, rng) if coords == "helix" else T.compact_coords(s, rng, box=36, min_dist=2.6)
    return s, x


def compact_allpairs(seed=1, n=40, ions=False, periodic=False, cutoff_ele=99.0):
    rng = np.random.default_rng(seed)
    seq = T.random_sequence(rng, n)
    s = T.build_rna_system([(seq, [])], all_pair_hbonds=True,
                           ions=(["Mg"] * 3 + ["K"] * 5 + ["Cl"] * 4) if ions else None,
                           box=(60.0, 60.0, 60.0) if periodic else None, cutoff_ele=cutoff_ele,
                           name=f"compact{seed}")
    x = T.compact_coords(s, rng, box=28, min_dist=2.4)
    return s, x


def hairpin_design(seed, n):
    rng = np.random.default_rng(seed)
    stem = int(rng.integers(4, 11))
    loop = int(rng.integers(3, 9))
    seq, pairs = T.hairpin_design(rng, n, stem, loop)
    s = T.build_rna_system([(seq, pairs)], name=f"hp{seed}")
    x = T.helix_coords(s, rng)
    return s, x


def small_ions_box(seed=5, n=20, box=60.0, alpha=0.2, rcut=25.0, hmax=6, form=ELE_EWALD):
    """A 20-nt chain (phosphate charge -1) with 24 ions in a 60 A periodic
    cube: Ewald (alpha 0.2, r_c 25 A, |h| <= 6) or plain Coulomb."""
    rng = np.random.default_rng(seed)
    seq = T.random_sequence(rng, n)
    ions = ["Mg"] * 2 + ["K"] * 12 + ["Cl"] * 10
    s = T.build_rna_system([(seq, [])], ions=ions, box=(box, box, box), ele_form=form, cutoff_ele=rcut,
                           ewld_alpha=alpha, ewld_hmax=hmax, phos_charge=-1.0, name="ions_box")
    x = T.compact_coords(s, rng, box=box * 0.9, min_dist=3.0)
    return s, x


def coulomb_chain(seed=21, n=160, rcut=40.0, box=60.0):
    """A compact n-nt chain (phosphate charge -1) with plain Coulomb
    electrostatics cut at 40 A, no box: 160 nt give three charge chunks of 64
    with a partial last one, 320 nt five."""
    rng = np.random.default_rng(seed)
    seq = T.random_sequence(rng, n)
    s = T.build_rna_system([(seq, [])], ele_form=ELE_COULOMB, cutoff_ele=rcut, phos_charge=-1.0,
                           name="coulomb_chain")
    x = T.compact_coords(s, rng, box=box, min_dist=2.6)
    return s, x


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bwyv19(ions=True):
    """The reference's own DT15-format fixture: ninfo/bwyv19.ninfo (28 nt,
    27 consecutive stacks, 108 hbonds of which 18 tertiary, one exclusive
    tertiary stack) on pdb/bwyv19_force_check.rna.pdb coordinates, with the
    206 Mg2+ of pdb/bwyv19_force_check.ion.pdb as a second unit, at the
    conditions of inp/bwyv19_force_check.inp (310.15 K, I = 0.057 M,
    i_DH_cutoff_type = 1, cutoff_ele = 50 A).  The files are kept under
    tests/golden/bwyv19/."""
    from tis_amd import ninfo as NI
    d = os.path.join(GOLDEN, "bwyv19")
    rec = NI.read_ninfo(os.path.join(d, "bwyv19.ninfo"))
    files = [os.path.join(d, "bwyv19_force_check.rna.pdb")]
    if ions:
        files.append(os.path.join(d, "bwyv19_force_check.ion.pdb"))
    beads = NI.read_pdb_cg(files)
    s = NI.system_from_ninfo(rec, beads, tempk=310.15, ionic_strength=0.057, dh_cutoff_type=1,
                             cutoff_ele=50.0, name="bwyv19")
    return s, beads.xyz.copy()


def tstack_hairpin(seed=61, n=48):
    """A compact hairpin with tertiary stacks: exclusive ones that switch
    off consecutive stacks on one or both sides, two that displace the same
    consecutive stack from either side, one whose partner stack does not
    exist (its side becomes non-exclusive), and one 12 A beyond its native
    distance (d - d0 >= 10: it displaces nothing)."""
    rng = np.random.default_rng(seed)
    seq, pairs = T.hairpin_design(rng, n, 6, 5)
    s = T.build_rna_system([(seq, pairs)], name="tstack_hairpin")
    x = T.compact_coords(s, rng, box=34, min_dist=2.6)
    B = lambda k: 3 * k + 1  # noqa: E731  base bead of nucleotide k (5'-S chain)
    stacks = [(B(3), B(20), +2, -2), (B(4), B(33), -2, +1), (B(12), B(40), +1, -1),
              (B(n - 2), B(25), +2, +2), (B(8), B(44), -2, +2)]
    s = T.add_tertiary_stacks(s, x, stacks, dist_shift=[0.3, -0.4, 0.2, 0.1, 12.0], rng=rng)
    return s, x


def valence_crowd(nc=72, seed=71):
    """One G whose three hbond sites each take part in nc formed hbonds (a
    G-C candidate with each of nc cytosines placed on a shell at the G-C
    native distance): the valence pass deletes nc - 1 of them, the first
    few from site sequences longer than 64.  The reference keeps at most 50
    candidates per site (MXMPHBNEIGHBOR, neighbor_list_hb.F90:124-131), so
    this case is checked against the oracle only."""
    import math
    rng = np.random.default_rng(seed)
    seq = "G" + "AAA" + "C" * nc + "A"
    s = T.build_rna_system([(seq, [(0, 4 + k) for k in range(nc)])], name="valence_crowd")
    x = np.zeros((s.nmp, 3))
    # Fibonacci sphere: consecutive nucleotides on neighbouring directions
    n = len(seq)
    ga = math.pi * (3.0 - math.sqrt(5.0))
    dirs = []
    for k in range(n):
        z = 1.0 - 2.0 * (k + 0.5) / n
        r = math.sqrt(1.0 - z * z)
        dirs.append((r * math.cos(ga * k), r * math.sin(ga * k), z))
    dirs = np.asarray(dirs)
    # sugars and phosphates off the base's radial line, so that no three
    # beads of an hbond's angles or dihedrals are collinear
    def perp(d, k):
        t = np.cross(d, (0.3, 0.5, 0.81) if k % 2 else (0.7, -0.2, 0.68))
        return t / np.linalg.norm(t)

    for k in range(n):
        sb, bb = 3 * k, 3 * k + 1
        if k == 0:
            x[bb] = 0.0
            x[sb] = -4.5 * dirs[n // 2] + 1.7 * perp(dirs[n // 2], 1)
            continue
        x[bb] = (5.655 + rng.uniform(-0.6, 0.6)) * dirs[k]
        x[sb] = (9.8 + rng.uniform(-0.3, 0.3)) * dirs[k] + 2.6 * perp(dirs[k], k)
        mid = (dirs[k] + dirs[k - 1]) / np.linalg.norm(dirs[k] + dirs[k - 1])
        x[sb - 1] = 12.5 * mid + 2.0 * perp(mid, k + 1)
    return s, x


def parity_cases(ewald=True):
    out = [("hairpin64_helix",) + hairpin64(0, "helix"),
           ("hairpin64_compact",) + hairpin64(3, "compact")]
    for k, kw in enumerate([{}, {"cutoff_ele": 2.0}, {"ions": True}, {"ions": True, "periodic": True}]):
        out.append((f"compact{k}",) + compact_allpairs(seed=10 + k, **kw))
    # several 64-charge chunks: pair chunks I < J of the electrostatic sweep
    out.append(("hairpin300_helix",) + hairpin_design(7, 300))
    # tertiary stacks: the reference's bwyv19 fixture and a generated hairpin
    out.append(("bwyv19",) + bwyv19())
    out.append(("tstack_hairpin",) + tstack_hairpin())
    out.append(("coulomb_chain160",) + coulomb_chain())
    out.append(("coulomb_chain320",) + coulomb_chain(seed=22, n=320, box=76.0))
    if ewald:
        out.append(("ions_box_ewald",) + small_ions_box())
        out.append(("ions_box_coulomb",) + small_ions_box(form=ELE_COULOMB))
        out.append(("c2_ewald_565ions",) + T.config2_ewald())
    return out


def force_error(f, fo):
    """Per-bead relative force error |dF_i| / max(|F_i|, floor), floor = 1e-3 max|F|."""
    nf = np.linalg.norm(fo, axis=1)
    floor = 1e-3 * max(nf.max(), 1e-12)
    return np.linalg.norm(f - fo, axis=1) / np.maximum(nf, floor)


def two_chain_system(seed=31):
    """Two hairpins (units 0 and 1) and a unit of ions: every off-diagonal
    unit pair has exv and DH energies."""
    rng = np.random.default_rng(seed)
    chains = [T.hairpin_design(rng, 24, 5, 4), T.hairpin_design(rng, 18, 4, 4)]
    s = T.build_rna_system(chains, ions=["Mg"] * 2 + ["K"] * 3 + ["Cl"] * 3, name="two_chains")
    x = T.compact_coords(s, rng, box=34, min_dist=2.9)
    return s, x


def unit_cases(cases=None):
    """Systems of several units for energy_unit: two chains + ions, the
    ion-rich compact chains (one periodic) and an Ewald box."""
    cases = parity_cases(ewald=False) if cases is None else cases
    return ([("two_chains",) + two_chain_system()]
            + [c for c in cases if c[0] in ("compact2", "compact3", "bwyv19")]
            + [("ions_box_ewald",) + small_ions_box()])

