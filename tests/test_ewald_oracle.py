"""Pins of the oracle's plain-Coulomb and Ewald terms: the reference's own
Fortran routines (where oracle/_ref builds), the rock-salt Madelung constant,
the reciprocal-vector count, and force = -grad(energy)."""
import math

import numpy as np
import pytest

from oracle import Oracle, RefLib
from tis_amd import topology as T
from tis_amd.sysdesc import E_ELE, ELE_COULOMB, ELE_EWALD

from systems import small_ions_box

needs_ref = pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (reference absent)")

MADELUNG_NACL = 1.747564594633182  # rock salt, per ion pair, in units of q^2 / r0


def small_ewald(form=ELE_EWALD):
    return small_ions_box(form=form)


def test_kvec_count_and_coefficients():
    s, _ = small_ewald()
    o = Oracle(s)
    for hmax, want in ((6, None), (10, 4168)):
        s.params.ewld_hmax = hmax
        o = Oracle(s)
        h, c = o.ewald_kvec()
        g = np.arange(-hmax, hmax + 1)
        hx, hy, hz = np.meshgrid(g, g, g, indexing="ij")
        r2 = hx ** 2 + hy ** 2 + hz ** 2
        count = int(np.count_nonzero((r2 <= hmax * hmax) & (r2 > 0)))
        assert len(h) == count
        if want is not None:
            assert count == want
        h2 = (h ** 2).sum(axis=1)
        assert np.all(np.diff(h2) >= 0)                  # ordered by |h|
        assert len({tuple(v) for v in h}) == count        # no duplicates, +h and -h both present
        L, a = s.params.box[0], s.params.ewld_alpha
        k2 = (2 * math.pi / L) ** 2 * h2
        np.testing.assert_allclose(c, 2 * math.pi / L ** 3 * np.exp(-k2 / (4 * a * a)) / k2, rtol=1e-14)


def rock_salt(ncell=8, a0=4.0):
    """ncell^3 alternating Na+/Cl- ions on a simple cubic lattice of spacing a0."""
    g = np.arange(ncell)
    ix, iy, iz = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1) * a0
    sign = (ix + iy + iz).ravel() % 2
    ions = ["Na" if v == 0 else "Cl" for v in sign]
    L = ncell * a0
    s = T.build_rna_system([], ions=ions, box=(L, L, L), ele_form=ELE_EWALD, cutoff_ele=L / 2 - 0.01,
                           ewld_alpha=4.0 / (L / 2 - 0.01), ewld_hmax=11, name="rocksalt")
    return s, pos - L / 2 + a0 / 2


def test_madelung_rock_salt():
    """E = -(N/2) M C q^2 / r0 for the rock-salt lattice (Fourier + real + self)."""
    s, x = rock_salt()
    o = Oracle(s)
    o.neighbor(x)
    e = o.energy(x)[E_ELE]
    n = s.nmp
    want = -0.5 * n * MADELUNG_NACL * s.params.ele_coef / 4.0
    assert abs(e - want) <= 2e-7 * abs(want)
    # a plausible slip (the self term dropped, or its sign flipped) is far off
    parts = o.ewald_energy_parts(x)
    assert abs(parts.sum() - e) <= 1e-9 * abs(e)
    assert abs((parts[0] + parts[1] - parts[2]) - want) > 1e-2 * abs(want)


def _fd_force(o, x, eps=1e-5, beads=None):
    f = np.zeros_like(x)
    for i in beads:
        for d in range(3):
            e = []
            for sgn in (1, -1):
                xs = x.copy()
                xs[i, d] += sgn * eps
                o.neighbor(xs)   # periodic terms read the wrapped copy made here
                e.append(o.energy(xs)[E_ELE])
            f[i, d] = -(e[0] - e[1]) / (2 * eps)
    o.neighbor(x)
    return f


@pytest.mark.parametrize("form", [ELE_EWALD, ELE_COULOMB])
def test_force_is_minus_gradient(form):
    s, x = small_ewald(form=form)
    o = Oracle(s)
    o.neighbor(x)
    # keep every listed pair clear of the cutoff (the real-space energy steps there)
    pairs, _ = o.lists(1)
    L = s.params.box[0]
    d = x[pairs[:, 1]] - x[pairs[:, 0]]
    d -= L * np.rint(d / L)
    r = np.linalg.norm(d, axis=1)
    assert np.min(np.abs(r - s.params.cutoff_ele)) > 1e-3
    beads = np.nonzero(np.abs(s.charge) > 0)[0][::3]
    name = "ewald" if form == ELE_EWALD else "coulomb"
    f = o.term(name, x)
    fd = _fd_force(o, x, beads=beads)
    err = np.abs(f[beads] - fd[beads]).max() / np.abs(f[beads]).max()
    assert err < 1e-6
    assert np.abs(f.sum(axis=0)).max() <= 1e-9 * np.abs(f).max()   # no net force on the box


@needs_ref
@pytest.mark.parametrize("form", [ELE_EWALD, ELE_COULOMB])
def test_matches_reference(form):
    s, x = small_ewald(form=form)
    o = Oracle(s)
    o.neighbor(x)
    R = RefLib()
    R.load(s)
    R.neighbor(x)
    pr, cr = R.lists(1)
    po, co = o.lists(1)
    key = lambda p: np.lexsort((p[:, 1], p[:, 0]))
    assert np.array_equal(pr[key(pr)], po[key(po)])
    name = "ewald" if form == ELE_EWALD else "coulomb"
    fr = R.force_term(name)
    fo = o.term(name, x)
    assert np.abs(fr - fo).max() <= 1e-11 * np.abs(fr).max()
    er = R.energy()[E_ELE]
    eo = o.energy(x)[E_ELE]
    assert abs(er - eo) <= 1e-11 * abs(er)


@needs_ref
def test_matches_reference_full_box():
    """The 76-nt chain with 565 ions in the 150 A cube (alpha 0.1, r_c 30, |h| <= 10)."""
    s, x = T.config2_ewald()
    o = Oracle(s)
    o.neighbor(x)
    R = RefLib()
    R.load(s)
    R.neighbor(x)
    fr = R.force_term("ewald")
    fo = o.term("ewald", x)
    assert np.abs(fr - fo).max() <= 1e-11 * np.abs(fr).max()
    assert abs(R.energy()[E_ELE] - o.energy(x)[E_ELE]) <= 1e-11 * abs(o.energy(x)[E_ELE])
