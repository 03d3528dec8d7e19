"""Pins of the oracle's Brownian-dynamics-with-HI functions: the reference's
own simu_hydro_tensors / force_fene / force_wca (where oracle/_ref builds),
continuity of the Zuk et al. RPY tensor at its branch points, the Oseen far
field, Cholesky against LAPACK, and the Stokes drift of the step."""
import math

import numpy as np
import pytest

import oracle
from oracle import RefLib
from tis_amd.bdhi import (HI_ERMAK_OVER, HI_RPY, HI_RPY_OVER, HI_ZUK_RPY, BDSystem, bead_spring_chain,
                          stokes_diffusion)
from tis_amd.topology import langevin_visc

needs_ref = pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (reference absent)")


def two_beads(r, ai=3.5, aj=3.5, htype=HI_ZUK_RPY):
    s = BDSystem(radius=[ai, aj], fene_mp=np.zeros((0, 2)), fene_nat=[], fene_coef=[], fene_dist2=[],
                 wca_mp=np.zeros((0, 2)), wca_nat2=[], wca_coef=np.zeros((0, 2)), cutoff_wca=1.0,
                 tempk=300.0, viscosity=langevin_visc(1e-5), hydro_tensor=htype)
    x = np.zeros((2, 3))
    x[1] = (r * 0.6, r * 0.8, 0.0)
    return s, x


def test_zuk_tensor_is_continuous_at_its_branch_points():
    for ai, aj in ((3.5, 3.5), (2.0, 5.0)):
        for r in (ai + aj, abs(aj - ai)):
            if r == 0.0:
                continue
            s, x_in = two_beads(r * (1 - 1e-9), ai, aj)
            _, x_out = two_beads(r * (1 + 1e-9), ai, aj)
            Din = oracle.rpy_tensor(s, x_in)
            Dout = oracle.rpy_tensor(s, x_out)
            assert np.abs(Din - Dout).max() <= 1e-7 * np.abs(Din).max()


def test_oseen_far_field():
    r = 2000.0
    s, x = two_beads(r)
    D = oracle.rpy_tensor(s, x)
    from tis_amd.topology import BOLTZ_KCAL_MOL
    kt_8pi_eta_r = BOLTZ_KCAL_MOL * s.tempk / (8 * math.pi * s.viscosity * r)
    rhat = x[1] / np.linalg.norm(x[1])
    oseen = kt_8pi_eta_r * (np.eye(3) + np.outer(rhat, rhat))
    assert np.abs(D[0:3, 3:6] - oseen).max() <= 1e-5 * np.abs(oseen).max()
    assert abs(D[0, 0] - stokes_diffusion(s)) <= 1e-15 * D[0, 0]
    np.testing.assert_array_equal(D, D.T)


@pytest.mark.parametrize("htype", [HI_ZUK_RPY, HI_RPY, HI_RPY_OVER, HI_ERMAK_OVER])
def test_cholesky_matches_lapack(htype):
    # the Ermak overlap form is not positive definite for touching beads
    # (the reference stops there too): keep that case's beads apart
    s, x = bead_spring_chain(30, seed=2, hydro_tensor=htype, bond=8.0 if htype == HI_ERMAK_OVER else 7.0)
    D = oracle.rpy_tensor(s, x)
    L = oracle.cholesky(D)
    Lr = np.linalg.cholesky(D)
    assert np.abs(L - Lr).max() <= 1e-12 * np.abs(Lr).max()
    assert np.abs(np.triu(L, 1)).max() == 0.0


def test_step_drift_is_stokes_mobility():
    """Same noise twice: the difference of a step with and without a force
    is (dt / kT) D F; for a lone bead D F = F kT / (6 pi eta a)."""
    s, x = two_beads(5000.0)
    s.fene_mp = np.array([[0, 1]], np.int32)
    s.fene_nat = np.array([4999.0]); s.fene_coef = np.array([0.3]); s.fene_dist2 = np.array([1e8])
    s0, _ = two_beads(5000.0)
    dt = 0.05
    xa = oracle.bdhi_run(s, x, 1, tstep=dt, seed=4)
    xb = oracle.bdhi_run(s0, x, 1, tstep=dt, seed=4)
    f = oracle.bd_forces(s, x)
    from tis_amd.topology import BOLTZ_KCAL_MOL
    want = dt / (BOLTZ_KCAL_MOL * s.tempk) * stokes_diffusion(s) * f[0]
    got = (xa - xb)[0]
    assert np.abs(got - want).max() <= 2e-3 * np.abs(want).max()   # far-field coupling ~ a/r


@needs_ref
@pytest.mark.parametrize("htype", [HI_ZUK_RPY, HI_RPY, HI_RPY_OVER, HI_ERMAK_OVER])
def test_tensor_and_factor_match_reference(htype):
    s, x = bead_spring_chain(25, seed=3, hydro_tensor=htype, bond=8.0 if htype == HI_ERMAK_OVER else 7.0)
    if htype == HI_ZUK_RPY:
        s.radius = np.linspace(2.5, 4.5, s.nmp)   # unequal radii and overlaps
        x[5] = x[4] + (1.0, 0.5, 0.0)
    D = oracle.rpy_tensor(s, x)
    L = oracle.cholesky(D)
    Dr, Lr = RefLib().hydro(s, x)
    assert np.abs(D - Dr).max() <= 1e-14 * np.abs(Dr).max()
    assert np.abs(L - Lr).max() <= 1e-12 * np.abs(Lr).max()


@needs_ref
def test_fene_wca_match_reference():
    s, x = bead_spring_chain(60, seed=5)
    x += np.random.default_rng(1).normal(scale=0.8, size=x.shape)   # stretch bonds, press contacts
    f = oracle.bd_forces(s, x)
    fr = RefLib().bd_forces(s, x)
    assert np.abs(f - fr).max() <= 1e-12 * np.abs(fr).max()
    assert np.abs(f.sum(axis=0)).max() <= 1e-9 * np.abs(f).max()
