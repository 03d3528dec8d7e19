"""GPU parity of the Brownian-dynamics-with-HI path (through the C-ABI)
against the FP64 oracle: tensor fill, blocked Cholesky with FP32
matrix-core trailing updates, FENE/WCA forces, the Ermak-McCammon step,
and the statistics of the correlated noise."""
import numpy as np
import pytest

import oracle
from tis_amd.bdhi import HI_ERMAK_OVER, HI_RPY, HI_RPY_OVER, HI_ZUK_RPY, bead_spring_chain
from tis_amd.engine import BDEngine

pytestmark = pytest.mark.gpu


def _chain(n, htype=HI_ZUK_RPY, seed=3, unequal=False):
    s, x = bead_spring_chain(n, seed=seed, hydro_tensor=htype, bond=8.0 if htype == HI_ERMAK_OVER else 7.0)
    if unequal:
        s.radius = np.linspace(2.5, 4.5, s.nmp)
        x[5] = x[4] + (1.0, 0.5, 0.0)
    return s, x


@pytest.mark.parametrize("htype", [HI_ZUK_RPY, HI_RPY, HI_RPY_OVER, HI_ERMAK_OVER])
def test_tensor_fill_matches_oracle(htype):
    s, x = _chain(41, htype, unequal=htype == HI_ZUK_RPY)
    print(f"[bdhi fill] type {htype} n={s.nmp}", flush=True)
    e = BDEngine(s, key=1, seed=2)
    e.set_coords(x)
    e.tensors()
    D = e.tensor(0)
    Do = oracle.rpy_tensor(s, x)
    assert np.abs(D - Do).max() <= 1e-14 * np.abs(Do).max()
    e.close()


@pytest.mark.parametrize("n", [21, 64, 300, 1000, 3000])
def test_factor_against_fp64(n):
    """L from FP64 diagonal blocks + FP32 matrix-core trailing updates against
    the FP64 factor; ||L L^T - D|| / ||D|| on the device equals the host value."""
    s, x = _chain(n)
    print(f"[bdhi factor] n={n} (3n={3 * n})", flush=True)
    e = BDEngine(s, key=1, seed=2)
    e.set_coords(x)
    e.tensors()
    D = e.tensor(0)
    L = e.tensor(1)
    L64 = np.linalg.cholesky(D)
    rel_l = np.abs(L - L64).max() / np.abs(L64).max()
    err_host = np.linalg.norm(L @ L.T - D) / np.linalg.norm(D)
    err_dev = e.factor_error()
    print(f"   max|L-L64|/max|L64| {rel_l:.3e}  ||LL^T-D||/||D|| host {err_host:.3e} device {err_dev:.3e}",
          flush=True)
    assert np.abs(np.triu(L, 1)).max() == 0.0
    assert rel_l <= 2e-5
    assert err_host <= 1e-6
    assert abs(err_dev - err_host) <= 1e-3 * err_host + 1e-15
    e.close()


def test_forces_match_oracle():
    s, x = _chain(120, seed=5)
    x = x + np.random.default_rng(1).normal(scale=0.8, size=x.shape)
    e = BDEngine(s, key=1, seed=2)
    e.set_coords(x)
    f = e.forces()
    fo = oracle.bd_forces(s, x)
    assert np.abs(f - fo).max() <= 1e-12 * np.abs(fo).max()
    e.close()


@pytest.mark.parametrize("refactor", [1, 5])
def test_steps_match_oracle(refactor):
    s, x = _chain(40, seed=7)
    print(f"[bdhi steps] refactor every {refactor}", flush=True)
    e = BDEngine(s, key=9, seed=11)
    e.set_coords(x)
    e.steps(20, tstep=0.05, refactor_every=refactor, step0=100)
    xg = e.coords()
    xo = oracle.bdhi_run(s, x, 20, tstep=0.05, refactor_every=refactor, seed=11, key=9, step0=100)
    moved = np.abs(xo - x).max()
    err = np.abs(xg - xo).max()
    print(f"   moved {moved:.3f} A, |x_gpu - x_fp64| {err:.3e} A", flush=True)
    assert moved > 0.1
    assert err <= 1e-5
    e.close()


def test_correlated_noise_statistics():
    """Noise L xi with the device factor: sample covariance matches D (FP64),
    and the whitened noise L64^-1 L xi is standard normal (KS)."""
    from scipy import stats
    s, x = _chain(30, seed=4)
    e = BDEngine(s, key=1, seed=2)
    e.set_coords(x)
    e.tensors()
    D = e.tensor(0)
    L = e.tensor(1)
    n3 = D.shape[0]
    xi = np.stack([np.concatenate([oracle.gauss3(5, 3, t, b) for b in range(s.nmp)]) for t in range(20000)])
    y = xi @ L.T
    cov = y.T @ y / len(y)
    rel = np.linalg.norm(cov - D) / np.linalg.norm(D)
    w = np.linalg.solve(np.linalg.cholesky(D), y.T).ravel()
    ks = stats.kstest(w, "norm")
    print(f"   cov rel err {rel:.3e} (20000 samples, {n3} dims), KS p {ks.pvalue:.3f}", flush=True)
    assert rel < 0.05
    assert ks.pvalue > 1e-3
    e.close()
