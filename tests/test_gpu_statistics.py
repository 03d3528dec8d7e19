"""Statistical trajectory parity: GPU Langevin trajectories against FP64
oracle trajectories run with a different seed (the random sequences differ,
so trajectories are compared by distributions): kinetic temperature, radius
of gyration, stack and hbond energies.  And the integrator's kinetic
temperature against the reference's own temperature check."""
import numpy as np
import pytest

import oracle
from tis_amd import topology as T
from tis_amd.engine import Engine, MODE_GRAPH
from tis_amd.sysdesc import E_HBOND, E_STACK

pytestmark = pytest.mark.gpu

KB = T.BOLTZ_KCAL_MOL


def _observables(s, x, v, e):
    m = s.mass[:, None]
    tk = float((m * v * v).sum() / (3 * s.nmp * KB))
    rg = float(np.sqrt(((x - x.mean(axis=0)) ** 2).sum(axis=1).mean()))
    return tk, rg, e[E_STACK], e[E_HBOND]


def test_trajectory_statistics_match_oracle():
    from scipy import stats
    nrep, burn, chunk, nsamp = 32, 2000, 200, 40
    systems, x0 = [], []
    for k in range(nrep):
        rng = np.random.default_rng(40 + k)
        seq, pairs = T.hairpin_design(rng, 30, 6, 4)
        s = T.build_rna_system([(seq, pairs)], tempk=310.0)
        systems.append(s)
        x0.append(T.helix_coords(s, rng))
    print(f"[stats] {nrep} x {systems[0].nmp} beads, {burn} + {chunk}x{nsamp} steps", flush=True)
    eng = Engine(systems, keys=list(range(nrep)), seed=101)
    for k, x in enumerate(x0):
        eng.set_coords(k, x)
    eng.langevin_init(tstep=0.05, rebuild_every=10)
    eng.langevin_steps(burn, MODE_GRAPH)
    gpu = [[] for _ in range(nrep)]
    for _ in range(nsamp):
        eng.langevin_steps(chunk, MODE_GRAPH)
        e = eng.energy_all()
        for k in range(nrep):
            gpu[k].append(_observables(systems[k], eng.coords(k), eng.get_state(k, 1), e[k]))
    ora = [[] for _ in range(nrep)]
    for k, s in enumerate(systems):
        o = oracle.Oracle(s)
        x, v, a = x0[k], np.zeros_like(x0[k]), np.zeros_like(x0[k])
        x, v, a, _ = o.langevin(x, v, a, burn, tstep=0.05, rebuild_every=10, seed=202, sys_id=k, step0=0)
        t = burn
        for _ in range(nsamp):
            x, v, a, _ = o.langevin(x, v, a, chunk, tstep=0.05, rebuild_every=10, seed=202, sys_id=k, step0=t,
                                    init_accel=False)
            t += chunk
            o.neighbor(x)
            o.force(x, seed=202, sys_id=k, step=t)
            ora[k].append(_observables(s, x, v, o.energy(x, v)))
    g = np.asarray(gpu)   # (nrep, nsamp, 4)
    o = np.asarray(ora)
    names = ["T_kin", "Rg", "E_stack", "E_hbond"]
    # samples of one replica are correlated in time; the replicas are
    # independent, so the tests run on one time average per replica
    for q, name in enumerate(names):
        gm, om = g[:, :, q].mean(axis=1), o[:, :, q].mean(axis=1)
        se = np.sqrt(gm.var(ddof=1) / nrep + om.var(ddof=1) / nrep)
        ks = stats.ks_2samp(gm, om)
        print(f"   {name}: gpu {gm.mean():.4f} oracle {om.mean():.4f} (se {se:.4f}), KS p {ks.pvalue:.3g}",
              flush=True)
        assert abs(gm.mean() - om.mean()) <= 4.0 * se + 1e-12
        assert ks.pvalue > 1e-3


def test_kinetic_temperature_of_the_integrator():
    """<T_kin> of the Langevin integrator at T = 300 K, eta = 1e-5 Pa s for a
    4-nt DT15 chain, dt = 0.05 and 0.01 tau.  The reference's own DT15 check
    (integrator_test/Temperature_check, 'azo' at 310 K, dt 0.05) reports
    309.2-309.7 K, within 0.3 % of the set temperature; its DT13 'uuuu'
    runs (284 K at dt 0.05) use a model that is not on this path."""
    ncopy, burn_tau, run_tau = 512, 300.0, 600.0
    out = {}
    for dt in (0.05, 0.01):
        burn, steps = int(burn_tau / dt), int(run_tau / dt)
        systems = []
        for k in range(ncopy):
            s = T.build_rna_system([("UUUU", [])], tempk=300.0, ionic_strength=0.1)
            systems.append(s)
        eng = Engine(systems, keys=list(range(ncopy)), seed=7)
        rng = np.random.default_rng(3)
        for k, s in enumerate(systems):
            eng.set_coords(k, T.helix_coords(s, rng))
        eng.langevin_init(tstep=dt, rebuild_every=10)
        eng.langevin_steps(burn, MODE_GRAPH)
        tks = []
        for _ in range(20):
            eng.langevin_steps(steps // 20, MODE_GRAPH)
            e = eng.energy_all()
            tks.append(np.mean([2.0 * e[k, 1] / (3 * systems[k].nmp * KB) for k in range(ncopy)]))
        out[dt] = float(np.mean(tks))
        eng.close()
        print(f"   dt {dt}: <T_kin> = {out[dt]:.2f} K", flush=True)
    for dt, tk in out.items():
        assert abs(tk - 300.0) < 3.0, (dt, tk)
