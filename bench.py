"""Benchmark: batched DT15 Langevin bead-steps/s on one MI355X.

Workload (the batched configuration): 512 independent DT15 hairpin-design
chains with lengths lognormal(mu=ln 120, sigma=0.7) clipped to 20-600 nt, plus
64 replicas of the 64-nt hairpin on a geometric 290-400 K ladder; FP32 pair
arithmetic with FP64 accumulation and integration; Langevin dt = 0.05 tau,
neighbour lists rebuilt every 10 steps; all systems advance in one
hipGraph-captured launch sequence per step.  Synthetic seeded sequences and
extended-helix starting coordinates (no external data).

Prints one JSON line (see the contract in the repository documentation):
value = whole-batch bead-steps/s; beside it the plain one-chain-per-launch
GPU path on the same workload, the force error against the FP64 oracle on a
sample of chains, roofline of the dominant kernel, and a CPU baseline (the
oracle timed on this host's cores on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tis_amd import topology as T  # noqa: E402
from tis_amd.engine import Engine, MODE_EAGER, MODE_GRAPH, MODE_PLAIN  # noqa: E402

HBM_PEAK_GBS = 8000.0
SEED = 20261202


def build_workload(nchains=512, nrep=64, seed=SEED, small=False):
    rng = np.random.default_rng(seed)
    lengths = T.config1_lengths(seed, nchains)
    systems, coords, labels = [], [], []
    for k, n in enumerate(lengths):
        n = int(n)
        stem = int(rng.integers(4, 11))
        loop = int(rng.integers(3, 9))
        while 2 * stem + loop + 2 > n:
            stem = max(2, stem - 1)
            loop = max(3, loop - 1)
        seq, pairs = T.hairpin_design(rng, n, stem, loop)
        s = T.build_rna_system([(seq, pairs)], tempk=310.0, ionic_strength=0.150, name=f"chain{k}")
        systems.append(s)
        coords.append(T.helix_coords(s, rng))
        labels.append("chain")
    seq0, pairs0 = T.config0_chain(seed)
    for r in range(nrep):
        tk = 290.0 * (400.0 / 290.0) ** (r / max(nrep - 1, 1))
        s = T.build_rna_system([(seq0, pairs0)], tempk=tk, ionic_strength=0.150, name=f"rep{r}")
        systems.append(s)
        coords.append(T.helix_coords(s, rng))
        labels.append("replica")
    return systems, coords, labels


def algorithmic_bytes(systems, eng):
    """Algorithmic HBM bytes of one step, split by kernel.  k_main: per bead
    169 (read x, v, a and the Langevin class byte; write F, v, a and the next
    x), per exv list entry 8.  k_force: per DH list entry 12, per term bond
    16, angle 20, stack 56, hbond and tertiary stack 76 (counts at the
    current lists)."""
    nb = sum(s.nmp for s in systems)
    nexv = nele = 0
    for k, s in enumerate(systems):
        nexv += len(eng.pairs(k, 0))
        nele += len(eng.pairs(k, 1)[0])
    nbd = sum(len(s.bd_nat) for s in systems)
    nba = sum(len(s.ba_nat) for s in systems)
    nst = sum(len(s.st_mp) for s in systems)
    nhb = sum(len(s.hb_mp) for s in systems)
    ntst = sum(len(s.tst_mp) for s in systems)
    main = nb * 169 + 8 * nexv
    force = 12 * nele + 16 * nbd + 20 * nba + 56 * nst + 76 * (nhb + ntst)
    return {"beads": nb, "exv_pairs": nexv, "ele_pairs": nele, "main_bytes": main, "force_bytes": force}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "dt15_batched_step_hbm_bytes.md")


def pmc_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated against a copy of known
    size); None when the summary or the row is absent."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 6 and cells[0] == kernel:
                    return int(cells[5])
    except OSError:
        return None
    return None


def force_check(systems, coords_now, eng, sample, energies=None):
    """Largest per-bead force error against the FP64 oracle over the sample
    (relative, floor 1e-3 max|F|) and, with the GPU energies of the same
    state, the largest per-term energy error (relative, floor 1e-6)."""
    import oracle
    worst = worst_e = 0.0
    for k in sample:
        s = systems[k]
        x = coords_now[k]
        o = oracle.Oracle(s)
        o.neighbor(x)
        fo, _ = o.force(x, seed=0, sys_id=k, step=0)
        fg = eng.get_state(k, 3)
        nf = np.linalg.norm(fo, axis=1)
        floor = 1e-3 * max(nf.max(), 1e-12)
        err = np.linalg.norm(fg - fo, axis=1) / np.maximum(nf, floor)
        worst = max(worst, float(err.max()))
        if energies is not None:
            eo = o.energy(x, eng.get_state(k, 1))
            worst_e = max(worst_e, float((np.abs(energies[k] - eo) / np.maximum(np.abs(eo), 1e-6)).max()))
    return worst, worst_e


def serving_stream(njobs=100, scale=0.01, inflight=5.0, round_steps=500, seed=SEED + 4):
    """Serving stream: jobs arrive as a Poisson process; each is a seeded
    hairpin-design chain, lognormal(ln 200, 1.0) clipped to 20-2000 nt, run
    for U[1e4, 1e6] x scale Langevin steps (a multiple of 10) at 150 mM and
    310 K.  The arrival rate is set from a calibration run so that about
    `inflight` jobs are on the GPU on average (Little's law)."""
    from tis_amd.serve import Job, Server, poisson_stream
    rng = np.random.default_rng(seed)

    def make(k, r):
        n = int(np.clip(np.round(r.lognormal(math.log(200.0), 1.0)), 20, 2000))
        stem = int(r.integers(4, 11))
        loop = int(r.integers(3, 9))
        while 2 * stem + loop + 2 > n:
            stem, loop = max(2, stem - 1), max(3, loop - 1)
        seq, pairs = T.hairpin_design(r, n, stem, loop)
        s = T.build_rna_system([(seq, pairs)], tempk=310.0, ionic_strength=0.150, name=f"job{k}")
        steps = int(r.uniform(1e4, 1e6) * scale) // 10 * 10
        return Job(key=100000 + k, system=s, x0=T.helix_coords(s, r), nsteps=max(steps, 10))

    # calibration: ms/step of a batch of `inflight` jobs of the same distribution
    cal = [make(k, np.random.default_rng(seed + 1 + k)) for k in range(int(round(inflight)))]
    eng = Engine([j.system for j in cal], keys=[j.key for j in cal], seed=SEED)
    for k, j in enumerate(cal):
        eng.set_coords(k, j.x0)
    eng.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
    eng.langevin_steps(100, MODE_GRAPH)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.langevin_steps(500, MODE_GRAPH)
    eng.synchronize()
    step_s = (time.perf_counter() - t0) / 500
    eng.close()
    mean_steps = 0.5 * (1e4 + 1e6) * scale
    rate = inflight / (mean_steps * step_s)
    jobs = poisson_stream(rng, njobs, rate, make)
    srv = Server(hbm_budget=64 << 30, round_steps=round_steps, clock="wall", seed=SEED)
    st = srv.run(jobs)
    out = st.summary()
    out.update({"arrival_rate_per_s": rate, "calibration_ms_per_step": step_s * 1e3, "round_steps": round_steps,
                "steps_scale": scale,
                "workload": f"{njobs} Poisson jobs, lognormal(ln200,1.0) 20-2000 nt hairpin designs, "
                            f"U[1e4,1e6]x{scale:g} steps, 150 mM, 310 K, target {inflight:g} in flight"})
    return out


def valence_leg(nsys=64, n=40, steps=200, seed=SEED + 7):
    """Chains whose hbond candidates can exceed their sites' valence (every
    base pair a candidate, compact seeded coordinates in a 28 A box): the
    batched Langevin step with the valence-limit pass of
    force_dtrna_hbond15.F90:373-510 at work, and its deletions per step."""
    rng = np.random.default_rng(seed)
    systems, coords = [], []
    for k in range(nsys):
        seq = T.random_sequence(rng, n)
        s = T.build_rna_system([(seq, [])], all_pair_hbonds=True, name=f"valence{k}")
        systems.append(s)
        coords.append(T.compact_coords(s, rng, box=28, min_dist=2.4))
    eng = Engine(systems, keys=list(range(nsys)), seed=seed)
    for k, x in enumerate(coords):
        eng.set_coords(k, x)
    eng.langevin_init(tstep=0.05, rebuild_every=10)
    eng.langevin_steps(20, MODE_GRAPH)
    eng.synchronize()
    c0 = eng.counters()
    t0 = time.perf_counter()
    eng.langevin_steps(steps, MODE_GRAPH)
    eng.synchronize()
    dt = time.perf_counter() - t0
    c1 = eng.counters()
    nb = sum(s.nmp for s in systems)
    finite = all(bool(np.isfinite(eng.coords(k)).all()) for k in range(nsys))
    eng.close()
    return {"workload": f"{nsys} compact {n}-nt chains, every base pair an hbond candidate (valence-limited)",
            "beads": nb, "ms_per_step": dt * 1e3 / steps, "bead_steps_per_s": nb * steps / dt,
            "hb_deleted_per_step": (c1["hb_deleted"] - c0["hb_deleted"]) / steps, "finite": finite}


def ewald_leg(steps=400, warmup=40):
    """Explicit ions with Ewald: the 76-nt chain + 565 ions in a 150 A cube
    (alpha 0.1, r_c 30 A, |h| <= 10), Langevin in hipGraph mode."""
    import oracle
    s, x = T.config2_ewald()
    eng = Engine([s], keys=[0], seed=SEED)
    eng.set_coords(0, x)
    eng.langevin_init(tstep=0.05, rebuild_every=10, step0=0, init_accel=True)
    eng.langevin_steps(warmup, MODE_GRAPH)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.langevin_steps(steps, MODE_GRAPH)
    eng.synchronize()
    dt = time.perf_counter() - t0
    eng.langevin_steps(50, MODE_EAGER)
    tm = eng.last_timing()
    eng.neighbor(0)
    fg = eng.force_sumup(0)
    xn = eng.coords(0)
    o = oracle.Oracle(s)
    o.neighbor(xn)
    fo, _ = o.force(xn, seed=0, sys_id=0, step=0)
    nf = np.linalg.norm(fo, axis=1)
    err = float((np.linalg.norm(fg - fo, axis=1) / np.maximum(nf, 1e-3 * nf.max())).max())
    eo = o.energy(xn)[7]
    eg = eng.energy_sumup(0)[7]
    eng.close()
    return {"workload": "76-nt chain + 10 Mg2+ 305 K+ 250 Cl- in a 150 A periodic cube; Ewald alpha 0.1, "
                        "r_c 30 A, |h| <= 10 (4168 vectors); Langevin dt 0.05 tau",
            "particles": s.nmp, "ms_per_step": dt * 1e3 / steps, "bead_steps_per_s": s.nmp * steps / dt,
            "force_max_rel_err_vs_fp64": err, "ele_energy_rel_err_vs_fp64": abs(eg - eo) / abs(eo),
            "eager_pre_ms": tm["pre_ms"], "eager_main_ms": tm["main_ms"]}


def bdhi_leg(sizes=(300, 1000, 3000), steps=50, refactor=10):
    """Brownian dynamics with HI: bead-spring chains (radius 3.5 A, FENE +
    WCA), RPY (Zuk et al.) tensor refactored every `refactor` steps."""
    from tis_amd.bdhi import bead_spring_chain
    from tis_amd.engine import BDEngine
    out = []
    for n in sizes:
        s, x = bead_spring_chain(n, seed=SEED % 1000 + n)
        e = BDEngine(s, key=n, seed=SEED)
        e.set_coords(x)
        e.tensors()                       # capture + first run
        e.tensors()
        tt = e.timing()
        rel = e.factor_error()
        n3 = 3 * n
        flop = n3 ** 3 / 3.0
        rec = {"beads": n, "dim": n3, "fill_ms": tt["fill_ms"], "factor_ms": tt["factor_ms"],
               "factor_tflops": flop / (tt["factor_ms"] * 1e-3) / 1e12, "llt_rel_err": rel}
        D = e.tensor(0)
        L = e.tensor(1)
        L64 = np.linalg.cholesky(D)
        rec["l_rel_err_vs_fp64"] = float(np.abs(L - L64).max() / np.abs(L64).max())
        del D, L, L64
        e.steps(2, tstep=0.05, refactor_every=refactor)
        t0 = time.perf_counter()
        e.steps(steps, tstep=0.05, refactor_every=refactor)
        dt = time.perf_counter() - t0
        rec.update({"steps": steps, "refactor_every": refactor, "ms_per_step": dt * 1e3 / steps,
                    "bead_steps_per_s": n * steps / dt})
        e.close()
        out.append(rec)
        print(f"[bench] bdhi n={n}: fill {rec['fill_ms']:.3f} ms, factor {rec['factor_ms']:.3f} ms "
              f"({rec['factor_tflops']:.2f} TFLOP/s), ||LL^T-D||/||D|| {rel:.2e}, "
              f"{rec['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    return out


def cpu_baseline_reference(systems, coords, budget_s=15.0, share=(0, 1)):
    """The reference's own Fortran routines (oracle/_ref, compiled from its
    sources) on this host, one thread, over chains of the same workload until
    budget_s: per step the terms force_sumup evaluates for these chains
    (force_bond, force_bangle, force_dtrna_stack_nlocal + force_dtrna_stack,
    force_dtrna_hbond15, force_exv_dt15, force_ele_DH) and neighbor() every
    10 steps, at the sample's coordinates (the O(N) Langevin update is not
    counted).  share = (k, m): this process takes every m-th system of the
    sample from the k-th.  None when the library is absent."""
    import oracle
    if not oracle.RefLib.available():
        return None
    R = oracle.RefLib()
    steps = 300
    terms = [R.TERMS[t] for t in ("bond", "angle", "stack", "hbond", "exv", "dh")]
    order = list(range(0, len(systems), max(1, len(systems) // 64)))[share[0]::share[1]]
    res = {}

    def run():
        # one thread with a large stack for the whole loop (the reference's
        # list builders keep pair buffers on the stack)
        t0 = time.perf_counter()
        bead_steps = nch = 0
        for k in order:
            s = systems[k]
            x = np.ascontiguousarray(coords[k], dtype=np.float64)
            f = np.zeros((s.nmp, 3))
            R.load(s)
            for t in range(steps):
                if t % 10 == 0:
                    R.L.tisref_neighbor(oracle.fp(x))
                for term in terms:
                    R.L.tisref_force_term(term, oracle.fp(f))
            bead_steps += s.nmp * steps
            nch += 1
            if time.perf_counter() - t0 > budget_s:
                break
        res.update(dt=time.perf_counter() - t0, bead_steps=bead_steps, nch=nch)

    # the reference writes a warning line to standard output for each force
    # clamp: send its output to /dev/null while it runs (this script's
    # standard output carries the one JSON line)
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        oracle._on_big_stack(run)
        R.L.tisref_flush()
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)
    return {"value": res["bead_steps"] / res["dt"], "unit": "bead-steps/s", "cores": 1, "kind": "reference",
            "bead_steps": res["bead_steps"], "seconds": res["dt"], "systems": res["nch"],
            "sample": f"{res['nch']} systems of the workload x {steps} steps of the reference's force routines "
                      f"and neighbor() every 10 steps (oracle/_ref, 1 thread, fixed coordinates), {res['dt']:.1f} s"}


def cpu_baseline_reference_procs(nproc, budget_s):
    """cpu_baseline_reference in nproc worker processes at once (the
    reference's routines keep their state in module globals, so one process
    per core), each over its share of the 64-system sample for budget_s;
    throughput = all bead-steps / the longest worker's time."""
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", "", "--cpu-budget", str(budget_s)]
    procs = []
    for k in range(nproc):
        c = list(cmd)
        c[3] = f"{k},{nproc}"
        procs.append(subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True))
    outs = [json.loads(p.communicate(timeout=budget_s * 4 + 120)[0].strip().splitlines()[-1]) for p in procs]
    if any(o is None for o in outs):
        return None
    bead_steps = sum(o["bead_steps"] for o in outs)
    dt = max(o["seconds"] for o in outs)
    return {"value": bead_steps / dt, "unit": "bead-steps/s", "cores": nproc, "kind": "reference",
            "sample": f"{sum(o['systems'] for o in outs)} systems of the workload shared by {nproc} processes "
                      f"(one per core), 300 steps each of the reference's force routines and neighbor() every "
                      f"10 steps (oracle/_ref, fixed coordinates), {dt:.1f} s"}


def cpu_baseline(systems, coords, budget_s=15.0):
    """Oracle Langevin (FP64, O(N^2) list build, rebuild every 10 steps) on
    this host, one core, over chains of the same workload until budget_s."""
    import oracle
    t0 = time.perf_counter()
    bead_steps = 0
    nch = 0
    steps = 300
    order = list(range(0, len(systems), max(1, len(systems) // 64)))
    for k in order:
        s = systems[k]
        o = oracle.Oracle(s)
        x = coords[k]
        o.langevin(x, np.zeros_like(x), np.zeros_like(x), steps, tstep=0.05, rebuild_every=10, seed=1, sys_id=k)
        bead_steps += s.nmp * steps
        nch += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": bead_steps / dt, "unit": "bead-steps/s", "cores": 1, "kind": "port",
            "sample": f"{nch} systems of the workload x {steps} Langevin steps (FP64 oracle, 1 thread), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--eager-steps", type=int, default=200)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--cpu-procs", type=int, default=8, help="processes of the multi-core CPU baseline")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=100)
    ap.add_argument("--serve-scale", type=float, default=0.01)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-ewald", action="store_true")
    ap.add_argument("--no-valence", action="store_true")
    ap.add_argument("--no-bdhi", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--swap-every", type=int, default=1000)
    ap.add_argument("--trigger-steps", type=int, default=1000,
                    help="steps of the displacement-trigger leg (0 = skip)")
    ap.add_argument("--rebuild-every", type=int, default=10,
                    help="neighbour-list rebuild interval in steps; 0 = displacement trigger (margin 8 A)")
    ap.add_argument("--only", choices=["ewald", "bdhi", "copy", "serve"], default=None,
                    help="run one side leg alone (profiling); 'copy' is a plain 1 GiB device copy "
                         "that calibrates the HBM byte counters; 'serve' runs the serving stream with "
                         "--serve-jobs and --serve-scale (1 = the configured U[1e4, 1e6] steps per job)")
    args = ap.parse_args()
    if args.cpu_worker is not None:
        # one worker of the multi-core CPU baseline: no GPU work in this process
        k, m = (int(v) for v in args.cpu_worker.split(","))
        systems, coords, _ = build_workload()
        print(json.dumps(cpu_baseline_reference(systems, coords, args.cpu_budget, share=(k, m))))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the systems of this workload are independent and would be "
              "sharded by bead budget across GPUs, one process per GPU, with no per-step exchange.",
              file=sys.stderr)
        sys.exit(2)
    if args.only == "ewald":
        print(json.dumps({"ewald": ewald_leg()}))
        return
    if args.only == "serve":
        print(json.dumps({"serving": serving_stream(njobs=args.serve_jobs, scale=args.serve_scale)}))
        return
    if args.only == "bdhi":
        print(json.dumps({"bdhi": bdhi_leg(sizes=(300, 3000), steps=20)}))
        return
    if args.only == "copy":
        import torch
        nbytes = 1 << 30
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nbytes, "launches": 3}))
        return

    t_build = time.perf_counter()
    systems, coords, labels = build_workload()
    nbeads = sum(s.nmp for s in systems)
    eng = Engine(systems, keys=list(range(len(systems))), seed=SEED)
    for k, x in enumerate(coords):
        eng.set_coords(k, x)
    eng.langevin_init(tstep=0.05, rebuild_every=args.rebuild_every, step0=0, init_accel=True)
    rb0 = eng.counters()["rebuilds"]
    print(f"[bench] {len(systems)} systems, {nbeads} beads, setup {time.perf_counter() - t_build:.1f} s",
          file=sys.stderr, flush=True)
    # launch shape of the force pass for this batch (no result depends on it)
    tuner = None if args.no_tune else eng.tune(reps=20)
    if tuner:
        print(f"[bench] tuner: pairs/wave {tuner['pairs_per_wave']}, terms/lane {tuner['terms_per_lane']}, "
              f"off-diagonal first {tuner['offdiag_first']}; k_force ms {tuner['candidates_ms']}",
              file=sys.stderr, flush=True)

    # the 64-replica ladder: temperature labels swap every swap_every steps
    from tis_amd.remd import ReplicaLadder
    rep_slots = [k for k, lab in enumerate(labels) if lab == "replica"]
    ladder = ReplicaLadder(eng, rep_slots, systems[rep_slots[0]], [systems[k].tempk for k in rep_slots],
                           seed=SEED, key=1)
    eng.langevin_steps(args.warmup, MODE_GRAPH)
    eng.synchronize()
    t0 = time.perf_counter()
    done = 0
    dev_ms = 0.0
    while done < args.steps:
        chunk = min(args.swap_every - (done % args.swap_every), args.steps - done)
        eng.langevin_steps(chunk, MODE_GRAPH)
        dev_ms += eng.last_timing()["total_ms"]
        done += chunk
        if done % args.swap_every == 0:
            ladder.attempt()
    eng.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = nbeads * args.steps / (t1 - t0)
    rebuilds = eng.counters()["rebuilds"] - rb0
    if args.rebuild_every == 0:
        rebuild_note = (f"{rebuilds / len(systems) / (args.warmup + args.steps):.4f} rebuilds per system-step "
                        f"(displacement trigger)")
    else:
        rebuild_note = f"rebuild every {args.rebuild_every} steps"
    print(f"[bench] batched graph: {ms_per_step:.4f} ms/step, {value:.4e} bead-steps/s "
          f"(device {dev_ms / args.steps:.4f} ms/step); {rebuild_note}", file=sys.stderr, flush=True)

    # per-kernel averages with HIP events on the context's stream (eager launches)
    eng.langevin_steps(args.eager_steps, MODE_EAGER)
    tm = eng.last_timing()
    eager_ms = tm["total_ms"] / args.eager_steps
    print(f"[bench] eager: {eager_ms:.4f} ms/step, k_main {tm['main_ms']:.4f} ms, k_force {tm['force_ms']:.4f} ms, "
          f"all before k_main {tm['pre_ms']:.4f} ms", file=sys.stderr, flush=True)

    # plain one-chain-per-launch path on the same workload
    tp0 = time.perf_counter()
    eng.langevin_steps(args.plain_steps, MODE_PLAIN)
    eng.synchronize()
    tp1 = time.perf_counter()
    plain = nbeads * args.plain_steps / (tp1 - tp0)
    print(f"[bench] plain: {(tp1 - tp0) * 1e3 / args.plain_steps:.3f} ms/step, {plain:.4e} bead-steps/s",
          file=sys.stderr, flush=True)

    # the reference's own rebuild rule on the same batch: lists rebuilt per
    # system by the displacement trigger (margin 8 A) instead of every N steps
    trig = None
    if args.rebuild_every > 0 and args.trigger_steps > 0:
        eng.langevin_init(tstep=0.05, rebuild_every=0,
                          step0=args.warmup + args.steps + args.eager_steps + args.plain_steps, init_accel=False)
        rb_t = eng.counters()["rebuilds"]
        eng.langevin_steps(100, MODE_GRAPH)
        eng.synchronize()
        tt0 = time.perf_counter()
        eng.langevin_steps(args.trigger_steps, MODE_GRAPH)
        eng.synchronize()
        tt1 = time.perf_counter()
        nreb = eng.counters()["rebuilds"] - rb_t
        trig = {"ms_per_step": (tt1 - tt0) * 1e3 / args.trigger_steps,
                "bead_steps_per_s": nbeads * args.trigger_steps / (tt1 - tt0),
                "rebuilds_per_system_step": nreb / len(systems) / (100 + args.trigger_steps)}
        print(f"[bench] displacement trigger: {trig['ms_per_step']:.4f} ms/step, "
              f"{trig['rebuilds_per_system_step']:.4f} rebuilds per system-step", file=sys.stderr, flush=True)

    # lists and forces at the final state; energy_sumup of every system in
    # one launch (the replica loop of energy_allrep.F90:123-125) and
    # energy_unit of the largest system; errors against the FP64 oracle
    eng.neighbor(-1)
    eng.force_all()
    ab = algorithmic_bytes(systems, eng)
    energies = eng.energy_all()
    ncall = 20
    t0 = time.perf_counter()
    for _ in range(ncall):
        eng.energy_all()
    t_all = (time.perf_counter() - t0) / ncall
    big = int(np.argmax([s.nmp for s in systems]))
    eng.energy_unit(big)
    t0 = time.perf_counter()
    for _ in range(ncall):
        eng.energy_unit(big)
    t_unit = (time.perf_counter() - t0) / ncall
    energy = {"systems": len(systems), "ms_per_batch_call": round(t_all * 1e3, 4),
              "energy_unit_ms": round(t_unit * 1e3, 4), "energy_unit_beads": systems[big].nmp,
              "note": "host-timed calls including the copy of the results to the host"}
    print(f"[bench] energy_sumup of {len(systems)} systems: {t_all * 1e3:.3f} ms per call; energy_unit of "
          f"{systems[big].nmp} beads: {t_unit * 1e3:.3f} ms", file=sys.stderr, flush=True)
    ferr = None
    if not args.no_check:
        rng = np.random.default_rng(1)
        sample = sorted(set(rng.choice(len(systems), size=6, replace=False).tolist()) | {big})
        now = [eng.coords(k) for k in range(len(systems))]
        # replica slots carry the parameters of their current ladder label
        ferr, eerr = force_check(eng.systems, now, eng, sample, energies)
        energy["max_rel_err_vs_fp64"] = eerr
        print(f"[bench] force max rel err vs FP64 oracle on {len(sample)} systems: {ferr:.3e}; "
              f"energy terms {eerr:.3e}", file=sys.stderr, flush=True)

    # dominant kernel of the step by HIP-event time on the context's stream:
    # the force pass (k_force: hbond/tertiary-stack systems, electrostatic
    # pairs, bonded terms) or the per-bead pass (k_main: bonded slots,
    # excluded volume, DH sums, integrator)
    main_ms, pre_ms, ele_ms = tm["main_ms"], tm["pre_ms"], tm["force_ms"]
    dom = "k_main" if main_ms >= ele_ms else "k_force"
    dom_bytes = ab["main_bytes"] if dom == "k_main" else ab["force_bytes"]
    dom_ms = main_ms if dom == "k_main" else ele_ms
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic(dom if dom == "k_main" or tuner is None or
                                   (tuner["pairs_per_wave"], tuner["terms_per_lane"]) == (1, 1)
                                   else f"k_force<{tuner['pairs_per_wave']}, {tuner['terms_per_lane']}>"),
            "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
            "avg_launch_ms": round(dom_ms, 5), "bytes_per_launch": dom_bytes,
            "k_main_ms": round(main_ms, 5), "k_force_ms": round(ele_ms, 5), "pre_main_ms": round(pre_ms, 5)}

    eng.close()
    serve = None
    if not args.no_serve:
        serve = serving_stream(njobs=args.serve_jobs, scale=args.serve_scale)
        print(f"[bench] serving: {serve['jobs']} jobs, p50 {serve.get('p50_latency', 0):.3f} s, "
              f"p99 {serve.get('p99_latency', 0):.3f} s, {serve['bead_steps_per_s_busy']:.4e} bead-steps/s, "
              f"mean in flight {serve.get('mean_inflight', 0):.2f}", file=sys.stderr, flush=True)
    valence = None if args.no_valence else valence_leg()
    if valence:
        print(f"[bench] valence pass: {valence['ms_per_step']:.4f} ms/step, "
              f"{valence['hb_deleted_per_step']:.2f} hbonds deleted per step", file=sys.stderr, flush=True)
    ewald = None if args.no_ewald else ewald_leg()
    if ewald:
        print(f"[bench] ewald: {ewald['ms_per_step']:.4f} ms/step, force err {ewald['force_max_rel_err_vs_fp64']:.2e}",
              file=sys.stderr, flush=True)
    bdhi = None if args.no_bdhi else bdhi_leg()
    cpu = None
    if not args.no_cpu:
        cpu = None
        if args.cpu_procs > 1:
            one = cpu_baseline_reference(systems, coords, args.cpu_budget / 2)
            if one is not None:
                cpu = cpu_baseline_reference_procs(args.cpu_procs, args.cpu_budget / 2)
                if cpu is not None:
                    cpu["single_core_value"] = one["value"]
        cpu = cpu or cpu_baseline_reference(systems, coords, args.cpu_budget) or cpu_baseline(systems, coords,
                                                                                       args.cpu_budget)
    out = {
        "metric": "batched bead-steps/s vs plain GPU @ force err <=1e-5 vs FP64; p50/p99 job latency",
        "value": value,
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair / fp64 accumulate+integrate",
        "data": "synthetic (seeded hairpin designs, extended-helix start)",
        "config": {"workload": "512 lognormal(ln120,0.7) 20-600 nt DT15 chains + 64 replicas of the 64-nt "
                               "hairpin on a 290-400 K ladder with swap attempts every 1000 steps; "
                               "Langevin dt 0.05 tau; " + (
                                   "neighbour lists rebuilt by the displacement trigger (margin 8 A)"
                                   if args.rebuild_every == 0 else f"rebuild every {args.rebuild_every} steps"),
                   "systems": len(systems), "beads": nbeads, "parallelism": "single GPU, batched"},
        "plain_gpu_bead_steps_per_s": plain,
        "batched_over_plain": value / plain,
        "force_max_rel_err_vs_fp64": ferr,
        "energy": energy,
        "p50_job_latency_s": serve.get("p50_latency") if serve else None,
        "p99_job_latency_s": serve.get("p99_latency") if serve else None,
        "serving": serve,
        "remd": {"replicas": len(rep_slots), "swap_every": args.swap_every, "attempts": ladder.attempts,
                 "acceptance": ladder.acceptance},
        "displacement_trigger": trig,
        "valence": valence,
        "ewald": ewald,
        "bdhi": bdhi,
        "roofline": roof,
        "cpu_baseline": cpu,
        "tuner": tuner,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
