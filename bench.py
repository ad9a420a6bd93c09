"""Benchmark: LJ timestep throughput on one MI355X (atom-steps/s).

Workload (the reference's LJ melt at scale): fcc lattice at reduced density
0.8442, 80x80x80 unit cells = 2,048,000 atoms, periodic cube, eps = sigma =
mass = 1, cutoff 2.5, skin 0.3, full Verlet list rebuilt every 20 steps,
Gaussian velocities from the reference's seeded generator at T = 1.44
(zero net momentum), dt = 0.005, NVE, thermo every 100 steps.  A "step" is one MD
timestep of the whole system (integrate, list maintenance, force, thermo).

Prints ONE JSON line.  value = atom-steps/s of the headline variant
(N * steps / wall time of the timed steps, cabanamd_impl.h:406-407), with
all state resident in HBM.  The roofline object is for the fused LJ force +
integrate kernel, timed with HIP events on the engine's stream during the
timed steps.  cpu_baseline times the FP64 OpenMP oracle on a bounded sample
of the same workload on the host's cores.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("atom-steps/s/GPU (LJ 2M, NNP) + batched vs plain speedup, "
          "p50/p99 job latency")
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per atom of the fused full-list kernel: the LJ full
# force per-atom figure (x_i 24 + offset 4 + 78 neighbour indices x 4 +
# f_i 24 = 364 B) plus the fused integrator traffic (v read 24, v write 24,
# x write 24 = 72 B).
LJ_FULL_BYTES_PER_ATOM = 364
FUSED_NVE_BYTES_PER_ATOM = 72
ALG_BYTES_PER_ATOM = LJ_FULL_BYTES_PER_ATOM + FUSED_NVE_BYTES_PER_ATOM
# the LJ legs time one fused force launch in this many steps with HIP events
PROFILE_EVERY = 4


def _force_error(md):
    """Forces of the engine on the current (liquid) state against the FP64
    oracle's full-list LJ on the same positions: ||dF||_2 / ||F||_2 and
    max_i |dF_i| / RMS |F|.  (At step 0 the atoms sit on the lattice, where
    forces cancel, so the check runs after the timed steps.)"""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    md.compute_forces()
    x, v, f, t, i = md.arrays()
    n = md.n_local
    off, idx = O.neigh_build(x, n, 2.8, half=False)
    lj1, lj2, cs = O.lj_coeff(1, [(1, 1, 1.0, 1.0, 2.5)])
    fo = O.lj_force(x, t, n, 1, lj1, lj2, cs, off, idx, False)[:n]
    d = f[:n] - fo
    rms = float(np.sqrt((fo ** 2).sum(axis=1).mean()))
    return dict(rel_l2=float(np.linalg.norm(d) / np.linalg.norm(fo)),
                max_over_rms=float(np.sqrt((d ** 2).sum(axis=1)).max() / rms))


def _drift(e):
    """max |E - E_0| / |E_0| over the thermo rows, None with fewer than two."""
    if len(e) < 2:
        return None
    return float(abs(e - e[0]).max() / abs(e[0]))


def _variant(M, ncells, steps, warmup, half, prec, lanes=0, accuracy=False):
    p = M.lattice_params(ncells=ncells, half=half, thermo=100,
                         nsteps=steps, gaussian=True)
    md = M.CbnMD(p, precision=prec, lanes=lanes)
    t0 = time.time()
    md.init()
    t_init = time.time() - t0
    g, tile, tune_s, width = md.tune_info()
    n = md.n_local
    if warmup:
        md.run(warmup)
    M._lib.call("cmd_device_synchronize")
    # HIP events around one fused force launch in PROFILE_EVERY: each event
    # is a marker that idles the device a few microseconds, so timing every
    # launch would slow every step it times
    md.profile(PROFILE_EVERY)
    t0 = time.perf_counter()
    md.run(steps)
    M._lib.call("cmd_device_synchronize")
    dt = time.perf_counter() - t0
    launches, mean_ms = md.kernel_stats()
    ferr = _force_error(md) if accuracy else None
    th = md.thermo()
    out = dict(n=n, steps=steps, seconds=dt, ms_per_step=1e3 * dt / steps,
               atom_steps_per_s=n * steps / dt, kernel_launches=launches,
               kernel_mean_ms=mean_ms, init_s=t_init, tune_s=tune_s,
               lanes=g, tile=f"{width}x{width}x{tile}",
               energy_drift=_drift(th[:, 3]), thermo_rows=len(th),
               force_error_vs_fp64_oracle=ferr,
               timers=md.timers())
    del md
    return out


def _host_threads():
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return int(os.environ.get("OMP_NUM_THREADS", threads))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _oracle_run(ncells, steps, thermo=100, half=False, timers=False):
    """The FP64 OpenMP oracle's MD loop on the LJ workload with Gaussian
    velocities: (atom-steps/s of the step loop, seconds, thermo rows), and
    with timers its phase times (the reference's T_Force, T_Neigh, T_Comm,
    T_Int, T_Other split, cabanamd_impl.h:404-425)."""
    # threads bound one per core for the OpenMP runtime the oracle loads
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    md = O.MD(O.lj_params(ncells=ncells, thermo=thermo, nsteps=steps,
                          gaussian=True, half=half))
    md.init()
    t0 = time.perf_counter()
    md.run(steps)
    dt = time.perf_counter() - t0
    out = (md.n_local * steps / dt, dt, md.thermo())
    if timers:
        out += ({k: float(v) for k, v in md.timers().items()},)
    return out


def _cpu_baseline(ncells, steps):
    rate, dt, _, split = _oracle_run(ncells, steps, timers=True)
    n = 4 * ncells[0] * ncells[1] * ncells[2]
    return dict(value=rate, unit="atom-steps/s", cores=_host_threads(),
                kind="port", cpu_model=_cpu_model(), split_s=split,
                sample=(f"FP64 OpenMP oracle (restatement of the reference "
                        f"algorithm), same workload, {n} atoms, {steps} "
                        f"steps incl. {steps // 20} list rebuilds, "
                        f"{dt:.2f} s"))


def _c1_case(M, cpu=True, drift_steps=1000):
    """The reference's CPU-runnable LJ case: fcc rho 0.8442, 20^3 cells =
    32,000 atoms, full list rebuilt every 20 steps, Gaussian velocities at
    T 1.44, dt 0.005, 100 NVE steps, FP64, thermo every 100.  Timed on the
    GPU (FP64) and with the FP64 oracle on the host's cores; then the
    total-energy drift over drift_steps steps (thermo every 100) of the
    engine in both precisions against the oracle's drift on the same deck."""
    import numpy as np
    nc = (20, 20, 20)
    out = {}
    for name, prec in (("fp64", M.PREC_FP64), ("mixed", M.PREC_MIXED)):
        p = M.lattice_params(ncells=nc, thermo=100, nsteps=100, gaussian=True)
        # an untimed run of the same deck first: the first launch of each
        # kernel in the process loads its code, which is not the step's cost
        warm = M.CbnMD(p, precision=prec)
        warm.init()
        warm.run(100)
        del warm
        md = M.CbnMD(p, precision=prec)
        md.init()
        M._lib.call("cmd_device_synchronize")
        t0 = time.perf_counter()
        md.run(100)
        M._lib.call("cmd_device_synchronize")
        dt = time.perf_counter() - t0
        md.run(drift_steps - 100)
        e = md.thermo()[:, 3]
        out[name] = dict(n=md.n_local, steps=100, seconds=dt,
                         atom_steps_per_s=md.n_local * 100 / dt,
                         ms_per_step=10.0 * dt, lanes=md.tune_info()[0],
                         drift_steps=drift_steps, energy_drift=_drift(e),
                         e_final=float(e[-1]))
        del md
    if cpu:
        rate, dt, _, split = _oracle_run(nc, 100, timers=True)
        out["cpu_oracle"] = dict(atom_steps_per_s=rate, seconds=dt,
                                 cores=_host_threads(), kind="port",
                                 cpu_model=_cpu_model(), split_s=split)
        _, _, th = _oracle_run(nc, drift_steps)
        e = th[:, 3]
        out["cpu_oracle"].update(drift_steps=drift_steps,
                                 energy_drift=_drift(e), e_final=float(e[-1]))
        for name in ("fp64", "mixed"):
            out[name]["e_final_minus_oracle"] = out[name]["e_final"] - e[-1]
    return out


NNP_SIZES = {3000: (10, 15, 20), 24000: (20, 30, 40), 192000: (40, 60, 80)}


def _nnp_case(M, n, steps, warmup, prec, accuracy=False):
    """The NNP benchmark configuration: two elements on a jittered simple
    cubic lattice (spacing 2.5, jitter 0.1 s, 2:1 ratio, seeded), 30 radial +
    40 angular symmetry functions per element, cutoff 6.0, network
    70-25-25-1 with seeded weights and min-max scaling fitted to the initial
    frame; T = 0.05, dt = 0.0005, list rebuilt every 20 steps."""
    import numpy as np
    cells = NNP_SIZES[n]
    x, t, ids = M.nnp_lattice(cells, spacing=2.5, jitter=0.1, seed=87287)
    L = [2.5 * c for c in cells]
    s = M.System(2)
    s.create_domain([0, 0, 0], L, 6.3)
    s.upload(x, None, None, t, ids)
    comm = M.Comm(s, 6.3)
    comm.exchange_halo()
    nl = M.NeighborVerlet(6.3, False, 100, lanes=1)
    nl.create(s)
    F = M.ForceNNP(M.NNPModel.synthetic(seed=87287, rc=6.0), M.PREC_FP64)
    F.fit_scaling(s, nl)
    model = F.model()
    rng = np.random.default_rng(87287)
    v = rng.normal(0, np.sqrt(0.05), x.shape)
    v -= v.mean(axis=0)
    p = M.nnp_params(nsteps=steps, thermo=max(steps // 10, 1))
    md = M.CbnMD(p, precision=prec)
    md.set_nnp_model(model)
    md.set_initial_state(x, t, [0, 0, 0], L, v=v, id=ids)
    t0 = time.time()
    md.init()
    t_init = time.time() - t0
    if warmup:
        md.run(warmup)
    M._lib.call("cmd_device_synchronize")
    t0 = time.perf_counter()
    md.run(steps)
    M._lib.call("cmd_device_synchronize")
    dt = time.perf_counter() - t0
    # the force phase's GPU time, from a few more steps with the engine's
    # per-phase events on (kept out of the timed steps)
    md.profile(True)
    md.run(min(steps, 10))
    _, force_ms = md.kernel_stats()
    md.profile(False)
    # the work of one force call under the engine's per-term flop model
    # (FP32: add, multiply 1, fused multiply-add 2, transcendental 1)
    wk = (C.c_double * 6)()
    M._lib.call("cmd_force_nnp_work", md.force.handle, wk)
    gf = [wk[2] / 1e9, wk[3] / 1e9, wk[4] / 1e9]
    work = dict(neighbours_per_atom=wk[0] / max(wk[5], 1),
                triplets_per_atom=wk[1] / max(wk[5], 1),
                gflop_symmetry_functions=gf[0], gflop_force_chain=gf[1],
                gflop_layers=gf[2],
                force_phase_ms=force_ms,
                model_tflops=sum(gf) / force_ms if force_ms > 0 else None)
    th = md.thermo()
    pe = th[:, 2] * n / 27.211384021355236
    ke = (th[:, 3] - th[:, 2]) * n
    e = pe + ke
    out = dict(n=n, steps=steps, seconds=dt, ms_per_step=1e3 * dt / steps,
               atom_steps_per_s=n * steps / dt, init_s=t_init,
               energy_drift=float(np.abs(e - e[0]).max() / np.abs(e).max()),
               layers="bf16-mfma" if prec == M.PREC_MIXED else "fp32",
               work=work)
    if accuracy:
        out["force_error_vs_fp64_oracle"] = _nnp_force_error(M, md, model)
    del md
    return out


def _nnp_force_error(M, md, model):
    """Engine forces on the final state against the FP64 NNP oracle."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pyoracle as O
    from nnp_models import to_oracle
    md.compute_forces()
    x, v, f, t, i = md.arrays()
    n = md.n_local
    off, idx = md.neighbor.get()
    m = model.describe(2)
    m.update(act="ttl", smin=0.0, smax=1.0, scale=True)
    Mo = to_oracle(O, m)
    _, _, _, _, fo = Mo.compute(x, t, n, off, idx, ld=96)
    send = np.zeros(1, np.int32)
    import ctypes as C
    counts = np.zeros(6, np.int32)
    M._lib.call("cmd_comm_ghost_lists", md._sys, None,
                counts.ctypes.data_as(C.POINTER(C.c_int)))
    send = np.zeros(max(int(counts.sum()), 1), np.int32)
    M._lib.call("cmd_comm_ghost_lists", md._sys,
                send.ctypes.data_as(C.POINTER(C.c_int)),
                counts.ctypes.data_as(C.POINTER(C.c_int)))
    O.ghost_fold(fo, n, send[:counts.sum()], counts)
    d = f[:n] - fo[:n]
    rms = float(np.sqrt((fo[:n] ** 2).sum(axis=1).mean()))
    return dict(rel_l2=float(np.linalg.norm(d) / np.linalg.norm(fo[:n])),
                max_over_rms=float(np.sqrt((d ** 2).sum(axis=1)).max() / rms))


def _batch_specs(nsys, seed):
    """The batched-ensemble workload: per system k, N_k log-uniform in
    [500, 20000], fcc nx*ny*nz cells with 4 nx ny nz ~ N_k (near-cubic),
    density U[0.70, 0.90], temperature U[0.8, 2.0], seed + 1000003 k."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nsys):
        nk = float(np.exp(rng.uniform(np.log(500.0), np.log(20000.0))))
        rho = float(rng.uniform(0.70, 0.90))
        temp = float(rng.uniform(0.8, 2.0))
        nc = max(1, round(nk / 4.0))
        nx = max(1, round(nc ** (1.0 / 3.0)))
        ny = max(1, round((nc / nx) ** 0.5))
        nz = max(1, round(nc / (nx * ny)))
        out.append(((nx, ny, nz), rho, temp, seed + 1000003 * k))
    return out


def _batch_case(M, nsys, steps, prec, plain_every, seed=87287):
    """All systems in one batch (one launch per kernel per step for the
    whole ensemble) against the plain loop: the single-system engine run
    on each system in turn (timed on every plain_every-th system)."""
    import numpy as np
    specs = _batch_specs(nsys, seed)
    p = M.lattice_params(thermo=100, every=20)
    B = M.Batch(p, prec, lanes=0)
    states, handles = [], []
    for spec in specs:
        st = M.fcc_system(*spec)
        states.append(st)
        x, v, t, i, lo, hi = st
        handles.append(B.add(x, t, lo, hi, v=v, id=i))
    t0 = time.perf_counter()
    B.run(0)  # join: ghosts, lists, lane autotune, step-0 forces
    t_join = time.perf_counter() - t0
    natoms = B.counts()["atoms"]
    M._lib.call("cmd_device_synchronize")
    t0 = time.perf_counter()
    B.run(steps)
    M._lib.call("cmd_device_synchronize")
    dt = time.perf_counter() - t0
    drift = []
    for h in handles:
        e = B.thermo(h)[:, 3]
        drift.append(float(np.abs(e - e[0]).max() / abs(e[0])))
    plain_as, plain_t, nplain = 0.0, 0.0, 0
    for k in range(0, nsys, plain_every):
        x, v, t, i, lo, hi = states[k]
        pk = M.lattice_params(thermo=100, every=20, nsteps=steps)
        md = M.CbnMD(pk, precision=prec, lanes=0)
        md.set_initial_state(x, t, lo, hi, v=v, id=i)
        md.init()
        M._lib.call("cmd_device_synchronize")
        t1 = time.perf_counter()
        md.run(steps)
        M._lib.call("cmd_device_synchronize")
        plain_t += time.perf_counter() - t1
        plain_as += len(x) * steps
        nplain += 1
        del md
    batched = natoms * steps / dt
    plain = plain_as / plain_t
    return dict(systems=nsys, atoms=int(natoms), steps=steps,
                seconds=dt, ms_per_step=1e3 * dt / steps,
                batched_atom_steps_per_s=batched,
                plain_atom_steps_per_s=plain, speedup=batched / plain,
                plain_sample=f"{nplain} of {nsys} systems (every "
                             f"{plain_every}th), {int(plain_as)} atom-steps",
                lanes=B.counts()["lanes"], join_s=t_join,
                drift_median=float(np.median(drift)),
                drift_max=float(np.max(drift)), timers=B.timers())


_MODEL = {}


def _nnp_bench_model(M):
    """The NNP benchmark model (seeded synthetic, 70 functions per element)
    with min-max scaling fitted to a 16x16x16 jittered lattice frame."""
    if "m" not in _MODEL:
        x, t, ids = M.nnp_lattice((16, 16, 16), spacing=2.5, jitter=0.1,
                                  seed=87287)
        L = [40.0] * 3
        s = M.System(2)
        s.create_domain([0, 0, 0], L, 6.3)
        import numpy as np
        s.upload(np.mod(x, 40.0), None, None, t, ids)
        comm = M.Comm(s, 6.3)
        comm.exchange_halo()
        nl = M.NeighborVerlet(6.3, False, 100, lanes=1)
        nl.create(s)
        F = M.ForceNNP(M.NNPModel.synthetic(seed=87287, rc=6.0), M.PREC_FP64)
        F.fit_scaling(s, nl)
        _MODEL["m"] = F.model()
        _MODEL["f"] = F
    return _MODEL["m"]


def _near_cube(n):
    nx = max(1, round(n ** (1.0 / 3.0)))
    ny = max(1, round((n / nx) ** 0.5))
    nz = max(1, round(n / (nx * ny)))
    return (nx, ny, nz)


def _serve_jobs(M, njobs, seed):
    """Served request stream: 70 % LJ / 30 % NNP, N log-uniform in
    [1000, 200000], steps uniform in [100, 1000]; LJ jobs are fcc boxes
    (density U[0.70, 0.90], T U[0.8, 2.0]), NNP jobs jittered lattices at
    spacing 2.5 and T 0.05; seeds seed + 1000003 k."""
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(njobs):
        nnp = rng.uniform() < 0.3
        n = float(np.exp(rng.uniform(np.log(1000.0), np.log(200000.0))))
        steps = int(rng.integers(100, 1001))
        sk = seed + 1000003 * k
        if nnp:
            st = M.nnp_system(_near_cube(n), temp=0.05, seed=sk)
            spec = None
        else:
            rho = float(rng.uniform(0.70, 0.90))
            temp = float(rng.uniform(0.8, 2.0))
            spec = (_near_cube(n / 4.0), rho, temp, sk)
            st = M.fcc_system(*spec)
        jobs.append((2 if nnp else 0, steps, st, spec))
    return jobs


def _oracle_lj_drift(spec, steps):
    """Energy drift of the FP64 oracle's MD loop on an LJ job's state."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    cells, rho, temp, seed = spec
    md = O.MD(O.lj_params(ncells=cells, rho=rho, temp=temp, seed=seed,
                          gaussian=True, thermo=20, every=20))
    md.init()
    md.run(steps)
    e = md.thermo()[:, 3]
    return float(np.abs(e - e[0]).max() / np.abs(e).max())


def _serve_run(M, model, jobs, arrivals, prec, oracle_jobs=0):
    import numpy as np
    S = M.Server(M.lattice_params(thermo=100, every=20),
                 M.nnp_params(thermo=100), model, precision=prec, lanes=0)
    t0 = S.clock()
    ids = []
    for (kind, steps, st, _), a in zip(jobs, arrivals):
        x, v, t, i, lo, hi = st
        ids.append(S.submit(kind, x, t, lo, hi, steps, v=v, id=i,
                            arrival=t0 + a))
    S.drain(900)
    lat, drift = [], {0: [], 2: []}
    vs_oracle = []
    t_end = 0.0
    for j, (kind, steps, st, spec) in zip(ids, jobs):
        info = S.job(j)
        lat.append(info["done"] - info["arrival"])
        t_end = max(t_end, info["done"])
        th = S.result(j)[3]
        n = info["n"]
        if kind == 2:
            e = th[:, 2] * n / 27.211384021355236 + (th[:, 3] - th[:, 2]) * n
        else:
            e = th[:, 3] * n
        drift[kind].append(float(np.abs(e - e[0]).max() / np.abs(e).max()))
        if kind == 0 and len(vs_oracle) < oracle_jobs and n <= 20000:
            vs_oracle.append(dict(n=n, steps=steps, gpu=drift[0][-1],
                                  oracle_fp64=_oracle_lj_drift(spec, steps)))
    atom_steps = sum(len(st[0]) * steps for _, steps, st, _ in jobs)
    span = t_end - (t0 + arrivals[0])
    return dict(jobs=len(jobs), atom_steps=atom_steps, makespan_s=span,
                atom_steps_per_s=atom_steps / span,
                jobs_per_s=len(jobs) / span,
                p50_latency_s=float(np.percentile(lat, 50)),
                p99_latency_s=float(np.percentile(lat, 99)),
                drift_lj_median=float(np.median(drift[0])) if drift[0] else None,
                drift_lj_max=float(np.max(drift[0])) if drift[0] else None,
                drift_nnp_median=float(np.median(drift[2])) if drift[2] else None,
                drift_nnp_max=float(np.max(drift[2])) if drift[2] else None,
                lj_drift_vs_fp64_oracle=vs_oracle, stats=S.stats())


def _serve_case(M, ncap, njobs, prec, seed=87287):
    """Capacity: ncap jobs all arriving at once (jobs/s at saturation);
    then njobs Poisson arrivals at 0.7x that rate, reporting p50/p99 latency
    (arrival to result), atom-steps/s and per-job energy drift."""
    import numpy as np
    model = _nnp_bench_model(M)
    cap_jobs = _serve_jobs(M, ncap, seed)
    cap = _serve_run(M, model, cap_jobs, [0.0] * ncap, prec)
    rate = 0.7 * cap["jobs_per_s"]
    jobs = _serve_jobs(M, njobs, seed + 7)
    rng = np.random.default_rng(seed + 11)
    arr = np.cumsum(rng.exponential(1.0 / rate, njobs))
    arr -= arr[0]
    served = _serve_run(M, model, jobs, list(arr), prec, oracle_jobs=4)
    served["arrival_rate_jobs_per_s"] = rate
    served["capacity"] = cap
    return served


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=80)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--c1", type=int, default=1,
                    help="time the 32,000-atom case and its 1000-step drift")
    ap.add_argument("--variants", default="mixed,fp64,half")
    ap.add_argument("--nnp", default="3000,24000,192000",
                    help="NNP system sizes to time ('' to skip)")
    ap.add_argument("--nnp-steps", type=int, default=200)
    ap.add_argument("--batch", type=int, default=512,
                    help="systems in the batched-ensemble leg (0 to skip)")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--batch-plain-every", type=int, default=4,
                    help="time the plain per-system loop on every k-th system")
    ap.add_argument("--serve", type=int, default=40,
                    help="jobs in the served-stream leg (0 to skip)")
    ap.add_argument("--serve-capacity-jobs", type=int, default=24)
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a 2 GiB streaming copy (k_copy_probe) "
                         "to calibrate byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding of one system is not built", file=sys.stderr)
        sys.exit(2)

    import cabanamd_amd as M
    if M.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    nc = (args.cells,) * 3
    copy_gbs = None
    if args.calibrate:
        import ctypes as C
        g, ms = C.c_double(), C.c_double()
        M._lib.call("cmd_bandwidth_probe", C.c_int64(1 << 31), 10,
                    C.byref(g), C.byref(ms))
        copy_gbs = g.value
        print(f"# copy probe: 2 GiB read + 2 GiB write, {ms.value:.4f} ms, "
              f"{copy_gbs:.1f} GB/s", file=sys.stderr, flush=True)
    variants = args.variants.split(",")
    res = {}
    for v in variants:
        half = v == "half"
        prec = M.PREC_MIXED if v == "mixed" else M.PREC_FP64
        res[v] = _variant(M, nc, args.steps, args.warmup, half, prec,
                          accuracy=not args.no_cpu)
        print(f"# {v}: {res[v]['atom_steps_per_s']:.4e} atom-steps/s, "
              f"{res[v]['ms_per_step']:.4f} ms/step, kernel "
              f"{res[v]['kernel_mean_ms']:.4f} ms x {res[v]['kernel_launches']}"
              f", lanes {res[v]['lanes']}", file=sys.stderr, flush=True)
    head = variants[0]
    h = res[head]
    n = h["n"]
    achieved = (ALG_BYTES_PER_ATOM * n / (h["kernel_mean_ms"] * 1e-3) / 1e9
                if h["kernel_mean_ms"] > 0 else None)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "lj_full_fused_2M_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    c1 = None
    if args.c1:
        c1 = _c1_case(M, cpu=not args.no_cpu)
        print(f"# C1 32,000 atoms FP64: {c1['fp64']['atom_steps_per_s']:.4e} "
              f"atom-steps/s GPU"
              + (f", {c1['cpu_oracle']['atom_steps_per_s']:.4e} CPU oracle; "
                 f"1000-step drift GPU {c1['fp64']['energy_drift']:.3e} "
                 f"oracle {c1['cpu_oracle']['energy_drift']:.3e}"
                 if 'cpu_oracle' in c1 else ""), file=sys.stderr, flush=True)
    nnp = {}
    for sz in [int(v) for v in args.nnp.split(",") if v]:
        for prec in (M.PREC_MIXED, M.PREC_FP64):
            r = _nnp_case(M, sz, args.nnp_steps, 5, prec,
                          accuracy=(not args.no_cpu and sz == 3000))
            nnp[f"{sz}-{r['layers']}"] = r
            w = r["work"]
            print(f"# nnp {sz} {r['layers']}: {r['atom_steps_per_s']:.4e} "
                  f"atom-steps/s, {r['ms_per_step']:.3f} ms/step, drift "
                  f"{r['energy_drift']:.2e}; per force call "
                  f"{w['gflop_symmetry_functions']:.2f} + "
                  f"{w['gflop_force_chain']:.2f} + {w['gflop_layers']:.3f} "
                  f"GFLOP in {w['force_phase_ms']:.3f} ms = "
                  f"{w['model_tflops'] or 0:.1f} TFLOP/s (model), "
                  f"{w['neighbours_per_atom']:.1f} neighbours and "
                  f"{w['triplets_per_atom']:.0f} triplets per atom",
                  file=sys.stderr, flush=True)
    batch = None
    if args.batch > 0:
        batch = _batch_case(M, args.batch, args.batch_steps, M.PREC_MIXED,
                            args.batch_plain_every)
        print(f"# batch {batch['systems']} systems / {batch['atoms']} atoms: "
              f"{batch['batched_atom_steps_per_s']:.4e} atom-steps/s batched,"
              f" {batch['plain_atom_steps_per_s']:.4e} plain, speedup "
              f"{batch['speedup']:.2f}, drift max {batch['drift_max']:.2e}",
              file=sys.stderr, flush=True)
    serve = None
    if args.serve > 0:
        serve = _serve_case(M, args.serve_capacity_jobs, args.serve,
                            M.PREC_MIXED)
        print(f"# serve {serve['jobs']} jobs at {serve['arrival_rate_jobs_per_s']:.2f}"
              f" jobs/s: p50 {serve['p50_latency_s']:.3f} s, p99 "
              f"{serve['p99_latency_s']:.3f} s, {serve['atom_steps_per_s']:.4e} "
              f"atom-steps/s (capacity {serve['capacity']['jobs_per_s']:.2f} "
              f"jobs/s)", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(nc, args.cpu_steps)
    line = {
        "metric": METRIC,
        "value": h["atom_steps_per_s"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-pair/f64" if head == "mixed" else "f64",
        "data": "synthetic (seeded fcc lattice + Gaussian velocities)",
        "config": {
            "workload": ("LJ fcc rho 0.8442, 80^3 cells = 2,048,000 atoms, "
                         "rc 2.5, skin 0.3, full Verlet list rebuilt every 20"
                         " steps, dt 0.005, NVE, thermo every 100"),
            "n_atoms": n, "list": "half" if head == "half" else "full",
            "precision": head, "lanes_per_atom": h["lanes"]},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": traffic,
            "kernel": "k_tile_force (LDS-tiled full-list LJ fused with "
                      "the NVE half-kicks, drift and wrap), mean of "
                      f"{h['kernel_launches']} launches (one in "
                      f"{PROFILE_EVERY} of the timed steps) = "
                      f"{h['kernel_mean_ms']:.4f} ms, "
                      f"{ALG_BYTES_PER_ATOM} algorithmic B/atom"},
        "cpu_baseline": cpu,
        "c1_32k": c1,
        "nnp": nnp,
        "batch": batch,
        "serve": serve,
        "copy_probe_gbs": copy_gbs,
        "variants": {k: {kk: vv for kk, vv in r.items() if kk != "timers"}
                     for k, r in res.items()},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
