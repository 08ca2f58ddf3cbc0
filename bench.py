"""Benchmark: SCF molecules/s on one MI355X, mixed precision against FP64, with p50/p99 latency and ERI/s.

Default workload (the serving configuration): RHF requests drawn from a seeded mix of water clusters (1-12
waters) and n-alkanes (C1-C16) in 6-31G (13-212 Cartesian functions), served by the continuous-batching server
(32 slots stepped in lockstep — each phase of an SCF iteration one launch over every admitted request — each
request's ERIs evaluated once into an in-core store in HBM, device-resident SCF state).
A step is one batch of --batch requests (default 100): --steps 20 is the 2000-request workload.  The timed region
is the mixed-precision tier serving all steps x batch requests queued at once (closed loop: the server's
sustained rate); the same requests are then served in FP64, and both tiers again under Poisson arrivals at 0.8x
their measured capacity for p50/p99 latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
    python bench.py --workload decane|water|water32|eicosyl      (single-molecule SCF legs)

Prints ONE JSON line.  value = mixed-tier molecules/s counting only requests that converged (1e-8 Eh) with
|E_mixed - E_FP64| <= 1e-6 Eh.  `roofline` is the workload's dominant kernel (most device time in the warm-up
pass), timed with HIP events on its own launch stream during the timed region.  `cpu_baseline` is the oracle
(a CPU restatement of the reference's algorithm: per-component HGP ERIs into the 8-fold packed store, then an
unscreened N^4 Fock loop per iteration) on a bounded sample of the same request mix, one SCF per process on P
host cores as the reference runs one SCF per thread, extrapolated to molecules/s.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector (and matrix)
PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector
PEAK_HBM_GBS = 8000.0

# SCF settings of every leg: the reference's SCF (DIIS 6/5, core guess) with normalIterAfterDIIS = 0 and the
# configurations' convergence criterion of 1e-8 Eh; the reference's density and DIIS-error criteria (1e-10,
# HartreeFock/Constants.h:6-10) are loosened to 1e-6, which the 1e-8 Eh energy criterion dominates
SCF_KW = dict(alpha=0.5, init_guess=0.0, use_diis=1, normal_iter_after_diis=0,
              energy_tolerance=1e-8, density_tolerance=1e-6, diis_tolerance=1e-6)
SCF_DESC = ("RHF, core guess, alpha 0.5, DIIS 6 kept / 5 to extrapolate, normalIterAfterDIIS 0; converged when "
            "|dE| <= 1e-8 Eh, ||dD||_F <= 1e-6 and DIIS error <= 1e-6")


def hip_sync():
    ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------------------------- CPU baseline
def _cpu_molecule_seconds(bfile, atoms, iterations, budget_s):
    """Oracle (reference algorithm) on one core for one molecule, sampled: every stride-th (i,j,k) triple of the
    packed ERI loop with its whole l loop (~60% of the budget) and some rows of one N^4 Fock loop (~40%),
    extrapolated to the molecule's ERI phase plus `iterations` Fock builds.  Returns (seconds, description)."""
    import numpy as np
    from oracle import refcpu
    mol = refcpu.Molecule(bfile, atoms)
    n = mol.nbf
    unique = refcpu.packed_index(n - 1, n - 1, n - 1, n - 1) + 1
    triples = n * (n + 1) // 2 * n
    stride = max(1, int(triples / max(1.0, 0.6 * budget_s / 19e-6 / (n / 2.0))))
    t0 = time.perf_counter()
    eri, cnt = mol.eri_packed_sample(stride, stride // 2)
    t_eri = time.perf_counter() - t0
    eri_phase = t_eri / max(cnt, 1) * unique
    rows = max(2, min(n, int(0.4 * budget_s / (20.0 / 194.0 * (n / 194.0) ** 3))))
    D = np.eye(n) * 0.1
    t0 = time.perf_counter()
    refcpu.rhf_g(n, eri, D, rows=(0, rows))
    t_g = time.perf_counter() - t0
    fock = t_g * n / rows
    total = eri_phase + iterations * fock
    return total, ("n=%d: %d of %d unique ERIs (%.1fs -> ERI phase %.0fs) + %d of %d Fock rows (%.2fs -> %.2fs per "
                   "iteration) x %d iterations = %.0fs" % (n, cnt, unique, t_eri, eri_phase, rows, n, t_g, fock,
                                                          iterations, total))


def _cpu_worker(job):
    bfile, mols, budget_s = job
    return [_cpu_molecule_seconds(bfile, atoms, it, budget_s)[0] for atoms, it in mols]


def host_cores():
    """Cores this process may use, capped at the host share of one GPU (16)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline_serving(bfile, sample, budget_s):
    """The oracle on P host processes at once (one SCF per core, as the reference runs one SCF per std::thread,
    HartreeFock/HartreeFockDoc.cpp:379-401): each process times two molecules of the sample (spread over the
    size range of the mix) within budget_s, and the mean extrapolated time per molecule gives P / mean
    molecules/s."""
    import concurrent.futures as cf
    import multiprocessing as mp
    P = host_cores()
    per = max(1, len(sample) // P)
    jobs = [(bfile, sample[w * per:(w + 1) * per], budget_s / per) for w in range(P) if sample[w * per:(w + 1) * per]]
    t0 = time.perf_counter()
    with cf.ProcessPoolExecutor(max_workers=len(jobs), mp_context=mp.get_context("spawn")) as ex:
        secs = [s for r in ex.map(_cpu_worker, jobs) for s in r]
    wall = time.perf_counter() - t0
    mean = sum(secs) / len(secs)
    return {"value": len(jobs) / mean, "unit": "molecules/s", "cores": len(jobs), "kind": "port",
            "sample": ("oracle (reference algorithm, single-threaded per molecule) on %d molecules of the timed mix "
                       "spread over its size range, %d processes at once for %.0fs: sampled ERI phase + sampled N^4 "
                       "Fock loop x the GPU run's iterations per molecule, extrapolated to %.0fs per molecule on "
                       "average; throughput = processes / mean time" % (len(secs), len(jobs), wall, mean))}


def cpu_baseline_single(bfile, atoms, iterations, budget_s):
    sec, desc = _cpu_molecule_seconds(bfile, atoms, iterations, budget_s)
    return {"value": 1.0 / sec, "unit": "molecules/s", "cores": 1, "kind": "port", "sample": "oracle on one core: " + desc}


# ---------------------------------------------------------------------------------------------- serving
SERVE_SLOTS = 32       # bench.py --slots N changes it
SERVE_LOCKSTEP = 2     # batch threads of the lockstep server (bench.py --lockstep k); 0: one worker thread per slot
SERVE_PREP = None      # preparation threads of the lockstep server (bench.py --prep-threads p; None: the server's 8)
SLOT_MODE_SLOTS = 8    # slots of the slot-thread comparison leg (its measured saturation point is 4-8)


def _pass(srv, bfile, reqs, rate, seed, tag):
    """One pass of reqs through a running server: all queued at once (rate None) or as a Poisson process at rate/s.
    Returns the results by tag and the wall time from the first submission to the last completion."""
    import numpy as np
    rng = np.random.default_rng(seed)
    got, shown = {}, 0
    gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))
    hip_sync()
    t0 = time.perf_counter()
    due = 0.0
    for (rt, atoms), gap in zip(reqs, gaps):
        due += gap
        while time.perf_counter() - t0 < due:
            for r in srv.poll(64, 0):
                got[r.tag] = r
            time.sleep(min(0.0005, max(0.0, due - (time.perf_counter() - t0))))
        srv.submit(bfile, atoms, tag=rt)
    while len(got) < len(reqs):
        for r in srv.poll(64, 1000):
            got[r.tag] = r
        if len(got) // 500 != shown:
            shown = len(got) // 500
            log("serving %s: %d of %d done" % (tag, len(got), len(reqs)))
    hip_sync()
    wall = time.perf_counter() - t0
    log("serving %s: %d requests in %.1f s" % (tag, len(reqs), wall))
    return got, wall


def serve(hf, bfile, reqs, prec, rate=None, slots=None, incore=1, tag="", profile=0, profile_class=-1,
          seed=77, max_iterations=300, deterministic=0, lockstep=None, warm=None, poisson_factor=None):
    """Serve reqs = [(tag, atoms)] on a new server: first the warm-up requests `warm` (untimed: a server's slots
    reach their working sizes), then reqs all queued at once (rate None) or as a Poisson process at `rate`/s.  With
    poisson_factor, reqs run once more on the same server as a Poisson process at that fraction of the rate just
    measured (converged requests / wall time).  Returns the results by tag, the wall time from the first
    submission to the last completion, the kernel statistics of that pass and the Poisson pass's results (or
    None)."""
    ls = SERVE_LOCKSTEP if lockstep is None else lockstep
    srv = hf.Server(slots=slots or SERVE_SLOTS, options=dict(precision=prec, fp32_threshold=1e-6, exact_transfer=1,
                                                             profile=profile, profile_class=profile_class,
                                                             deterministic=deterministic),
                    scf=dict(SCF_KW, incore=incore, max_iterations=max_iterations), lockstep=ls if ls else False,
                    prep_threads=SERVE_PREP)
    try:
        if warm:
            _pass(srv, bfile, warm, None, seed, tag + " (warm-up)")
            srv.kernel_stats(reset=True)
        got, wall = _pass(srv, bfile, reqs, rate, seed, tag)
        stats = srv.kernel_stats()
        extra = None
        if poisson_factor:
            conv = sum(1 for r in got.values() if r.status == 0 and r.converged)
            extra = _pass(srv, bfile, reqs, poisson_factor * conv / wall, seed + 1, tag + " poisson")
    finally:
        srv.close()
    return got, wall, stats, extra


def kernel_kinds(stats):
    """Device time, launches, work and bound per kernel kind (ERI class x precision, in-core contraction FP64/FP32)."""
    out = {}
    for s in stats:
        if s.la < 0:
            bpe = 4.0 if s.fp32 else 8.0
            out[s.name()] = dict(ms=s.ms, launches=s.launches, work=s.eris * bpe, bound="hbm", unit="GB/s",
                                 peak=PEAK_HBM_GBS, kernel="incore_jk_kernel<%s>" % ("float" if s.fp32 else "double"),
                                 key=-2)
        else:
            key = s.la * 1000 + s.lb * 100 + s.lc * 10 + s.ld
            out[s.name()] = dict(ms=s.ms, launches=s.launches, work=s.flops, bound="fp32-valu" if s.fp32 else
                                 "fp64-valu", unit="TFLOP/s", peak=PEAK_FP32_TFLOPS if s.fp32 else PEAK_FP64_TFLOPS,
                                 kernel="eri_class_kernel" + s.name(), key=key)
    return out


def roofline_of(kind, stats):
    k = kernel_kinds(stats).get(kind)
    if not k or k["ms"] <= 0:
        return None
    per_ms = k["ms"] / k["launches"]
    per_work = k["work"] / k["launches"]
    scale = 1e9 if k["unit"] == "GB/s" else 1e12
    achieved = per_work / (per_ms * 1e-3) / scale
    return {"bound": k["bound"], "kernel": k["kernel"], "achieved": achieved, "peak": k["peak"], "unit": k["unit"],
            "frac": achieved / k["peak"], "traffic": None, "launches": k["launches"], "ms_per_launch": per_ms,
            ("bytes_per_launch" if k["unit"] == "GB/s" else "flops_per_launch"): per_work}


def pct(vals, p):
    import numpy as np
    return float(np.percentile(np.asarray(vals), p)) if len(vals) else None


def serving_bench(a):
    import hartreefock_amd as hf
    from hartreefock_amd import geometry as g
    bfile = hf.basis_path("6-31g.1.nw")
    K, W, B = a.steps, a.warmup, a.batch
    mols = g.serving_mix(W * B + K * B, seed=2024)
    warm = [(i, atoms) for i, (_, atoms) in enumerate(mols[:W * B])]
    timed = [(i, atoms) for i, (_, atoms) in enumerate(mols[W * B:], start=W * B)]
    # dominant kernel (untimed): one batch per tier with every kernel timed
    dominant = {}
    for prec in (1, 0):
        _, _, st, _ = serve(hf, bfile, warm[:B], prec, tag="kernel pick %s" % ("mixed" if prec else "fp64"), profile=1)
        kinds = kernel_kinds(st)
        dominant[prec] = max(kinds, key=lambda k: kinds[k]["ms"]) if kinds else None
        log("kernel pick %s: dominant kernel %s (%s)" % ("mixed" if prec else "fp64", dominant[prec],
                                                        ", ".join("%s %.0f ms" % (k, v["ms"]) for k, v in
                                                                  sorted(kinds.items(), key=lambda x: -x[1]["ms"])[:5])))
    # timed region: the mixed tier on a server warmed up with W batches, every timed request queued at once; only the
    # dominant kernel carries events (profile_class: its class key, or -2 for the in-core contraction); then the same
    # requests under Poisson arrivals at 0.8x the measured capacity, on the same server, for latency
    dom_key = None
    if dominant[1]:
        nm, sp = dominant[1], "spd"
        dom_key = -2 if nm.startswith("incore") else sp.index(nm[1]) * 1000 + sp.index(nm[2]) * 100 + \
            sp.index(nm[4]) * 10 + sp.index(nm[5])
    res_m, wall_m, st_m, pm = serve(hf, bfile, timed, 1, tag="timed mixed", profile=1 if dom_key is not None else 0,
                                    profile_class=dom_key if dom_key is not None else -1, warm=warm,
                                    poisson_factor=0.8)
    # the same requests in FP64 (closed loop, then Poisson), on a server warmed up the same way
    res_d, wall_d, st_d, pd = serve(hf, bfile, timed, 0, tag="fp64", warm=warm, poisson_factor=0.8)
    conv_d = {t: r.energy for t, r in res_d.items() if r.status == 0 and r.converged}
    conv_m = {t: r.energy for t, r in res_m.items() if r.status == 0 and r.converged}
    # the |dE| reference: the same requests in FP64 deterministic mode (fixed-point J/K and V sums, whose value
    # does not depend on the order of the GPU's additions); untimed
    res_r, wall_r, _, _ = serve(hf, bfile, timed, 0, tag="fp64 deterministic reference", deterministic=1)
    conv_r = {t: r.energy for t, r in res_r.items() if r.status == 0 and r.converged}
    good_m = [t for t, e in conv_m.items() if t in conv_r and abs(e - conv_r[t]) <= 1e-6]
    dE = [abs(conv_m[t] - conv_r[t]) for t in conv_m if t in conv_r]
    dE_d = [abs(conv_d[t] - conv_r[t]) for t in conv_d if t in conv_r]
    cap_m, cap_d = len(good_m) / wall_m, len(conv_d) / wall_d
    lat = {}
    for name, (r, w) in (("mixed", pm), ("fp64", pd)):
        ms = [x.done_ms - x.arrival_ms for x in r.values()]
        conv = sum(1 for x in r.values() if x.status == 0 and x.converged)
        lat[name] = {"rate_per_s": 0.8 * (cap_m if name == "mixed" else cap_d), "p50_ms": pct(ms, 50),
                     "p99_ms": pct(ms, 99), "converged": conv, "molecules_per_s": conv / w}
    # the slot-thread server (one worker thread and context per slot, each request's SCF alone), FP64, same requests
    slot_leg = None
    if SERVE_LOCKSTEP and not a.no_slot_leg:
        res_s, wall_s, _, _ = serve(hf, bfile, timed, 0, slots=SLOT_MODE_SLOTS, lockstep=0, tag="fp64 slot threads",
                                    warm=warm[:B])
        conv_s = sum(1 for r in res_s.values() if r.status == 0 and r.converged)
        slot_leg = {"molecules_per_s": conv_s / wall_s, "wall_s": wall_s, "converged": conv_s,
                    "slots": SLOT_MODE_SLOTS}
    # the plain FP64 path, one molecule at a time with direct J/K (the north star's reference point)
    plain_reqs = timed[:a.plain_requests]
    res_p, wall_p, _, _ = serve(hf, bfile, plain_reqs, 0, slots=1, incore=0, tag="plain fp64", lockstep=0,
                                warm=warm[:10])
    cap_p = sum(1 for r in res_p.values() if r.status == 0 and r.converged) / wall_p
    eris_m = sum(s.eris for s in st_m if s.la >= 0)
    roof = roofline_of(dominant[1], st_m) if dominant[1] else None
    roof = with_traffic(roof, "hbm_traffic_serving_631g.json")
    nbf = [r.nbf for r in res_m.values()]
    out = {
        "metric": "SCF molecules/s, mixed precision vs FP64 on GPU, at |dE| <= 1e-6 Eh (serving, continuous batching)",
        "value": cap_m, "unit": "molecules/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": wall_m / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64/f32 (mixed tier: FP32 quartets below a 1e-6 Schwarz bound and an FP32 store for density "
                 "changes, FP64 finish)",
        "data": "synthetic (seeded geometries)",
        "config": {"workload": "serving mix: %d RHF requests (%d steps x %d), water clusters 1-12 / n-alkanes C1-C16, "
                               "6-31G, %d-%d functions" % (K * B, K, B, min(nbf), max(nbf)),
                   "batch_per_step": B, "slots": SERVE_SLOTS,
                   "server": ("lockstep: %d batch threads, the SCF steps of each batch's requests advance together, one "
                              "launch per phase for all of them" % SERVE_LOCKSTEP) if SERVE_LOCKSTEP else
                             "slot threads: one worker thread per slot",
                   "jk": "in-core ERI store per request",
                   "eri_convention": "exact (6-31G has no d shells: identical to the reference's)",
                   "schwarz_threshold": 1e-12, "scf": SCF_DESC, "max_iterations": 300,
                   "arrival": "all queued at once (closed loop) in the timed region, on a server warmed up with "
                              "--warmup x --batch requests; Poisson at 0.8x capacity on the same server for latency",
                   "parallelism": "single GPU"},
        "mixed": {"molecules_per_s": cap_m, "wall_s": wall_m, "converged": len(conv_m), "within_1e-6_of_fp64":
                  len(good_m), "max_abs_dE_vs_fp64_Eh": max(dE) if dE else None,
                  "dE_reference": "FP64 deterministic mode, same requests (%d converged)" % len(conv_r),
                  "latency": lat["mixed"], "eri_per_s": eris_m / wall_m},
        "fp64": {"molecules_per_s": cap_d, "wall_s": wall_d, "converged": len(conv_d), "latency": lat["fp64"],
                 "max_abs_dE_vs_fp64_deterministic_Eh": max(dE_d) if dE_d else None,
                 "eri_per_s": sum(s.eris for s in st_d if s.la >= 0) / wall_d},
        "fp64_deterministic_reference": {"wall_s": wall_r, "converged": len(conv_r)},
        "mixed_vs_fp64": cap_m / cap_d,
        "plain_fp64_one_at_a_time": {"molecules_per_s": cap_p, "requests": len(plain_reqs), "jk": "direct", "slots": 1},
        "fp64_slot_threads": slot_leg,
        "mixed_batched_vs_plain_fp64": cap_m / cap_p, "fp64_batched_vs_plain_fp64": cap_d / cap_p,
        "latency_ms": {"p50": lat["mixed"]["p50_ms"], "p99": lat["mixed"]["p99_ms"]},
        "eri_per_s": eris_m / wall_m,
        "not_converged": sorted(set(t for t in res_d if t not in conv_d) | set(t for t in res_m if t not in conv_m)),
        "roofline": roof,
        "dominant_kernel": {"mixed": dominant[1], "fp64": dominant[0]},
    }
    if not a.no_subleg:
        out["decane_cfg2"] = single_molecule("decane", 3, 1, cpu=False)
    if not a.no_cpu_baseline:
        # sample spread over the size range; iterations from the FP64 run of each molecule
        import numpy as np
        P = host_cores()
        order = sorted(conv_d, key=lambda t: res_d[t].nbf)
        pick = [order[int(x)] for x in np.linspace(0, len(order) - 1, 2 * P)]
        sample = [(mols[t][1], res_d[t].iterations) for t in pick]
        out["cpu_baseline"] = cpu_baseline_serving(bfile, sample, a.cpu_budget)
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------------------------- single molecule
def workload(name):
    from hartreefock_amd import geometry as g
    import hartreefock_amd as hf
    if name == "decane":
        return ("C10H22/6-31G* RHF (194 functions)", hf.basis_path("6-31g_st_.1.nw"), g.decane_cfg2())
    if name == "water":
        return ("H2O/STO-3G RHF", hf.basis_path("sto3g.txt"), g.water_cfg1())
    if name == "water32":
        return ("(H2O)32/6-31G** RHF (800 functions)", hf.basis_path("6-31g_st__st_.nw"), g.water32_cfg3())
    if name == "eicosyl":
        return ("C20H41/6-31+G* UHF (462 functions)", hf.basis_path("6-31+g_st_.1.nw"), g.eicosyl_cfg5())
    raise SystemExit("unknown workload " + name)


def timed_scfs(eng, steps, warmup, **extra):
    kw = dict(SCF_KW, **extra)
    for _ in range(warmup):
        eng.scf(check=False, **kw)
    eng.reset_kernel_stats()
    lat, res = [], []
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        res.append(eng.scf(check=False, **kw))
        lat.append(time.perf_counter() - s0)
    hip_sync()
    return time.perf_counter() - t0, lat, res


def single_molecule(name, steps, warmup, cpu=True, cpu_budget=20.0, max_iterations=0, legs=""):
    """One molecule's SCF per step.  Legs: direct FP64 (the plain method), in-core FP64 (the ERIs evaluated once per
    SCF into HBM, inside the step), mixed direct and mixed in-core; each with molecules/s, iterations, energy and
    |dE| against direct FP64 and against the deterministic-mode FP64 direct reference (one untimed SCF).  Roofline of the direct leg's dominant ERI class (HIP events, FP64 VALU) and of the
    in-core contraction (HBM)."""
    import hartreefock_amd as hf
    label, bfile, atoms = workload(name)
    uhf = name == "eicosyl"
    kw = dict(max_iterations=max_iterations or (200 if uhf else 300), unrestricted=uhf)
    all_legs = [("fp64_direct", 0, 0), ("fp64_incore", 0, 1), ("mixed_direct", 1, 0), ("mixed_incore", 1, 1)]
    if legs:
        all_legs = [l for l in all_legs if l[0] in legs.split(",")]
    out = {"workload": label, "scf": SCF_DESC, "schwarz_threshold": 1e-12, "eri_convention": "exact"}
    # the |dE| reference: FP64 direct in deterministic mode (order-independent fixed-point J/K and V sums), so
    # the reference itself does not move with the order of the GPU's atomic additions
    ref = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1, deterministic=1).set_molecule(bfile, atoms)
    log("%s: fp64 deterministic reference" % name)
    t0 = time.perf_counter()
    rr = ref.scf(check=False, **dict(SCF_KW, incore=0, **kw))
    out["fp64_direct_deterministic"] = {"energy_Eh": rr.energy, "iterations": rr.iterations,
                                        "converged": rr.converged, "ms_per_scf": (time.perf_counter() - t0) * 1e3}
    if uhf:
        out["fp64_direct_deterministic"]["s2"] = rr.s2
    ref.close()
    e_ref = rr.energy
    e0 = None
    for leg, prec, inc in all_legs:
        eng = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1, precision=prec,
                        fp32_threshold=1e-6).set_molecule(bfile, atoms)
        out["nbf"] = eng.nbf
        if inc and eng.nbf > 512:
            eng.close()
            out[leg] = "store does not fit (n > 512)"
            continue
        log("%s: %s" % (name, leg))
        el, lat, res = timed_scfs(eng, steps, warmup, incore=inc, **kw)
        r = res[-1]
        if e0 is None:
            e0 = r.energy
        nb = max(1, r.full_builds + r.incremental_builds)
        out[leg] = {"molecules_per_s": steps / el, "ms_per_scf": el / steps * 1e3, "iterations": r.iterations,
                    "converged": r.converged, "energy_Eh": r.energy, "abs_dE_vs_fp64_direct_Eh": abs(r.energy - e0),
                    "abs_dE_vs_fp64_deterministic_Eh": abs(r.energy - e_ref),
                    "jk_ms_per_build": r.jk_ms / nb, "incore": r.incore, "eri_store_ms": r.eri_ms,
                    "fp32_builds": r.fp32_builds, "p50_ms": pct([x * 1e3 for x in lat], 50),
                    "p99_ms": pct([x * 1e3 for x in lat], 99)}
        if uhf:
            out[leg]["s2"] = r.s2
        eng.close()
    # roofline passes: every class timed to find the dominant one, then only that class (launched first in each
    # build, so its events bracket the kernel rather than its wait for CUs); the in-core contraction's events
    if "fp64_direct" in out and isinstance(out["fp64_direct"], dict):
        p = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1, profile=1).set_molecule(bfile, atoms)
        p.scf(check=False, **dict(SCF_KW, **kw))
        kinds = kernel_kinds(p.kernel_stats())
        dom = max((k for k in kinds if not k.startswith("incore")), key=lambda k: kinds[k]["ms"])
        p.close()
        q = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1, profile=1,
                      profile_class=kinds[dom]["key"]).set_molecule(bfile, atoms)
        q.scf(check=False, **dict(SCF_KW, **kw))
        q.reset_kernel_stats()
        q.scf(check=False, **dict(SCF_KW, **kw))
        out["fp64_direct"]["roofline"] = roofline_of(dom, q.kernel_stats())
        out["fp64_direct"]["eri_per_s"] = sum(s.eris for s in q.kernel_stats()) / (out["fp64_direct"]["ms_per_scf"] * 1e-3)
        q.close()
    if "fp64_incore" in out and isinstance(out["fp64_incore"], dict):
        p = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1, profile=1, profile_class=-2).set_molecule(bfile, atoms)
        p.scf(check=False, incore=1, **dict(SCF_KW, **kw))
        p.reset_kernel_stats()
        p.scf(check=False, incore=1, **dict(SCF_KW, **kw))
        out["fp64_incore"]["roofline"] = with_traffic(roofline_of("incore_jk", p.kernel_stats()),
                                                      "hbm_traffic_incore_decane_631gs.json")
        p.close()
    if cpu and isinstance(out.get("fp64_direct"), dict):
        out["cpu_baseline"] = cpu_baseline_single(bfile, atoms, out["fp64_direct"]["iterations"], cpu_budget)
    return out


def serving_batch(a):
    """The serving mix solved by the lockstep batch entry point (hf_scf_batch) from one host thread: --steps batches of
    --batch molecules (after --warmup untimed batches), each batch one hf_scf_batch call on contexts kept between
    batches, in-core stores, mixed or FP64 (--precision).  One host thread issues everything, so a kernel trace sees
    exactly the lockstep phases, the in-core contraction and the store evaluation."""
    import hartreefock_amd as hf
    from hartreefock_amd import geometry as g
    bfile = hf.basis_path("6-31g.1.nw")
    K, W, B = a.steps, a.warmup, a.batch
    mols = g.serving_mix((W + K) * B, seed=2024)
    engs = [hf.Engine(precision=a.precision, fp32_threshold=1e-6, exact_transfer=1) for _ in range(B)]
    kw = dict(SCF_KW, incore=1, max_iterations=300)
    out_res = []

    def run(chunk):
        for e, (_, atoms) in zip(engs, chunk):
            e.set_molecule(bfile, atoms)
        return hf.scf_batch(engs[:len(chunk)], **kw)

    for w in range(W):
        run(mols[w * B:(w + 1) * B])
    hip_sync()
    t0 = time.perf_counter()
    for k in range(K):
        out_res += run(mols[(W + k) * B:(W + k + 1) * B])
    hip_sync()
    wall = time.perf_counter() - t0
    conv = sum(1 for st, r in out_res if st == 0 and r.converged)
    its = [r.iterations for st, r in out_res if st == 0 and r.converged]
    store_ms = [r.eri_ms for st, r in out_res if st == 0 and r.incore]
    f32 = [r.fp32_builds for st, r in out_res if st == 0 and r.converged]
    # the in-core contraction's launches and algorithmic bytes (the store rows it streams) over every batch,
    # warm-up included, as a counter run of this command sees them (--pmc-summary takes this line)
    incore = {}
    for e in engs:
        for x in e.kernel_stats():
            if x.la < 0:
                k = incore.setdefault("incore_jk_kernel<%s>" % ("float" if x.fp32 else "double"),
                                      {"launches": 0, "bytes": 0.0})
                k["launches"] += x.launches
                k["bytes"] += x.eris * (4.0 if x.fp32 else 8.0)
    print(json.dumps({"metric": "SCF molecules/s, serving mix through hf_scf_batch (one host thread)",
                      "value": conv / wall, "unit": "molecules/s", "steps": K, "warmup": W, "batch": B,
                      "precision": "mixed" if a.precision else "fp64", "wall_s": wall, "converged": conv,
                      "mean_iterations": sum(its) / max(1, len(its)),
                      "mean_fp32_builds": sum(f32) / max(1, len(f32)),
                      "mean_store_ms": sum(store_ms) / max(1, len(store_ms)), "incore_stats": incore}), flush=True)
    for e in engs:
        e.close()


def incore_shape_sweep():
    """In-core contraction: every tile shape (hf_options.incore_shape) at sizes across the serving range and C10H22,
    one and two densities; HIP-event time per launch and the store bytes over it (GB/s), printed as one JSON line."""
    import numpy as np
    import hartreefock_amd as hf
    from hartreefock_amd import geometry as g
    cases = [("(H2O)4/6-31G", "6-31g.1.nw", g.water_cluster(4, seed=1)),
             ("C6H14/6-31G", "6-31g.1.nw", g.alkane_zigzag(6, seed=1)),
             ("(H2O)8/6-31G", "6-31g.1.nw", g.water_cluster(8, seed=1)),
             ("C11H24/6-31G", "6-31g.1.nw", g.alkane_zigzag(11, seed=1)),
             ("C10H22/6-31G*", "6-31g_st_.1.nw", g.decane_cfg2()),
             ("C16H34/6-31G", "6-31g.1.nw", g.alkane_zigzag(16, seed=1)),
             ("(H2O)24/6-31G", "6-31g.1.nw", g.water_cluster(24, seed=42))]
    rows = []
    for label, b, atoms in cases:
        for shape in range(8):
            eng = hf.Engine(schwarz_threshold=1e-12, profile=1, profile_class=-2,
                            incore_shape=shape).set_molecule(hf.basis_path(b), atoms)
            n = eng.nbf
            store = hf.lib().hf_pair_matrix_len(n) * 8.0
            rng = np.random.default_rng(1)
            X = rng.normal(size=(2, n, n)) * 0.1
            D2 = X + X.transpose(0, 2, 1)
            for nd in (1, 2):
                D = D2 if nd == 2 else D2[0]
                for _ in range(2):
                    eng.build_jk_incore(D)
                eng.reset_kernel_stats()
                for _ in range(10):
                    eng.build_jk_incore(D)
                st = [x for x in eng.kernel_stats() if x.la < 0 and not x.fp32][0]
                ms = st.ms / st.launches
                rows.append({"molecule": label, "nbf": n, "densities": nd, "shape": shape, "us_per_launch": ms * 1e3,
                             "store_bytes": store, "GB_per_s": store / (ms * 1e-3) / 1e9})
            eng.close()
            log("%s shape %d: %s" % (label, shape, ", ".join("nd %d %.0f GB/s" % (r["densities"], r["GB_per_s"])
                                                            for r in rows[-2:])))
    print(json.dumps({"incore_shape_sweep": rows, "peak_GB_per_s": PEAK_HBM_GBS}), flush=True)


def _pair_pad_base(k):
    return k * (k + 2) // 2 if k % 2 == 0 else (k + 1) * (k + 1) // 2


def _blocks_before(i, nr):
    q, r = divmod(i, nr)
    return nr * q * (q + 1) // 2 + r * (q + 1)


def _counter_rows(root, counter):
    import csv
    import glob
    rows = []
    for f in sorted(glob.glob(os.path.join(root, "**", "*counter_collection*.csv"), recursive=True)):
        with open(f) as fh:
            rows += [r for r in csv.DictReader(fh) if r.get("Counter_Name") == counter]
    if not rows:
        raise SystemExit("no %s rows under %s" % (counter, root))
    return rows


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, workload_json=None):
    """HBM traffic of the in-core contraction from rocprofv3 --pmc runs (FETCH_SIZE and WRITE_SIZE in KiB per
    dispatch), scaled by what the same counters read on the plain 1 GiB device copy of --copy-calibration.  The
    kernel's algorithmic bytes (the store rows it streams, 8 or 4 bytes per element) come from the profiled command's
    own JSON line when it carries them (workload_json: the serving-batch workload's "incore_stats", launches and
    bytes over the whole command, batched launches covering several stores); otherwise from each single-store
    dispatch's grid (n from the blocks of NR store rows, NR the sixth template argument), as the whole store np^2."""
    import re
    gib_kib = float(1 << 20)
    scale = {}
    for counter, d in (("FETCH_SIZE", copy_fetch_dir), ("WRITE_SIZE", copy_write_dir)):
        v = [float(r["Counter_Value"]) for r in _counter_rows(d, counter) if "copyBuffer" in r["Kernel_Name"]]
        scale[counter] = gib_kib / (sum(v) / len(v))
    alg_total = {}
    if workload_json:
        with open(workload_json) as fh:
            line = [ln for ln in fh if ln.lstrip().startswith("{")][-1]
        alg_total = json.loads(line).get("incore_stats", {})
    out = {"copy_scale": scale, "kernels": {}}
    for counter, d, key in (("FETCH_SIZE", fetch_dir, "fetch"), ("WRITE_SIZE", write_dir, "write")):
        for r in _counter_rows(d, counter):
            m = re.search(r"incore_jk_(batch_)?kernel<(\d+), (\d+), (\d+), (\d+), (true|false), (\d+), (double|float)>",
                          r["Kernel_Name"])
            if not m:
                continue
            nr, t = int(m.group(7)), m.group(8)
            name = "incore_jk_kernel<%s>" % t
            k = out["kernels"].setdefault(name, {"fetch_bytes": 0.0, "fetch_alg": 0.0, "write_bytes": 0.0,
                                                 "write_alg": 0.0, "dispatches": 0, "batched_dispatches": 0})
            if not alg_total:
                if m.group(1):
                    raise SystemExit("batched in-core dispatches: give the profiled command's JSON line")
                blocks = int(r["Grid_Size"]) // 256
                n = next(x for x in range(1, 513) if _blocks_before(x, nr) >= blocks)
                k[key + "_alg"] += _pair_pad_base(n) ** 2 * (8 if t == "double" else 4)
            k[key + "_bytes"] += float(r["Counter_Value"]) * 1024.0 * scale[counter]
            if key == "fetch":
                k["dispatches"] += 1
                k["batched_dispatches"] += 1 if m.group(1) else 0
    for name, k in out["kernels"].items():
        if alg_total:
            a = alg_total.get(name, {})
            if not a.get("launches") or a["launches"] > k["dispatches"]:
                raise SystemExit("%s: %s dispatches counted, %s launches in the command's statistics"
                                 % (name, k["dispatches"], a.get("launches")))
            # the autotuner's timing runs (the first build of each kind of store in the process) are dispatches the
            # statistics do not count: they are taken at the command's mean bytes per launch
            k["tuning_dispatches"] = k["dispatches"] - a["launches"]
            k["fetch_alg"] = k["write_alg"] = a["bytes"] * k["dispatches"] / a["launches"]
        k["fetch_ratio"] = k["fetch_bytes"] / k["fetch_alg"] if k["fetch_alg"] else None
        k["write_ratio"] = k["write_bytes"] / k["write_alg"] if k["write_alg"] else None
        k["traffic_ratio"] = (k["fetch_ratio"] or 0.0) + (k["write_ratio"] or 0.0)
        k["mean_alg_bytes_per_dispatch"] = k["fetch_alg"] / max(1, k["dispatches"])
        k["mean_traffic_bytes_per_dispatch"] = (k["fetch_bytes"] + k["write_bytes"]) / max(1, k["dispatches"])
        k["alg_bytes"] = "the store rows streamed (command statistics)" if alg_total else "the whole store (grid)"
    print(json.dumps(out), flush=True)


def pmc_table(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir):
    """HBM bytes per dispatch of every kernel from rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch), each scaled by what the counter reads on the plain 1 GiB device copy of --copy-calibration; a text table
    of the median over each kernel's dispatches (robust to a minority of atypical launches of the same kernel, such as
    the fixed-point accumulation of a deterministic-mode SCF in the same command) and the mean, largest write first."""
    gib_kib = float(1 << 20)
    scale = {}
    for counter, d in (("FETCH_SIZE", copy_fetch_dir), ("WRITE_SIZE", copy_write_dir)):
        v = [float(r["Counter_Value"]) for r in _counter_rows(d, counter) if "copyBuffer" in r["Kernel_Name"]]
        scale[counter] = gib_kib / (sum(v) / len(v))
    agg = {}
    for counter, d, key in (("FETCH_SIZE", fetch_dir, 0), ("WRITE_SIZE", write_dir, 1)):
        for r in _counter_rows(d, counter):
            name = r["Kernel_Name"].replace("(anonymous namespace)::", "").replace("void ", "").replace("hfgpu::", "")
            name = name.split("(")[0] if not name.startswith("(") else name
            agg.setdefault(name, ([], []))[key].append(float(r["Counter_Value"]) * 1024.0 * scale[counter])
    print("# copy scale: FETCH_SIZE %.4f, WRITE_SIZE %.4f; bytes per dispatch" % (scale["FETCH_SIZE"],
                                                                                  scale["WRITE_SIZE"]))
    print("%-56s %8s %11s %11s %11s %11s" % ("kernel", "dispatch", "fetch med", "write med", "fetch mean",
                                             "write mean"))

    def med(v):
        v = sorted(v)
        return 0.5 * (v[(len(v) - 1) // 2] + v[len(v) // 2])

    rows = sorted(((n, f, w) for n, (f, w) in agg.items() if f and w), key=lambda x: -med(x[2]))
    for name, f, w in rows:
        print("%-56s %8d %11.4g %11.4g %11.4g %11.4g" % (name[:56], len(f), med(f), med(w), sum(f) / len(f),
                                                           sum(w) / len(w)))


def trace_summary(root, table=False):
    """Per-kernel resources and durations from a rocprofv3 --kernel-trace directory (kernel_trace.csv): calls, mean
    duration, VGPRs (arch + accumulation), SGPRs, LDS and scratch per workgroup/lane, workgroup size, and the
    wavefronts per SIMD that the VGPR count allows (512 VGPRs per SIMD lane, granule 8, at most 8); also the trace's
    span and the time the device ran any kernel (the union of the kernels' intervals).  table: a text table of time
    by kernel instead of the JSON line."""
    import csv
    import glob
    agg = {}
    iv = []
    for f in sorted(glob.glob(os.path.join(root, "**", "*kernel_trace.csv"), recursive=True)):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                iv.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
                name = r["Kernel_Name"]
                k = agg.setdefault(name, {"calls": 0, "ns": 0.0, "vgpr": int(r.get("VGPR_Count", 0) or 0),
                                          "agpr": int(r.get("Accum_VGPR_Count", 0) or 0),
                                          "sgpr": int(r.get("SGPR_Count", 0) or 0),
                                          "lds": int(r.get("LDS_Block_Size", 0) or 0),
                                          "scratch": int(r.get("Scratch_Size", 0) or 0),
                                          "wg": int(r.get("Workgroup_Size", 0) or 0), "grid_max": 0})
                k["calls"] += 1
                k["ns"] += float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
                k["grid_max"] = max(k["grid_max"], int(r.get("Grid_Size", 0) or 0))
    rows = []
    for name, k in agg.items():
        regs = max(8, -(-(k["vgpr"] + k["agpr"]) // 8) * 8)
        rows.append(dict(kernel=name, calls=k["calls"], mean_us=k["ns"] / k["calls"] / 1e3, vgpr=k["vgpr"],
                         agpr=k["agpr"], sgpr=k["sgpr"], lds_bytes=k["lds"], scratch_bytes_per_lane=k["scratch"],
                         workgroup=k["wg"], max_grid=k["grid_max"], waves_per_simd=min(8, 512 // regs)))
    rows.sort(key=lambda x: -x["calls"] * x["mean_us"])
    iv.sort()
    busy, span = 0, 0
    if iv:
        cs, ce = iv[0]
        for a, b in iv[1:]:
            if a > ce:
                busy += ce - cs
                cs, ce = a, b
            else:
                ce = max(ce, b)
        busy += ce - cs
        span = max(b for _, b in iv) - iv[0][0]
    if not table:
        print(json.dumps({"span_ms": span / 1e6, "busy_ms": busy / 1e6, "kernels": rows}), flush=True)
        return
    total = sum(x["calls"] * x["mean_us"] for x in rows) / 1e3
    print("# kernel%s calls    total_ms   mean_us  share" % (" " * 62))
    print("kernels %d, total %.1f ms" % (sum(x["calls"] for x in rows), total))
    for x in rows[:40]:
        name = x["kernel"].replace("(anonymous namespace)::", "").replace("void ", "").replace("hfgpu::", "")
        name = name.split("(")[0] if not name.startswith("(") else name
        ms = x["calls"] * x["mean_us"] / 1e3
        print("%-70s %7d %9.2f ms %8.1f us %5.1f%%" % (name[:70], x["calls"], ms, x["mean_us"], 100.0 * ms / total))
    print("span %.1f ms, device busy (union of kernels) %.1f ms = %.0f%%" % (span / 1e6, busy / 1e6,
                                                                             100.0 * busy / max(span, 1)))


def with_traffic(roof, fname):
    """roofline.traffic = the kernel's measured HBM bytes over its algorithmic bytes (profiles/fname: counter runs of
    a workload of the same kind) times this run's algorithmic bytes per launch; for a summary whose algorithmic bytes
    are whole stores (the rows the screening zeroed not left out), its mean HBM bytes per dispatch instead."""
    path = os.path.join(ROOT, "profiles", fname)
    if roof and roof.get("unit") == "GB/s" and os.path.exists(path):
        with open(path) as fh:
            rec = json.load(fh).get("kernels", {}).get(roof["kernel"])
        if rec and rec.get("alg_bytes", "").startswith("the store rows") and rec.get("traffic_ratio"):
            roof["traffic"] = rec["traffic_ratio"] * roof["bytes_per_launch"]
            roof["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of the same kind of workload: %.3f x "
                                      "the streamed store bytes, profiles/%s" % (rec["traffic_ratio"], fname))
        elif rec and rec.get("mean_traffic_bytes_per_dispatch"):
            roof["traffic"] = rec["mean_traffic_bytes_per_dispatch"]
            roof["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of the same workload, mean HBM bytes "
                                      "per dispatch (%.3f x this run's streamed bytes per launch), profiles/%s"
                                      % (rec["mean_traffic_bytes_per_dispatch"] / roof["bytes_per_launch"], fname))
    return roof


def copy_calibration():
    """A plain device copy of known size: 1 GiB read + 1 GiB written per launch, three launches.
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this gives the counters' scale for HBM bytes."""
    import torch
    n = (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="serving: batches of --batch requests in the timed region")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=100, help="serving: requests per step")
    ap.add_argument("--slots", type=int, default=32, help="serving: concurrent requests (server slots)")
    ap.add_argument("--lockstep", type=int, default=2, help="serving: batch threads of the lockstep server (default 2); "
                    "0: one worker thread per slot")
    ap.add_argument("--prep-threads", type=int, default=None,
                    help="serving: preparation threads of the lockstep server (default: the server's 8)")
    ap.add_argument("--no-slot-leg", action="store_true", help="serving: skip the FP64 slot-thread comparison leg")
    ap.add_argument("--workload", default="serving", help="serving (default) | decane | water | water32 | eicosyl | "
                    "incore-shapes (in-core contraction per tile shape and size) | serving-batch (the serving mix "
                    "through hf_scf_batch from one host thread)")
    ap.add_argument("--precision", type=int, default=1, help="serving-batch: 1 mixed tier (default), 0 FP64")
    ap.add_argument("--plain-requests", type=int, default=300,
                    help="serving: requests of the one-at-a-time plain FP64 direct run")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU work for the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-subleg", action="store_true", help="serving: skip the C10H22 single-molecule sub-leg")
    ap.add_argument("--legs", default="", help="single-molecule workloads: fp64_direct,fp64_incore,mixed_direct,"
                    "mixed_incore (default all)")
    ap.add_argument("--max-iterations", type=int, default=0)
    ap.add_argument("--pmc-summary", nargs="+", metavar="DIR",
                    help="only summarise rocprofv3 --pmc output: FETCH_DIR WRITE_DIR COPY_FETCH_DIR COPY_WRITE_DIR "
                         "[WORKLOAD_JSON]: in-core contraction HBM traffic against its algorithmic bytes (from the "
                         "profiled command's JSON line when given), scaled by the copy calibration")
    ap.add_argument("--pmc-table", nargs=4, metavar="DIR",
                    help="FETCH_SIZE, WRITE_SIZE, copy FETCH_SIZE and copy WRITE_SIZE rocprofv3 --pmc directories: "
                         "HBM bytes per dispatch of every kernel, as a text table")
    ap.add_argument("--trace-table", metavar="TRACE_DIR",
                    help="a rocprofv3 --kernel-trace directory: text table of kernel time by kernel, span and busy time")
    ap.add_argument("--trace-summary", metavar="TRACE_DIR",
                    help="only summarise a rocprofv3 --kernel-trace directory: per-kernel resources and durations")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration for rocprofv3 --pmc)")
    a = ap.parse_args()
    if a.pmc_table:
        pmc_table(*a.pmc_table)
        return
    if a.trace_summary or a.trace_table:
        trace_summary(a.trace_summary or a.trace_table, table=bool(a.trace_table))
        return
    if a.pmc_summary:
        if len(a.pmc_summary) not in (4, 5):
            raise SystemExit("--pmc-summary takes 4 directories and optionally the workload's JSON output")
        pmc_summary(*a.pmc_summary)
        return
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (serving and single-molecule SCF per GPU; multi-GPU J/K "
              "sharding is a design note, not implemented). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "incore-shapes":
        incore_shape_sweep()
        return
    if a.workload == "serving-batch":
        serving_batch(a)
        return
    if a.workload == "serving":
        global SERVE_SLOTS, SERVE_LOCKSTEP, SERVE_PREP
        SERVE_PREP = a.prep_threads
        SERVE_SLOTS = a.slots
        SERVE_LOCKSTEP = a.lockstep
        serving_bench(a)
        return
    res = single_molecule(a.workload, a.steps, a.warmup, cpu=not a.no_cpu_baseline, cpu_budget=a.cpu_budget,
                          max_iterations=a.max_iterations, legs=a.legs)
    d = res["fp64_direct"] if isinstance(res.get("fp64_direct"), dict) else next(
        v for v in res.values() if isinstance(v, dict) and "molecules_per_s" in v)
    out = {"metric": "SCF molecules/s (plain FP64 direct on GPU)", "value": d["molecules_per_s"], "unit": "molecules/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": d["ms_per_scf"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded geometry)",
           "config": {"workload": res["workload"], "nbf": res.get("nbf"), "scf": SCF_DESC, "eri_convention": "exact",
                      "schwarz_threshold": 1e-12, "parallelism": "single GPU"},
           "roofline": d.get("roofline"), "cpu_baseline": res.pop("cpu_baseline", None), "legs": res}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
