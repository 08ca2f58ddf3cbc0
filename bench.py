"""Benchmark: SCF molecules/s on one MI355X.

Workload (default): plain direct RHF in FP64 on the GPU for n-decane C10H22 in 6-31G*
(194 Cartesian functions), seeded zigzag geometry (C-C 1.54 A, C-H 1.09 A, tetrahedral
angles, +-0.05 A jitter), Schwarz threshold 1e-12, converged to 1e-8 Eh.  A step is one
complete SCF solve (integrals recomputed every iteration: direct J/K).  The molecule's
shells, primitive pairs and Schwarz table are resident in HBM before timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload decane|water|serving|water32|eicosyl]

Prints ONE JSON line (rank 0).  `roofline` is the dominant ERI class kernel: algorithmic
flops (exact operation count of the generated code, per launch) over its average launch
duration from HIP events recorded on its launch stream during the timed steps.
`cpu_baseline` times the oracle (a CPU restatement of the reference's algorithm:
per-component HGP ERIs into the 8-fold packed store, then an unscreened N^4 Fock loop
per iteration) on a bounded sample of the same workload, single core, and extrapolates
to one full SCF with the GPU run's iteration count.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def hip_sync():
    try:
        ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()
    except OSError:
        pass


def workload(name):
    from hartreefock_amd import geometry as g
    import hartreefock_amd as hf
    if name == "decane":
        return ("C10H22/6-31G* direct RHF FP64", hf.basis_path("6-31g_st_.1.nw"), g.decane_cfg2())
    if name == "water":
        return ("H2O/STO-3G RHF FP64", hf.basis_path("sto3g.txt"), g.water_cfg1())
    if name == "water32":
        return ("(H2O)32/6-31G** direct RHF FP64", hf.basis_path("6-31g_st__st_.nw"), g.water32_cfg3())
    if name == "eicosyl":
        return ("C20H41/6-31+G* UHF FP64", hf.basis_path("6-31+g_st_.1.nw"), g.eicosyl_cfg5())
    raise SystemExit("unknown workload " + name)


SCF_KW = dict(alpha=0.5, init_guess=0.0, use_diis=1, normal_iter_after_diis=0,
              energy_tolerance=1e-8, density_tolerance=1e-6, diis_tolerance=1e-6)


def cpu_baseline(bfile, atoms, iterations, budget_s):
    """Oracle (reference algorithm) on one core: sampled ERI phase + sampled N^4 Fock loop."""
    import numpy as np
    from oracle import refcpu
    mol = refcpu.Molecule(bfile, atoms)
    n = mol.nbf
    unique = refcpu.packed_index(n - 1, n - 1, n - 1, n - 1) + 1
    triples = n * (n + 1) // 2 * n
    # ERI phase: every stride-th (i,j,k) triple with its whole l loop (~60% of the budget)
    stride = max(1, int(triples / max(1.0, 0.6 * budget_s / 19e-6 / (n / 2.0))))
    t0 = time.perf_counter()
    eri, cnt = mol.eri_packed_sample(stride, stride // 2)
    t_eri = time.perf_counter() - t0
    eri_phase = t_eri / max(cnt, 1) * unique
    # Fock loop: rows of the reference's G loop (~40% of the budget at ~0.1 s/row for n=194)
    rows = max(2, min(n, int(0.4 * budget_s / (20.0 / 194.0 * (n / 194.0) ** 3))))
    D = np.eye(n) * 0.1
    t0 = time.perf_counter()
    refcpu.rhf_g(n, eri, D, rows=(0, rows))
    t_g = time.perf_counter() - t0
    fock = t_g * n / rows
    total = eri_phase + iterations * fock
    return {"value": 1.0 / total, "unit": "molecules/s", "cores": 1, "kind": "port",
            "sample": ("oracle on %d of %d unique ERIs (every %d-th (i,j,k) triple, %.1fs -> ERI phase %.0fs) + "
                       "%d of %d rows of one N^4 Fock loop (%.1fs -> %.1fs/iteration) x %d iterations = %.0fs/molecule"
                       % (cnt, unique, stride, t_eri, eri_phase, rows, n, t_g, fock, iterations, total))}


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


def serving(a):
    """Serving configuration: a seeded mix of water clusters (1-12) and alkanes (C1-C16) in 6-31G,
    RHF to 1e-8 Eh.  Phase 1 measures the server's capacity with every request queued at once;
    phase 2 submits the requests as a Poisson process at 0.8 x that capacity (recorded seed) and
    reports molecules/s (completions over first arrival to last completion) and the p50/p99 of
    completion - arrival, for FP64 and for the mixed-precision tier."""
    import numpy as np
    import hartreefock_amd as hf
    from hartreefock_amd import geometry as g
    bfile = hf.basis_path("6-31g.1.nw")
    scf = dict(SCF_KW, max_iterations=300)   # a request that has not converged by then is reported as failed
    mols = g.serving_mix(a.requests, seed=2024)
    rng = np.random.default_rng(77)

    def run(prec, rate, tag, slots=None, incore=None, reqs=None):
        shown = 0
        reqs = mols if reqs is None else reqs
        srv = hf.Server(slots=slots or a.slots, options=dict(precision=prec, fp32_threshold=1e-6, exact_transfer=1),
                        scf=dict(scf, incore=a.serve_incore if incore is None else incore))
        try:
            gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))
            got = []
            t0 = time.perf_counter()
            due = 0.0
            for (name, atoms), gap in zip(reqs, gaps):
                due += gap
                while time.perf_counter() - t0 < due:
                    got += srv.poll(64, 0)
                    time.sleep(min(0.001, max(0.0, due - (time.perf_counter() - t0))))
                srv.submit(bfile, atoms)
            while len(got) < len(reqs):
                got += srv.poll(64, 1000)
                if len(got) // 250 != shown:
                    shown = len(got) // 250
                    print("serving %s: %d of %d done" % (tag, len(got), len(reqs)), file=sys.stderr, flush=True)
        finally:
            srv.close()
        print("serving %s: finished in %.1f s" % (tag, time.perf_counter() - t0), file=sys.stderr, flush=True)
        bad = [{"id": r.id, "nbf": r.nbf, "status": r.status, "iterations": r.iterations, "energy": r.energy}
               for r in got if r.status != 0 or not r.converged]
        first = min(r.arrival_ms for r in got)
        last = max(r.done_ms for r in got)
        lat = np.array(sorted(r.done_ms - r.arrival_ms for r in got))
        nbf = [r.nbf for r in got]
        return {"molecules_per_s": len(got) / ((last - first) * 1e-3), "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)), "not_converged": bad, "nbf_min": min(nbf), "nbf_max": max(nbf),
                "energies": {r.id: r.energy for r in got if r.status == 0 and r.converged}}

    # the plain FP64 path: one molecule at a time, direct J/K, on the first plain_requests requests
    plain = run(0, None, "plain", slots=1, incore=0, reqs=mols[:a.plain_requests])["molecules_per_s"]
    cap = run(0, None, "capacity")["molecules_per_s"]
    cap_m = run(1, None, "capacity mixed")["molecules_per_s"]
    rate = 0.8 * cap
    r64 = run(0, rate, "fp64")
    r32 = run(1, rate, "mixed")
    common = set(r64["energies"]) & set(r32["energies"])
    dmax = max(abs(r64["energies"][k] - r32["energies"][k]) for k in common)
    for r in (r64, r32):
        r.pop("energies")
    out = {"metric": "SCF molecules/s under Poisson arrivals (serving, continuous batching)",
           "value": r64["molecules_per_s"], "unit": "molecules/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "f64", "data": "synthetic (seeded geometries)",
           "config": {"workload": "serving mix: %d RHF requests, water clusters 1-12 / alkanes C1-C16, 6-31G" % a.requests,
                      "slots": a.slots, "arrival": "Poisson, rate 0.8 x closed-loop capacity", "seed": 77,
                      "jk": "in-core ERI store" if a.serve_incore else "direct",
                      "closed_loop_capacity": cap, "rate": rate},
           "plain_fp64_one_at_a_time": {"molecules_per_s": plain, "requests": a.plain_requests,
                                        "jk": "direct", "slots": 1},
           "capacity": {"fp64": cap, "mixed": cap_m},
           "mixed_batched_vs_plain_fp64": cap_m / plain, "fp64_batched_vs_plain_fp64": cap / plain,
           "fp64": r64, "mixed": r32, "mixed_vs_fp64": r32["molecules_per_s"] / r64["molecules_per_s"],
           "max_abs_dE_mixed_vs_fp64_Eh": dmax}
    print(json.dumps(out), flush=True)


def large(a):
    """The large single-molecule configurations, one SCF per leg: (H2O)32/6-31G** RHF (800 functions) and
    C20H41/6-31+G* UHF (462 functions, 200 iterations at most: with the reference's SCF settings it does
    not meet 1e-8 Eh).  Legs: direct FP64, in-core FP64 (when the store fits: n <= 512), mixed direct and
    mixed in-core; each reports wall time, iterations, energy and |dE| against direct FP64."""
    import hartreefock_amd as hf
    label, bfile, atoms = workload(a.workload)
    uhf = a.workload == "eicosyl"
    kw = dict(SCF_KW, max_iterations=a.max_iterations or (200 if uhf else 300))
    legs = [("fp64_direct", 0, 0), ("fp64_incore", 0, 1), ("mixed_direct", 1, 0), ("mixed_incore", 1, 1)]
    if a.legs:
        legs = [l for l in legs if l[0] in a.legs.split(",")]
    out = {"metric": "SCF wall time per molecule", "unit": "s", "higher_is_better": False, "n_gpus": 1,
           "dtype": "f64", "data": "synthetic (seeded geometry)",
           "config": {"workload": label, "schwarz_threshold": 1e-12, "energy_tolerance": 1e-8,
                      "eri_convention": "exact", "scf": "alpha=0.5, core guess, DIIS(6/5), normalIterAfterDIIS=0"}}
    e0 = None
    for name, prec, inc in legs:
        eng = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1, precision=prec,
                        fp32_threshold=1e-6).set_molecule(bfile, atoms)
        out["config"]["nbf"] = eng.nbf
        if inc and eng.nbf > 512:
            eng.close()
            out[name] = "store does not fit (n > 512)"
            continue
        print("large %s: %s" % (a.workload, name), file=sys.stderr, flush=True)
        hip_sync()
        t0 = time.perf_counter()
        r = eng.scf(unrestricted=uhf, check=False, incore=inc, **kw)
        hip_sync()
        el = time.perf_counter() - t0
        eng.close()
        if e0 is None:
            e0 = r.energy
        out[name] = {"s": el, "iterations": r.iterations, "converged": r.converged, "energy_Eh": r.energy,
                     "abs_dE_vs_fp64_direct_Eh": abs(r.energy - e0), "incore": r.incore, "eri_store_ms": r.eri_ms,
                     "jk_ms": r.jk_ms, "fp32_builds": r.fp32_builds}
        if uhf:
            out[name]["s2"] = r.s2
    first = next(l[0] for l in legs if isinstance(out.get(l[0]), dict))
    out["value"] = out[first]["s"]
    out["config"]["max_iterations"] = kw["max_iterations"]
    print(json.dumps(out), flush=True)


def timed_scfs(eng, steps, warmup, **extra):
    kw = dict(SCF_KW, **extra)
    for _ in range(warmup):
        eng.scf(**kw)
    eng.reset_kernel_stats()
    lat, res = [], []
    hip_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        r = eng.scf(**kw)
        lat.append(time.perf_counter() - s0)
        res.append(r)
    hip_sync()
    return time.perf_counter() - t0, lat, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="decane")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU work for the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--requests", type=int, default=2000, help="serving workload: number of requests")
    ap.add_argument("--slots", type=int, default=8, help="serving workload: concurrent requests")
    ap.add_argument("--legs", default="", help="water32/eicosyl: comma list of legs (fp64_direct, fp64_incore, "
                    "mixed_direct, mixed_incore); default all")
    ap.add_argument("--max-iterations", type=int, default=0, help="water32/eicosyl: SCF iteration cap (0: 300 / 200)")
    ap.add_argument("--plain-requests", type=int, default=300,
                    help="serving workload: requests of the one-at-a-time plain FP64 reference run")
    ap.add_argument("--serve-incore", type=int, default=1,
                    help="serving workload: 1 = each request's ERIs evaluated once into an in-core store (0: direct)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration for rocprofv3 --pmc)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (one molecule's SCF per GPU; "
              "multi-GPU J/K sharding is a design note, not implemented). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if a.workload == "serving":
        serving(a)
        return
    if a.workload in ("water32", "eicosyl"):
        large(a)
        return
    import hartreefock_amd as hf
    label, bfile, atoms = workload(a.workload)
    # exact integrals: the reference's (dd|dd) transfer convention makes this SCF diverge
    # untimed probe SCF with every class kernel timed: contracted ERIs per SCF, the kernels' share
    # and the dominant class (HIP events around all 21 launches of a build cost it several per cent)
    probe = hf.Engine(schwarz_threshold=1e-12, profile=1, exact_transfer=1).set_molecule(bfile, atoms)
    probe.scf(**SCF_KW)
    probe.reset_kernel_stats()
    t0 = time.perf_counter()
    probe.scf(**SCF_KW)
    probe_s = time.perf_counter() - t0
    pstats = probe.kernel_stats()
    eris_per_scf = sum(s.eris for s in pstats)
    kernel_share = sum(s.ms for s in pstats) * 1e-3 / probe_s
    top = max(pstats, key=lambda s: s.ms)
    key = top.la * 1000 + top.lb * 100 + top.lc * 10 + top.ld
    probe.close()
    # timed region: no events
    eng = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1).set_molecule(bfile, atoms)
    elapsed, lat, res = timed_scfs(eng, a.steps, a.warmup)
    energies = [r.energy for r in res]
    iters = [r.iterations for r in res]
    nbf = eng.nbf
    eng.close()
    # roofline pass (two more SCFs): only the dominant class carries HIP events, and it is launched
    # first in each build, so the events bracket its execution rather than a wait for CUs
    rl = hf.Engine(schwarz_threshold=1e-12, profile=1, profile_class=key, exact_transfer=1).set_molecule(bfile, atoms)
    rl.scf(**SCF_KW)
    rl.reset_kernel_stats()
    rl.scf(**SCF_KW)
    dom = [s for s in rl.kernel_stats() if s.ms > 0 and not s.fp32][0]
    rl.close()
    per_launch_flops = dom.flops / dom.launches
    per_launch_ms = dom.ms / dom.launches
    achieved = per_launch_flops / (per_launch_ms * 1e-3) / 1e12
    lat_sorted = sorted(lat)
    # HBM bytes per launch of the dominant kernel: FETCH_SIZE + WRITE_SIZE from the committed
    # rocprofv3 --pmc runs on this workload (calibrated against a 1 GiB copy), when available
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic_decane_631gs.json")
    if a.workload == "decane" and os.path.exists(tfile):
        with open(tfile) as fh:
            rec = json.load(fh)["kernels"].get(dom.name())
        traffic = rec["total"] if rec else None

    def pct(p):
        k = min(len(lat_sorted) - 1, max(0, int(round(p / 100.0 * (len(lat_sorted) - 1)))))
        return lat_sorted[k] * 1e3

    out = {
        "metric": "SCF molecules/s (plain FP64 direct RHF on GPU)",
        "value": a.steps / elapsed,
        "unit": "molecules/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded geometry)",
        "config": {"workload": label, "nbf": nbf, "schwarz_threshold": 1e-12, "energy_tolerance": 1e-8,
                   "eri_convention": "exact",
                   "scf": "alpha=0.5, core guess, DIIS(6/5), normalIterAfterDIIS=0", "parallelism": "single GPU"},
        "energy_Eh": energies[-1],
        "iterations": iters[-1],
        "latency_ms": {"p50": pct(50), "p99": pct(99)},
        "eri_per_s": eris_per_scf * a.steps / elapsed,
        "eri_kernel_share": kernel_share,
        "jk_ms_per_build": sum(r.jk_ms for r in res) / max(1, sum(r.full_builds + r.incremental_builds for r in res)),
        "roofline": {"bound": "mfma", "pipe": "fp64-valu", "kernel": "eri_class_kernel" + dom.name() + " (J/K)",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "flops_per_launch": per_launch_flops, "ms_per_launch": per_launch_ms,
                     "launches": dom.launches},
    }
    # in-core leg: the ERIs evaluated once per SCF into a pair-matrix store in HBM (inside the timed
    # step), every J/K build contracts the store; roofline of the contraction kernel from its HIP
    # events in a profiled SCF (one launch per build on the SCF stream)
    ei = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1).set_molecule(bfile, atoms)
    el_i, lat_i, res_i = timed_scfs(ei, a.steps, a.warmup, incore=1)
    ei.close()
    if res_i[-1].incore:
        ep = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1, profile=1).set_molecule(bfile, atoms)
        ep.scf(incore=1, **SCF_KW)
        ep.reset_kernel_stats()
        ep.scf(incore=1, **SCF_KW)
        ist = [s for s in ep.kernel_stats() if s.la < 0 and not s.fp32][0]
        ep.close()
        bytes_launch = ist.eris / ist.launches * 8.0
        ms_launch = ist.ms / ist.launches
        itraffic = None
        if a.workload == "decane" and os.path.exists(tfile):
            with open(tfile) as fh:
                rec = json.load(fh)["kernels"].get("incore_jk")
            itraffic = rec["total"] if rec else None
        out["incore"] = {
            "value": a.steps / el_i, "unit": "molecules/s", "vs_direct": (a.steps / el_i) / out["value"],
            "abs_dE_vs_direct_Eh": abs(res_i[-1].energy - energies[-1]), "iterations": res_i[-1].iterations,
            "eri_store_ms": res_i[-1].eri_ms, "store_bytes": bytes_launch,
            "jk_ms_per_build": res_i[-1].jk_ms / max(1, res_i[-1].full_builds + res_i[-1].incremental_builds),
            "roofline": {"bound": "hbm", "kernel": "incore_jk_kernel (J/K from the store)",
                         "achieved": bytes_launch / (ms_launch * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": bytes_launch / (ms_launch * 1e-3) / 1e9 / PEAK_HBM_GBS, "traffic": itraffic,
                         "ms_per_launch": ms_launch, "launches": ist.launches}}
    if not a.no_mixed:
        # mixed-precision tier: quartets with Schwarz bound < 1e-6 in FP32, final iterations FP64
        em = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1, precision=1,
                       fp32_threshold=1e-6).set_molecule(bfile, atoms)
        el_m, lat_m, res_m = timed_scfs(em, a.steps, a.warmup)
        out["mixed"] = {"value": a.steps / el_m, "unit": "molecules/s", "vs_fp64": (a.steps / el_m) / out["value"],
                        "abs_dE_vs_fp64_Eh": abs(res_m[-1].energy - energies[-1]), "iterations": res_m[-1].iterations,
                        "fp32_builds": res_m[-1].fp32_builds, "fp32_threshold": 1e-6, "final_fp64_iterations": 3}
        # mixed tier with the in-core store: an FP32 copy of the store until the final FP64 steps
        emi = hf.Engine(schwarz_threshold=1e-12, exact_transfer=1, precision=1,
                        fp32_threshold=1e-6).set_molecule(bfile, atoms)
        el_mi, lat_mi, res_mi = timed_scfs(emi, a.steps, a.warmup, incore=1)
        emi.close()
        out["mixed_incore"] = {"value": a.steps / el_mi, "unit": "molecules/s",
                               "vs_fp64_direct": (a.steps / el_mi) / out["value"],
                               "abs_dE_vs_fp64_Eh": abs(res_mi[-1].energy - energies[-1]),
                               "iterations": res_mi[-1].iterations, "fp32_builds": res_mi[-1].fp32_builds,
                               "incore": res_mi[-1].incore}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(bfile, atoms, iters[-1], a.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
