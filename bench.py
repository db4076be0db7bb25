#!/usr/bin/env python3
"""bench.py — J/K Fock builds per second of the MI355X engine, plain FP64 and
mixed precision, gated by the SCF energy error of the mixed runs.

Default workload (one GPU): the RHF two-electron Fock build J[P], K[P] of
all-trans n-C10H22 (idealised geometry: C-C 1.54 A, C-H 1.09 A, tetrahedral
angles) in the Cartesian cc-pVTZ basis (680 functions, f shells, so every
quartet class up to (ff|ff) runs), Schwarz threshold 1e-12.  A step is one
build from a device-resident density (seeded random symmetric matrix) to
device-resident J and K.

Measured, in this order:
  1. the plain FP64 build: W untimed + K timed steps (default kernels and
     launch shapes; --autotune lets the engine's autotuner pick them first);
  2. the mixed-precision build (quartets whose Schwarz bound lies below the
     FP32 cutoff computed in FP32, everything else FP64) at each cutoff of
     --cutoffs (default 1e-4 .. 1e-7), W + K steps each;
  3. the RHF SCF (core guess, DIIS 8) in FP64 and at each cutoff: wall time,
     iterations and |dE| against the FP64 SCF;
  4. the CPU baseline: the oracle's reference-faithful Fock build on a bounded
     sample of bra AO pairs on the host cores, scaled to one build.
The headline `value` is the fastest build whose SCF energy lies within 1e-6 Eh
of the plain FP64 GPU energy (the plain build itself qualifies).  The roofline
is that configuration's dominant ERI class: flop model v1 over its HIP-event
time in the timed steps.

Other workloads: c20h42-ccpvdz (510 functions, FP64 and mixed) and
water-clusters (256 seeded clusters in one batched, graph-replayed build).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "planck-refactored_amd"))

METRIC = "J/K Fock builds/s and SCF wall time at |ΔE|≤1e-6 Eh vs plain FP64 GPU"
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector / matrix peak
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak


WORKLOADS = {
    # name: (carbons, basis, description)
    "c20h42-ccpvdz": (20, "cc-pvdz", "RHF J/K build, all-trans n-C20H42 (C-C 1.54 A, C-H 1.09 A, tetrahedral), "
                                     "cc-pVDZ Cartesian (510 functions), Schwarz 1e-12"),
    "c10h22-ccpvtz": (10, "cc-pvtz", "RHF J/K build, all-trans n-C10H22 (C-C 1.54 A, C-H 1.09 A, tetrahedral), "
                                     "cc-pVTZ Cartesian (680 functions, f shells), Schwarz 1e-12"),
    "water-clusters": (None, "6-31g*", "RHF J/K build of 256 seeded water clusters of 1-16 molecules (6-31G*, "
                                       "19-304 functions each) in one batched, hipGraph-replayed build, Schwarz 1e-12"),
}


def print_classes(cls):
    """Per-class device time of one profiled build, slowest first (stderr)."""
    for c in sorted(cls, key=lambda c: -c["ms_avg"]):
        print("  class (%d%d|%d%d)%s%s x%d kernels: %8.3f ms  %.3e quartets  %.2f TFLOP/s  %.1f kets/wave" % (
            c["la"], c["lb"], c["lc"], c["ld"], " fp32" if c["precision"] else "", " alt" if c.get("alt") else "",
            c["kernels"], c["ms_avg"],
            c["quartets"], c["model_flops"] / (c["ms_avg"] * 1e-3) * 1e-12,
            c["quartets"] / max(c.get("waves", 0), 1)), file=sys.stderr, flush=True)


def bench_water_clusters(args):
    """One step = the J/K build of all 256 clusters: a batch context (merged
    per-class quartet lists), the engine's default launch shapes (autotuned with
    --autotune), hipGraph replay.  Beside it: the same 256 builds one context at
    a time."""
    import torch
    from pkfock.engine import FockEngine, PREC_MIXED
    from pkfock.host import Molecule
    from pkfock.molecules import water_cluster
    dev = torch.device("cuda:0")
    torch.cuda.init()
    mols = [water_cluster(s, n_min=1, n_max=16) for s in range(256)]
    shells = [Molecule(Z, X, "6-31g*").shells() for Z, X in mols]
    t = time.perf_counter()
    eng = FockEngine(shells, "os", 1e-12)
    if args.fp32_cutoff is not None:
        eng.set("os", 1e-12, precision=PREC_MIXED, fp32_cutoff=args.fp32_cutoff)
    import numpy as np
    rng = np.random.default_rng(0)
    Ps = []
    for n in eng.nbs:
        A = rng.normal(size=(int(n), int(n))) * 0.05
        Ps.append(A + A.T)
    P = torch.tensor(eng.pack(Ps), device=dev)
    J, K = torch.empty_like(P), torch.empty_like(P)
    stream = torch.cuda.current_stream(dev)
    eng.set_stream(stream.cuda_stream)
    eng.jk_device(P.data_ptr(), J.data_ptr(), K.data_ptr())  # work lists
    if args.autotune:
        eng.autotune(reps=2)
    eng.set_graph(True)
    setup_s = time.perf_counter() - t
    for _ in range(max(args.warmup, 1)):
        eng.jk_device(P.data_ptr(), J.data_ptr(), K.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.jk_device(P.data_ptr(), J.data_ptr(), K.data_ptr())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    st = eng.stats()
    eng.set_graph(False)
    eng.profile(True)
    eng.jk_device(P.data_ptr(), J.data_ptr(), K.data_ptr())
    torch.cuda.synchronize()
    cls = eng.class_stats()
    eng.profile(False)
    for c in cls:
        c["ms_avg"] = c["ms_total"] / max(c["builds"], 1)
    dom = max(cls, key=lambda c: c["ms_avg"])
    achieved = dom["model_flops"] / (dom["ms_avg"] * 1e-3) * 1e-12
    print_classes(cls)
    # the same builds one molecule at a time (one context each, eager launches)
    seq_ms, dmax = None, None
    if not args.no_seq:
        singles = [FockEngine(s, "os", 1e-12) for s in shells]
        Ps = eng.unpack(P.cpu().numpy())
        dPs = [torch.tensor(p, device=dev).contiguous() for p in Ps]
        dJs = [torch.empty_like(p) for p in dPs]
        dKs = [torch.empty_like(p) for p in dPs]
        for e, p, j, k in zip(singles, dPs, dJs, dKs):
            e.set_stream(stream.cuda_stream)
            e.jk_device(p.data_ptr(), j.data_ptr(), k.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nrep = max(1, args.steps // 4)
        for _ in range(nrep):
            for e, p, j, k in zip(singles, dPs, dJs, dKs):
                e.jk_device(p.data_ptr(), j.data_ptr(), k.data_ptr())
        torch.cuda.synchronize()
        seq_ms = (time.perf_counter() - t0) / nrep * 1e3
        dmax = max(float((torch.tensor(eng.unpack(J.cpu().numpy())[i]) - dJs[i].cpu()).abs().max())
                   for i in range(256))
        del singles, dPs, dJs, dKs
    print("bench: water clusters nb_total=%d setup %.2fs; batched %.2f ms/build (graph), one-at-a-time %s ms; "
          "max|dJ| batch vs single %s" % (eng.nbasis, setup_s, ms, "%.2f" % seq_ms if seq_ms else "-",
                                          "%.1e" % dmax if dmax is not None else "-"), file=sys.stderr, flush=True)
    mixed = args.fp32_cutoff is not None
    scf = None
    if not args.no_scf:
        mols_h = [Molecule(Z, X, "6-31g*") for Z, X in mols]
        opts = dict(tol_eri=1e-12, tol_energy=1e-8, tol_density=1e-6)
        t = time.perf_counter()
        rb = Molecule.rhf_batch(mols_h, **opts)
        t_batch = time.perf_counter() - t
        scf = {"batched_wall_s": round(t_batch, 3), "max_iterations": max(r["iterations"] for r in rb),
               "converged": all(r["rc"] == 0 for r in rb)}
        if not args.no_seq:
            t = time.perf_counter()
            rs = [m.rhf(**opts) for m in mols_h]
            scf["one_at_a_time_wall_s"] = round(time.perf_counter() - t, 3)
            scf["max_abs_dE_batch_vs_single"] = max(abs(a["energy"] - b["energy"]) for a, b in zip(rb, rs))
        if mixed:
            t = time.perf_counter()
            rm = Molecule.rhf_batch(mols_h, fp32_cutoff=args.fp32_cutoff, **opts)
            scf["mixed_batched_wall_s"] = round(time.perf_counter() - t, 3)
            scf["max_abs_dE_mixed_vs_fp64"] = max(abs(a["energy"] - b["energy"]) for a, b in zip(rm, rb))
        print("bench: water-cluster SCF %s" % scf, file=sys.stderr, flush=True)
    peak = PEAK_FP32_TFLOPS if dom["precision"] else PEAK_FP64_TFLOPS
    out = {
        "metric": METRIC, "value": args.steps / elapsed, "unit": "builds/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64+f32" if mixed else "f64", "data": "synthetic", "scf": scf,
        "config": {"workload": WORKLOADS["water-clusters"][2] + (", mixed FP32 below Schwarz %g" % args.fp32_cutoff
                                                               if mixed else ", FP64"),
                   "molecules": 256, "nbasis_total": eng.nbasis, "shell_quartets": st["quartets"],
                   "kernel_launches_per_build": st["kernel_launches"], "one_at_a_time_ms": seq_ms,
                   "batch_speedup": seq_ms / ms if seq_ms else None},
        "roofline": {"bound": "valu", "kernel": dom_name(dom), "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None, "kernel_ms": dom["ms_avg"]},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def cpu_info():
    """CPU model of this host (lscpu / /proc/cpuinfo) and its logical CPUs."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(Z, X, basis, seconds, threads):
    """Oracle Fock build (per-AO-function Obara-Saika, the reference's
    algorithm) on every stride-th bra AO pair, scaled by the stride."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    from refcpu import Ref
    from pkfock.molecules import ANGSTROM_TO_BOHR
    ref = Ref(basis, Z, X * ANGSTROM_TO_BOHR, engine="os", tol_eri=1e-12)
    n = ref.nbasis
    npairs = n * (n + 1) // 2
    rng = np.random.default_rng(0)
    A = rng.normal(size=(n, n)) * 0.05
    P = A + A.T
    ref.jk_sample(P, stride=npairs, offset=0, nthreads=threads)  # Schwarz table outside the timing
    stride = 65536
    for _ in range(4):
        t = time.perf_counter()
        _, _, nq = ref.jk_sample(P, stride=stride, offset=1, nthreads=threads)
        dt = time.perf_counter() - t
        if dt >= 0.5 * seconds or stride <= 64:
            break
        stride = max(64, int(stride * max(dt, 1e-3) / seconds))
    full = dt * stride
    model, ncpu = cpu_info()
    return {"value": 1.0 / full, "unit": "builds/s", "cores": threads, "kind": "port",
            "cpu_model": model, "host_logical_cpus": ncpu,
            "sample": "oracle J/K (reference-faithful per-function OS, OpenMP) over every %d-th of %d bra AO pairs "
                      "(%.0f AO quartets, %.1f s on %d threads), scaled x%d to one build" % (
                          stride, npairs, nq, dt, threads, stride)}


def timed_builds(eng, P, J, K, steps, warmup, profiled=2):
    """W untimed builds, then K timed ones as the engine runs them (class
    kernels spread over concurrent queues); then `profiled` more builds with
    per-class HIP events on the engine's stream, which run the class lists one
    at a time.  Returns (ms per timed build, seconds, class stats with ms_avg)."""
    import torch
    for _ in range(max(warmup, 1)):
        eng.jk_device(P.data_ptr(), J.data_ptr(), K.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.jk_device(P.data_ptr(), J.data_ptr(), K.data_ptr())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    eng.profile(True)
    for _ in range(profiled):
        eng.jk_device(P.data_ptr(), J.data_ptr(), K.data_ptr())
    torch.cuda.synchronize()
    eng.profile(False)
    cls = eng.class_stats()
    for c in cls:
        c["ms_avg"] = c["ms_total"] / max(c["builds"], 1)
    return elapsed / steps * 1e3, elapsed, cls


def bench_alkane(args):
    import numpy as np
    import torch
    from pkfock.engine import PREC_MIXED
    from pkfock.host import Molecule
    from pkfock.molecules import alkane

    dev = torch.device("cuda:0")
    ncarb, basis, desc = WORKLOADS[args.workload]
    Z, X = alkane(ncarb)
    mol = Molecule(Z, X, basis)
    t = time.perf_counter()
    eng = mol.engine("os", tol_eri=1e-12)
    n = eng.nbasis
    g = torch.Generator().manual_seed(0)
    A = torch.randn(n, n, generator=g, dtype=torch.float64) * 0.05
    P = (A + A.T).contiguous().to(dev)
    J, K = torch.empty_like(P), torch.empty_like(P)
    stream = torch.cuda.current_stream(dev)
    eng.set_stream(stream.cuda_stream)
    eng.jk_device(P.data_ptr(), J.data_ptr(), K.data_ptr())  # work lists
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t
    print("bench: %s nb=%d setup %.2fs" % (args.workload, n, setup_s), file=sys.stderr, flush=True)

    cutoffs = [float(c) for c in args.cutoffs.split(",")] if args.cutoffs else []
    runs = []  # one entry per precision configuration
    for cut in [None] + cutoffs:
        if cut is None:
            eng.set("os", 1e-12)
        else:
            eng.set("os", 1e-12, precision=PREC_MIXED, fp32_cutoff=cut)
        if args.autotune:
            eng.jk_device(P.data_ptr(), J.data_ptr(), K.data_ptr())  # work lists of this configuration
            eng.autotune(reps=2)
        ms, elapsed, cls = timed_builds(eng, P, J, K, args.steps, args.warmup)
        st = eng.stats()
        dom = max(cls, key=lambda c: c["ms_avg"])
        n32 = sum(c["quartets"] for c in cls if c["precision"])
        run = {"cutoff": cut, "ms": ms, "builds_per_s": args.steps / elapsed, "elapsed": elapsed, "cls": cls,
               "dom": dom, "stats": st, "fp32_quartet_frac": n32 / max(st["quartets"], 1)}
        runs.append(run)
        print("bench: %s %d steps %.2f ms/build (%.2f builds/s); model %.3e flop/build = %.2f TFLOP/s; "
              "FP32 quartets %.1f%%" % ("FP64" if cut is None else "mixed %g" % cut, args.steps, ms,
                                        run["builds_per_s"], st["model_flops"], st["model_flops"] / ms * 1e-9,
                                        100 * run["fp32_quartet_frac"]), file=sys.stderr, flush=True)
        if args.verbose or cut is None:
            print_classes(cls)
    eng.set("os", 1e-12)

    # UHF variant (one pass over the ERIs) and one-electron integrals
    t = time.perf_counter()
    mol.one_electron()
    one_ms = (time.perf_counter() - t) * 1e3
    Pb = (0.5 * P).contiguous()
    Jt, Ka, Kb = torch.empty_like(P), torch.empty_like(P), torch.empty_like(P)
    eng.jk_uhf_device(P.data_ptr(), Pb.data_ptr(), Jt.data_ptr(), Ka.data_ptr(), Kb.data_ptr())
    torch.cuda.synchronize()
    t = time.perf_counter()
    nu = max(1, args.steps // 4)
    for _ in range(nu):
        eng.jk_uhf_device(P.data_ptr(), Pb.data_ptr(), Jt.data_ptr(), Ka.data_ptr(), Kb.data_ptr())
    torch.cuda.synchronize()
    uhf_ms = (time.perf_counter() - t) / nu * 1e3
    print("bench: one-electron S, T, V %.2f ms (incl. copy to host); UHF J/K %.2f ms per call (RHF %.2f)" % (
        one_ms, uhf_ms, runs[0]["ms"]), file=sys.stderr, flush=True)

    # SCF per configuration: |dE| against the plain FP64 GPU SCF
    if not args.no_scf:
        scf_kw = dict(engine="os", tol_eri=1e-12, tol_energy=args.scf_tol_energy, tol_density=args.scf_tol_density)
        for run in runs:
            r = mol.rhf(fp32_cutoff=run["cutoff"], **scf_kw)
            run["scf"] = {"wall_s": round(r["wall_s"], 3), "iterations": r["iterations"],
                          "converged": bool(r["converged"]), "energy": r["energy"],
                          "fock_ms_total": round(r["fock_ms"], 1)}
            run["scf"]["abs_dE_vs_fp64"] = abs(r["energy"] - runs[0]["scf"]["energy"])
            print("bench: SCF %s %s" % ("FP64" if run["cutoff"] is None else "mixed %g" % run["cutoff"], run["scf"]),
                  file=sys.stderr, flush=True)

    # headline: the fastest build whose SCF energy is within 1e-6 Eh of FP64
    ok = [r for r in runs if r["cutoff"] is None or
          (r.get("scf") and r["scf"]["converged"] and r["scf"]["abs_dE_vs_fp64"] <= 1e-6)]
    best = max(ok, key=lambda r: r["builds_per_s"])
    dom = best["dom"]
    achieved = dom["model_flops"] / (dom["ms_avg"] * 1e-3) * 1e-12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload.replace("-", "_"))
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dom_name(dom))

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline(Z, X, basis, args.cpu_seconds, threads)
        except Exception as e:  # the baseline is reported, never required
            print("bench: cpu baseline unavailable: %s" % e, file=sys.stderr)

    mixed = best["cutoff"] is not None
    peak = PEAK_FP32_TFLOPS if dom["precision"] else PEAK_FP64_TFLOPS
    sweep = []
    for r in runs:
        e = {"precision": "fp64" if r["cutoff"] is None else "mixed", "fp32_cutoff": r["cutoff"],
             "ms_per_build": round(r["ms"], 3), "builds_per_s": round(r["builds_per_s"], 4),
             "fp32_quartet_frac": round(r["fp32_quartet_frac"], 4),
             "speedup_vs_fp64": round(runs[0]["ms"] / r["ms"], 4)}
        if r.get("scf"):
            e.update({"scf_wall_s": r["scf"]["wall_s"], "scf_iterations": r["scf"]["iterations"],
                      "scf_energy": r["scf"]["energy"], "abs_dE_vs_fp64": r["scf"]["abs_dE_vs_fp64"],
                      "scf_converged": r["scf"]["converged"]})
        sweep.append(e)
    st = best["stats"]
    out = {
        "metric": METRIC, "value": best["builds_per_s"], "unit": "builds/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms"], "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64+f32" if mixed else "f64", "data": "synthetic",
        "config": {"workload": desc + (", mixed: FP32 below Schwarz %g" % best["cutoff"] if mixed else ", FP64"),
                   "nbasis": n, "precision": "mixed" if mixed else "fp64", "fp32_cutoff": best["cutoff"],
                   "shell_quartets": st["quartets"], "primitive_quartets": st["primitive_quartets"],
                   "model_flops_per_build": st["model_flops"],
                   "whole_build_tflops": st["model_flops"] / best["ms"] * 1e-9,
                   "plain_fp64_builds_per_s": runs[0]["builds_per_s"],
                   "speedup_vs_plain_fp64": runs[0]["ms"] / best["ms"],
                   "scf_wall_s": best["scf"]["wall_s"] if best.get("scf") else None,
                   "abs_dE_vs_fp64": best["scf"]["abs_dE_vs_fp64"] if best.get("scf") else None,
                   "scf_tol_energy": args.scf_tol_energy, "scf_tol_density": args.scf_tol_density},
        "sweep": sweep,
        "roofline": {"bound": "valu", "kernel": dom_name(dom), "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                     "kernel_ms": dom["ms_avg"], "kernel_model_flops": dom["model_flops"]},
        "cpu_baseline": cpu,
        "one_electron_ms": round(one_ms, 2),
        "uhf_ms_per_build": round(uhf_ms, 2),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-scf", action="store_true")
    ap.add_argument("--no-seq", action="store_true", help="water clusters: skip the one-at-a-time comparisons")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--verbose", action="store_true", help="per-class times of every configuration")
    ap.add_argument("--autotune", action="store_true",
                    help="let the engine's autotuner pick each class list's kernel and launch shape per configuration "
                         "(its trial builds then also appear in a kernel trace of the run)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c10h22-ccpvtz")
    ap.add_argument("--cutoffs", default="1e-4,1e-5,1e-6,1e-7",
                    help="FP32 Schwarz cutoffs of the mixed-precision sweep ('' for FP64 only)")
    ap.add_argument("--scf-tol-energy", type=float, default=1e-9)
    ap.add_argument("--scf-tol-density", type=float, default=1e-7)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 512 MiB device to device 3 times (scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--fp32-cutoff", type=float, default=None,
                    help="water clusters, mixed precision: quartets with Schwarz bound below this run in FP32")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine measures one GPU; multi-GPU runs are not part of this build "
              "(the J/K build would shard by quartet task list over GPUs)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        n = 64 * 1024 * 1024
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8, "copies": 3}), flush=True)
        return
    if args.workload == "water-clusters":
        return bench_water_clusters(args)
    return bench_alkane(args)


def dom_name(c):
    """Kernel that runs a class list (pk_class_stats.kind)."""
    cls = "%d,%d,%d,%d" % (c["la"], c["lb"], c["lc"], c["ld"])
    kind = c.get("kind", 0)
    if kind == 2:
        return "k_generic (%s)" % cls
    return "k_jk<%s,%s,%s>" % ("float" if kind >= 3 else "double", cls, "OS" if kind in (0, 3) else "Rys")


if __name__ == "__main__":
    main()
