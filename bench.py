#!/usr/bin/env python3
"""Headline benchmark: (k, iw) inversions/s of the lattice Green's-function
loss + gradient step on one MI355X.

Workload (the precision configuration): 14-orbital d+p lattice, correlated
5x5 d block, random Hermitian H(k) from seeded hoppings on the first two cubic
shells, 3-pole Sigma, Nk = 32^3, N_freq = 2048, beta = 50.  Inputs are
synthetic, generated deterministically, and resident in HBM before timing.

One step = what a Newton-Krylov DMFT update needs per call (H(k) is bound
once, with its Schur data eig(H_dd(k)), W = H_cd V_d, before timing):
  moments   fit of the Sigma tail moments (getAsymto_moments, GreenFtn.cpp:189-195)
            and the static eigen-decomposition of H(k) + P m0 P^H (:197-215)
  forward   GreenFtn_w: G_loc, NumMatrix and total N over Nk x 2 N_freq (k, iw)
  adjoint   VJP of L = sum_w ||G_loc - G_tgt||_F^2 + (N - N0)^2 -> dL/dSigma, dL/dmu
value = Nk * 2 N_freq / step time, for the FP32 + FP64-refinement Schur method;
the plain FP64 full-inverse path runs the same step for the ratio.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "(k,ω) inversions/s: FP32+refine & adjoint vs plain FP64, same GPU, err≤1e-9"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


WORKLOADS = {
    "dp14": "dp14: 14-orbital d+p, n_corr=5, Nk=32^3, N_freq=2048, beta=50, 3-pole Sigma",
    "t2g": "t2g: 3-orbital cubic t2g, diagonal Sigma, Nk=16^3, N_freq=1024, beta=40",
    "large32": "large32: 32 orbitals, two 5x5 correlated blocks (n_corr=10), Nk=40^3, N_freq=4096, beta=100",
    "batch64": "batch64: 64 lattices, n in {3,5,7,10,14,20}, Nk in {8^3,12^3,16^3,24^3}, N_freq=1024, beta=40",
}


def log(msg):
    """Progress of the long phases (setup, references, timing) on stderr."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def f_fwd(n):
    return 8.0 * n ** 3


def f_adj(n, nc):
    return 8.0 * n ** 3 + 16.0 * nc ** 3 + 8.0 * n * nc ** 2 + 8.0 * n ** 2


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--method", default="schur_f32r", choices=["schur_f32r", "schur_f64", "plain_f64"])
    ap.add_argument("--workload", default="dp14",
                    choices=["dp14", "t2g", "large32", "batch64", "density_matrix", "real_axis", "ft", "hk",
                             "upfold", "aux"],
                    help="dp14: the 14-orbital precision configuration (default, the headline); t2g: 3-orbital t2g, "
                         "Nk=16^3, N_freq=1024, beta=40; large32: 32 orbitals, two 5x5 correlated blocks, Nk=40^3, "
                         "N_freq=4096, beta=100; batch64: 64 uneven lattices in one replayable step; the others time "
                         "one path either side of the lattice loop (densityMatDFT, real axis, FT round trip, "
                         "H(R)->H(k), upfolding, moments + hybridisation + mu search), one JSON line each")
    ap.add_argument("--nk-side", type=int, default=None)
    ap.add_argument("--n-freq", type=int, default=None)
    ap.add_argument("--refine", type=int, default=1)
    ap.add_argument("--wg", type=int, default=0, choices=[0, 64, 128],
                    help="frequencies per workgroup of the Schur kernels (DMFTGPU_OPT_SCHUR_WG; 0 = the default "
                         "of each launch: adjoint 128, FP32R forward below N_freq as two wavefronts splitting the k "
                         "walk over 64 frequencies, the others 64)")
    ap.add_argument("--target", type=int, default=0,
                    help="work items per lattice launch, sets the k chunking (DMFTGPU_OPT_TARGET_WORKGROUPS; 0 = default)")
    ap.add_argument("--layout", type=int, default=0, choices=[0, 1],
                    help="Schur kernels for n_corr <= 6: 0 one (k,iw) matrix per lane, 1 one column per lane")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device 3 times (a plain copy kernel of known size "
                         "that calibrates HBM byte counters)")
    return ap.parse_args()


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: the lattice path runs on a single MI355X here "
              f"(requested --gpus {args.gpus}, WORLD_SIZE {world}); sharding the k-mesh over GPUs "
              f"(one rank per GPU, all-reduce of G_loc) is not part of this build.", file=sys.stderr)
        sys.exit(2)

    import torch
    from dmftpack_amd import capi, gpu as G, lattice as L

    methods = {"plain_f64": G.PLAIN_F64, "schur_f64": G.SCHUR_F64, "schur_f32r": G.SCHUR_F32R}
    torch.cuda.init()
    dev = torch.device("cuda:0")

    if args.workload == "batch64":
        return bench_batch(args, torch, capi, G, L)
    if args.workload in PATHS:
        return bench_path(args, torch, capi, G, L)

    # ---- synthetic workload -------------------------------------------------
    log(f"building workload {args.workload}")
    w = L.workload(args.workload, seed=1, nk_side=args.nk_side, n_freq=args.n_freq)
    Hk, sig, beta = w["Hk"], w["sigma"], w["beta"]
    mom = w.get("exact_moments")
    if mom is None:  # tail moments fitted on the GPU (getAsymto_moments)
        hm = G.Handle()
        mom = hm.get_asymto_moments(sig, beta)
        hm.close()
    nk, n = Hk.shape[0], Hk.shape[1]
    nc, nf = sig.shape[1], sig.shape[0]
    ncl = w.get("n_cluster", 1)
    N0 = w["n_target"]
    mu = 0.25 if args.workload == "dp14" else w.get("mu", 0.0)
    units = nk * 2 * nf
    # host -> device copies of the inputs (PCIe), outside every timed region
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hk_dev = G.to_dev_cm(Hk)
    torch.cuda.synchronize()
    upload_hk_ms = (time.perf_counter() - t0) * 1e3
    hk_dev._dmft_colmajor = True
    lat = G.LatticeGF(hk_dev, nc, nf, beta, mom, n_cluster=ncl)
    del w
    lat.set_option(capi.OPT_REFINE_STEPS, args.refine)
    lat.set_option(capi.OPT_SCHUR_LAYOUT, args.layout)
    lat.set_option(capi.OPT_SCHUR_WG, args.wg)
    lat.set_option(capi.OPT_TARGET_WORKGROUPS, args.target)
    s_dev = lat.sigma_dev(sig)
    # target G_loc from a seeded perturbation of Sigma (plain FP64, untimed)
    s_tgt = lat.sigma_dev(L.perturbed(sig, seed=7))
    log("target G_loc (plain FP64, untimed)")
    gtgt, _ = lat.greenftn_w_dev(s_tgt, mu, G.PLAIN_F64)
    gtgt = gtgt.clone()
    mom_dev = G.to_dev_cm(mom)
    gbuf = torch.empty((nf, nc, nc), dtype=torch.complex128, device=dev)
    nbuf = torch.empty((nc, nc), dtype=torch.complex128, device=dev)
    sbar = torch.empty_like(s_dev)

    def step(method):
        # a new Sigma brings new tail moments: the fit of getAsymto_moments (GreenFtn.cpp:189-195), then the
        # static eigen-decomposition of H(k) + P m0 P^H
        lat.check(lat.lib.dmftgpu_moments(lat.h, nc, nf, beta, G._ptr(s_dev), 0, G._ptr(mom_dev), None))
        lat.check(lat.lib.dmftgpu_update_moments(lat.h, G._ptr(mom_dev)))
        # forward (G_loc, NumMatrix, N), loss cotangents and adjoint (dmftgpu_gloc_loss_grad)
        _, _, N, _, _, mb = lat.loss_grad_dev(s_dev, mu, gtgt, N0, method, gloc=gbuf, num=nbuf, out=sbar)
        return N, mb

    def timed(method, steps, warmup, kernel_times=False):
        lat.enable_timing(kernel_times)
        for _ in range(warmup):
            step(method)
        torch.cuda.synchronize()
        kf, kv = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            step(method)
            if kernel_times:
                kf.append(lat.kernel_ms(capi.MODE_FWD))
                kv.append(lat.kernel_ms(capi.MODE_VJP))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        lat.enable_timing(False)
        return (t1 - t0) / steps, kf, kv

    head = methods[args.method]
    log(f"timing {args.method}: {args.warmup} warmup + {args.steps} steps")
    ms_head, kf, kv = timed(head, args.steps, args.warmup, kernel_times=True)
    value = units / ms_head

    # ---- comparisons on the same inputs ---------------------------------------
    extra = {}
    if not args.no_compare:
        res = {}
        for name in ("plain_f64", "schur_f64", "schur_f32r"):
            log(f"reference outputs {name}")
            m = methods[name]
            gl, num, N, _, sb, mb = lat.loss_grad_dev(s_dev, mu, gtgt, N0, m)
            res[name] = dict(gloc=gl, num=num, N=N, sb=sb, mb=mb)
        ref = res["plain_f64"]
        errs = {}
        for name in ("schur_f64", "schur_f32r"):
            r = res[name]
            errs[name] = dict(
                max_abs_dGloc=float(torch.max(torch.abs(r["gloc"] - ref["gloc"])).item()),
                max_abs_dNumMatrix=float(torch.max(torch.abs(r["num"] - ref["num"])).item()),
                abs_dN=abs(r["N"] - ref["N"]),
                rel_dSigmaBar=float((torch.max(torch.abs(r["sb"] - ref["sb"])) / torch.max(torch.abs(ref["sb"]))).item()),
                rel_dMuBar=abs(r["mb"] - ref["mb"]) / max(abs(ref["mb"]), 1e-300))
        extra["errors_vs_plain_fp64"] = errs
        others = {}
        for name in ("plain_f64", "schur_f64", "schur_f32r"):
            for layout in ((0,) if name == "plain_f64" or nc > 6 else (0, 1)):
                if name == args.method and layout == args.layout:
                    continue
                lat.set_option(capi.OPT_SCHUR_LAYOUT, layout)
                steps = args.plain_steps if name == "plain_f64" else max(3, args.steps // 2)
                log(f"timing {name} layout {layout}: {steps} steps")
                ms, pkf, pkv = timed(methods[name], steps, 1, kernel_times=(name == "plain_f64"))
                key = name if name == "plain_f64" else f"{name}_layout{layout}"
                others[key] = dict(value=units / ms, ms_per_step=ms * 1e3, steps=steps)
                if name == "plain_f64":
                    others[key].update(fwd_call_ms=float(np.mean(pkf)), vjp_call_ms=float(np.mean(pkv)))
        lat.set_option(capi.OPT_SCHUR_LAYOUT, args.layout)
        extra["other_methods"] = others
        if "plain_f64" in others:
            extra["speedup_vs_plain_fp64"] = value / others["plain_f64"]["value"]
            extra["plain_fp64_executed"] = plain_executed(others["plain_f64"], args.workload)

    # ---- roofline of the dominant kernel --------------------------------------
    # On the one-matrix-per-lane Schur path dmftgpu_gloc_loss_grad is fused: the
    # forward call is two launches (forward below N_freq; Hermitian-part forward +
    # tr G^2 above), the adjoint call one launch over the first N_freq frequencies,
    # which is the longest single launch of the step.  The library's events time
    # each call's main launches on its own stream.
    t_f, t_v = float(np.mean(kf)) / 1e3, float(np.mean(kv)) / 1e3
    fused = head != G.PLAIN_F64 and nc <= 6 and args.layout == 0
    if fused or t_v >= t_f:
        kname = "k_lattice_schur<VJP> (first N_freq)" if fused else "k_lattice_schur<VJP>"
        t_k, flop, pkey = t_v, f_adj(n, nc), ("VJP" if fused else "VJP_CALL")
        units_k = nk * nf if fused else units
    else:
        kname = "k_lattice_schur<FWD>"
        t_k, flop, pkey, units_k = t_f, f_fwd(n), "FWD_CALL", units
    if head == G.PLAIN_F64:
        kname = kname.replace("schur", "plain")
    # HBM bytes per launch: FETCH_SIZE / WRITE_SIZE counters of the same kernel,
    # taken on the dp14 workload only (profiles/pmc_traffic.json)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath) and args.workload == "dp14":
        try:
            traffic = json.load(open(tpath)).get(args.method, {}).get(pkey)
        except Exception:
            traffic = None
    # FLOPs the kernel executes per launch: rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 +
    # SQ_INSTS_VALU_FLOPS_FP32 (x 64 lanes) of the same kernel on the same workload,
    # summarised in profiles/pmc_flops.json; FP32 is not packed in this build, so
    # both issue at the FP64 VALU rate and share its 78.6 TFLOP/s peak
    executed = None
    fpath = os.path.join(ROOT, "profiles", "pmc_flops.json")
    if os.path.exists(fpath) and args.workload == "dp14":
        try:
            ex = json.load(open(fpath)).get(args.method, {}).get(pkey)
            if ex:
                executed = dict(flop_per_launch=ex["fp64_flops"] + ex["fp32_flops"], fp64=ex["fp64_flops"],
                                fp32=ex["fp32_flops"], counter_kernel=ex["kernel"],
                                counter_launch_ms=ex["launch_s"] * 1e3)
        except Exception:
            executed = None
    plain_eq = dict(flop_per_unit=flop, units_per_launch=units_k,
                    tflops=flop * units_k / t_k / 1e12, frac_of_fp64_peak=flop * units_k / t_k / 1e12 / PEAK_FP64_TFLOPS,
                    note="the plain algorithm's work (8n^3 per (k,iw); adjoint 8n^3+16nc^3+8n nc^2+8n^2) over this "
                         "kernel's time: a speed figure, not a fraction of peak, since the Schur kernels do less work")
    achieved = executed["flop_per_launch"] / t_k / 1e12 if executed else None
    roofline = dict(bound="fp64_valu", achieved=achieved, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                    frac=None if achieved is None else achieved / PEAK_FP64_TFLOPS, traffic=traffic, kernel=kname,
                    kernel_ms=t_k * 1e3, fwd_call_ms=t_f * 1e3, vjp_call_ms=t_v * 1e3,
                    flop_per_launch=None if executed is None else executed["flop_per_launch"],
                    flop_source=None if executed is None else
                    "profiles/pmc_flops.json: rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 + _FP32 x 64 lanes of "
                    + executed["counter_kernel"] + f" ({executed['counter_launch_ms']:.2f} ms under the profiler)",
                    plain_equivalent=plain_eq,
                    note="achieved = FLOPs the dominant kernel executes per launch (hardware counters, same kernel "
                         "and workload) / its live duration (HIP events on the library's stream); the kernels are "
                         "FP64/FP32 VALU code (no MFMA), bound by the VALU issue rate, 78.6 TFLOP/s")

    # ---- CPU baseline: the oracle's GreenFtn_w on a k subset ---------------------
    cpu = None
    if not args.no_cpu:
        try:
            from oracle import oracle as O
            log("CPU baseline (oracle)")
            cores = len(os.sched_getaffinity(0))
            probe_k = max(1, min(nk, 4 * cores))
            t0 = time.perf_counter()
            O.greenftn_w(Hk[:probe_k], sig, mom, mu, beta, nthreads=cores)
            tp = time.perf_counter() - t0
            kk = int(min(nk, max(probe_k, probe_k * args.cpu_seconds / max(tp, 1e-3))))
            kk = max(cores, (kk // cores) * cores)
            t0 = time.perf_counter()
            O.greenftn_w(Hk[:kk], sig, mom, mu, beta, nthreads=cores)
            tc = time.perf_counter() - t0
            # one core, on a smaller subset (about a quarter of the time budget)
            per_core_k_per_s = kk / cores / max(tc, 1e-3)
            k1 = max(1, min(nk, int(per_core_k_per_s * 0.25 * args.cpu_seconds)))
            t0 = time.perf_counter()
            O.greenftn_w(Hk[:k1], sig, mom, mu, beta, nthreads=1)
            t1c = time.perf_counter() - t0
            cpu = dict(value=kk * 2 * nf / tc, unit="(k,w) inversions/s", cores=cores, kind="port",
                       sample=f"oracle GreenFtn_w (partial-pivot LU inverse + G0 per (k,iw), OpenMP over k) on "
                              f"{kk} of {nk} k points x {2 * nf} frequencies, {tc:.1f} s; forward only",
                       value_1core=k1 * 2 * nf / t1c,
                       sample_1core=f"{k1} k points x {2 * nf} frequencies on 1 thread, {t1c:.1f} s")
        except Exception as e:  # the baseline is reported, never required
            cpu = dict(value=None, unit="(k,w) inversions/s", cores=None, kind="port", sample=f"failed: {e}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "(k,w) inversions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_head * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.method == "schur_f32r" else "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload],
                   "n": n, "n_corr": nc, "n_k": nk, "n_freq": nf, "beta": beta, "method": args.method,
                   "refine_steps": args.refine, "schur_layout": args.layout, "schur_wg": args.wg or "default (adjoint 128, forward 64, FP32R forward below N_freq in k pairs)",
                   "target_workgroups": args.target or "default",
                   "units_per_step": units,
                   "step": "tail-moment fit of Sigma + static eig for the tail + GreenFtn_w forward (G_loc, NumMatrix, "
                           "N) + loss cotangents + VJP (dL/dSigma, dL/dmu): dmftgpu_moments, dmftgpu_update_moments, "
                           "one dmftgpu_gloc_loss_grad call"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "host_to_device": {"hk_bytes": int(Hk.nbytes), "hk_upload_ms": upload_hk_ms,
                           "note": "H(k) copy over PCIe before timing; not part of value"},
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    lat.close()
    calibration_copy(args, torch)


def calibration_copy(args, torch):
    """A plain device copy of --calib-copy MiB, three times, after the timed work:
    the known-size kernel that scales HBM byte counters (FETCH_SIZE, WRITE_SIZE)."""
    if args.calib_copy <= 0:
        return
    src = torch.ones(args.calib_copy * (1 << 20) // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"calibration copy: {args.calib_copy} MiB read + {args.calib_copy} MiB written per launch", flush=True)


def plain_executed(rec, workload):
    """Executed fraction of FP64 peak of the plain comparator's two lattice calls:
    FLOPs per launch from rocprofv3 SQ_INSTS_VALU_FLOPS_FP64/_FP32 x 64 lanes of
    k_lattice_plain on the same workload (profiles/pmc_flops.json, "plain_f64")
    over this run's live call times (HIP events on the library's stream)."""
    fpath = os.path.join(ROOT, "profiles", "pmc_flops.json")
    if workload != "dp14" or not os.path.exists(fpath):
        return None
    try:
        pf = json.load(open(fpath)).get("plain_f64", {})
    except Exception:
        return None
    out = {}
    for call, key in (("fwd_call", "FWD_CALL"), ("vjp_call", "VJP_CALL")):
        ex, ms = pf.get(key), rec.get(call + "_ms")
        if not ex or not ms:
            continue
        fl = ex["fp64_flops"] + ex["fp32_flops"]
        out[call] = dict(ms=ms, flop_per_launch=fl, tflops=fl / (ms / 1e3) / 1e12,
                         frac=fl / (ms / 1e3) / 1e12 / PEAK_FP64_TFLOPS, counter_kernel=ex["kernel"])
    if not out:
        return None
    fl = sum(v["flop_per_launch"] for v in out.values())
    ms = sum(v["ms"] for v in out.values())
    out["both_calls_frac"] = fl / (ms / 1e3) / 1e12 / PEAK_FP64_TFLOPS
    out["source"] = "profiles/pmc_flops.json (plain_f64) over live call times"
    return out


def bench_batch(args, torch, capi, G, L):
    """64 independent lattices of uneven size in one replayable step (persistent
    launches per n_corr, hipGraph, autotuned tiles): prepare + forward + loss
    cotangents + adjoint for every system."""
    methods = {"plain_f64": G.PLAIN_F64, "schur_f64": G.SCHUR_F64, "schur_f32r": G.SCHUR_F32R}
    method = methods[args.method] if args.method != "plain_f64" else G.SCHUR_F64
    nf = args.n_freq or 1024
    systems = L.batch_systems(64, seed=1, n_freq=nf)
    h = G.Handle()
    h.set_option(capi.OPT_REFINE_STEPS, args.refine)
    units = 0
    for i, sd in enumerate(systems):
        sd["moments"] = sd["exact_moments"]
        sd["mu"] = 0.1 * (i % 3)
        sd["n_target"] = 0.5 * sd["n_corr"]
        # target: G_loc of a perturbed Sigma (single-lattice FP64, untimed)
        lat = G.LatticeGF(sd["Hk"], sd["n_corr"], nf, sd["beta"], sd["moments"])
        sd["gloc_target"] = lat.greenftn_w(L.perturbed(sd["sigma"], seed=100 + i), sd["mu"], method=G.SCHUR_F64)["gloc"]
        lat.close()
        units += sd["Hk"].shape[0] * 2 * nf
    b = G.Batch(h, systems, method=method, tune=True, graph=True, prepare=True)
    for _ in range(args.warmup):
        b.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    got = b.results()
    # per-system errors against the single-lattice FP64 path (untimed)
    worst = dict(max_abs_dGloc=0.0, rel_dNumMatrix=0.0, abs_dN=0.0, rel_dSigmaBar=0.0, rel_dMuBar=0.0)
    if not args.no_compare:
        for sd, r in zip(systems, got):
            lat = G.LatticeGF(sd["Hk"], sd["n_corr"], nf, sd["beta"], sd["moments"])
            ref = lat.greenftn_w(sd["sigma"], sd["mu"], method=G.SCHUR_F64)
            gbar = 2.0 * (ref["gloc"] - sd["gloc_target"])
            sb, mb = lat.vjp(sd["sigma"], sd["mu"], gbar, 2.0 * (ref["n_total"] - sd["n_target"]), method=G.SCHUR_F64)
            lat.close()
            worst["max_abs_dGloc"] = max(worst["max_abs_dGloc"], float(np.max(np.abs(r["gloc"] - ref["gloc"]))))
            worst["rel_dNumMatrix"] = max(worst["rel_dNumMatrix"], float(np.max(np.abs(r["num_matrix"] - ref["num_matrix"]))
                                                                       / np.max(np.abs(ref["num_matrix"]))))
            worst["abs_dN"] = max(worst["abs_dN"], abs(r["n_total"] - ref["n_total"]))
            worst["rel_dSigmaBar"] = max(worst["rel_dSigmaBar"], float(np.max(np.abs(r["sigma_bar"] - sb)) / np.max(np.abs(sb))))
            worst["rel_dMuBar"] = max(worst["rel_dMuBar"], abs(r["mu_bar"] - mb) / max(abs(mb), 1e-300))
    out = {
        "metric": METRIC, "value": units / sec, "unit": "(k,w) inversions/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if method == G.SCHUR_F32R else "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS["batch64"], "systems": len(systems), "units_per_step": units,
                   "method": args.method, "refine_steps": args.refine,
                   "step": "per system: prepare + GreenFtn_w forward + loss cotangents + VJP, one graph replay"},
        "worst_error_vs_single_fp64": worst if not args.no_compare else None,
        "tuning": b.tuning(),
    }
    print(json.dumps(out), flush=True)
    b.close()
    h.close()


# ---- the paths either side of the lattice hot path --------------------------------
PATHS = {
    "density_matrix": "densityMatDFT (GreenFtn.cpp:291-319): dp14 lattice, Nk=32^3, N_freq=2048, full n x n "
                      "inverse per (k,iw)",
    "real_axis": "retarded_GreenFtn2 k sums (qs_spectra.cpp:118-131): dp14 lattice, Nk=32^3, 1000 energies, "
                 "eta=0.05",
    "ft": "FourierTransform_ + FT_t_to_w round trip (FourierTransform.cpp:126-178,521-561): d=10, N_freq=4096, "
          "N_tau=8192, beta=100",
    "hk": "H(R) -> H(k) (ConstructModelHamiltonian.cpp:58-77): 32 orbitals, Nk=40^3, two cubic shells",
    "upfold": "upfolding_density (upfolding.cpp:38-139): NBAND=14, NumOrbit=28, Nk=32^3, overlap, "
              "R in {-1,0,1}^3 x all orbital pairs",
    "aux": "getAsymto_moments (d=5, N_freq=2048) + Construct_hyb_delta (5x5, N_freq=2048) + "
           "GreenFtn_w_adjust_mu (dp14 lattice, Newton on dN/dmu)",
}


def _events(torch, fn, steps, warmup):
    """Per-call GPU time (HIP events on the current stream, the stream the
    library's handle launches on) and wall time, after warmup calls."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return e0.elapsed_time(e1) / 1e3 / steps, wall


def _cpu_sample(fn, total_units, units_of, budget, probe):
    """Time the oracle on a bounded sample: a probe size first, then a sample
    sized to about `budget` seconds; returns (units/s, sample size, seconds)."""
    t0 = time.perf_counter()
    fn(probe)
    tp = max(time.perf_counter() - t0, 1e-4)
    size = int(max(probe, min(total_units, probe * budget / tp)))
    t0 = time.perf_counter()
    fn(size)
    tc = max(time.perf_counter() - t0, 1e-4)
    return units_of(size) / tc, size, tc


def _flop_roofline(achieved_tflops, kernel, note, flop_per_unit, units, plain_equivalent=False, executed=None):
    """Roofline of a path's main kernel.  achieved_tflops is flop_per_unit x units
    over the kernel time.  When flop_per_unit is the plain algorithm's 8n^3
    (plain_equivalent) and the kernel executes less, that rate is a speed figure,
    not a fraction of peak: it goes to its own field, and achieved / frac come
    from counter-executed FLOPs (executed = (TFLOP/s, source)) or stay null."""
    if not plain_equivalent:
        return dict(bound="fp64_valu", achieved=achieved_tflops, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                    frac=achieved_tflops / PEAK_FP64_TFLOPS, traffic=None, kernel=kernel,
                    flop_per_unit=flop_per_unit, units_per_launch=units, note=note)
    ach = executed[0] if executed else None
    return dict(bound="fp64_valu", achieved=ach, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                frac=None if ach is None else ach / PEAK_FP64_TFLOPS, traffic=None, kernel=kernel,
                flop_source=executed[1] if executed else None,
                plain_equivalent=dict(flop_per_unit=flop_per_unit, units_per_launch=units, tflops=achieved_tflops,
                                      note=note), units_per_launch=units)


def _density_lr_executed(t_gpu):
    """Executed FP64 FLOPs of k_density_lr per launch (MFMA + VALU) from the
    counter summary profiles/mfma_counters_density_matrix_dp14.txt, over t_gpu."""
    path = os.path.join(ROOT, "profiles", "mfma_counters_density_matrix_dp14.txt")
    try:
        for line in open(path):
            f = line.split()
            if "k_density_lr" in line and len(f) > 3 and not line.startswith("#"):
                gf = float(f[1]) + float(f[2])
                return gf / 1e3 / t_gpu, ("profiles/mfma_counters_density_matrix_dp14.txt: SQ_INSTS_VALU_MFMA_MOPS_F64 "
                                          "x 512 + SQ_INSTS_VALU_FLOPS_FP64 x 64 = %.1f GFLOP per launch" % gf)
    except OSError:
        pass
    return None


def bench_path(args, torch, capi, G, L):
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    budget = args.cpu_seconds
    methods = {"plain_f64": G.PLAIN_F64, "schur_f64": G.SCHUR_F64, "schur_f32r": G.SCHUR_F32R}
    wname = args.workload
    cpu, extra, dtype = None, {}, "f64"

    if wname in ("density_matrix", "real_axis", "aux"):
        log("building the dp14 lattice")
        w = L.workload("dp14", seed=1)
        Hk, sig, beta, mom = w["Hk"], w["sigma"], w["beta"], w["exact_moments"]
        nk, n = Hk.shape[0], Hk.shape[1]
        nc, nf = sig.shape[1], sig.shape[0]
        mu = 0.25
        lat = G.LatticeGF(Hk, nc, nf, beta, mom)
        s_dev = lat.sigma_dev(sig)

    if wname == "density_matrix":
        units = nk * 2 * nf
        out = torch.empty((nk, n, n), dtype=torch.complex128, device="cuda")
        fn = lambda: lat.check(lat.lib.dmftgpu_density_matrix(lat.h, G._ptr(s_dev), mu, G._ptr(out)))  # noqa: E731
        per = {}
        for kern, kname in ((1, "full_inverse"), (0, "block_mfma")):
            lat.set_option(capi.OPT_DENSITY_KERNEL, kern)
            log(f"timing density_matrix ({kname})")
            t_gpu, t_wall = _events(torch, fn, args.steps, args.warmup)
            per[kname] = dict(value=units / t_wall, ms_per_call=t_wall * 1e3, gpu_ms=t_gpu * 1e3)
            torch.cuda.synchronize()
            per[kname]["out"] = out.clone()
        diff = float(torch.max(torch.abs(per["block_mfma"]["out"] - per["full_inverse"]["out"])).item())
        scale = float(torch.max(torch.abs(per["full_inverse"]["out"])).item())
        for v in per.values():
            del v["out"]
        t_gpu = per["block_mfma"]["gpu_ms"] / 1e3
        metric, unit, value = "densityMatDFT (k,w) inversions/s", "(k,w) inversions/s", per["block_mfma"]["value"]
        extra["kernels"] = per
        extra["speedup_block_vs_full_inverse"] = per["block_mfma"]["value"] / per["full_inverse"]["value"]
        extra["max_abs_diff_block_vs_full_inverse"] = diff
        extra["rel_diff_block_vs_full_inverse"] = diff / max(scale, 1e-300)
        ach = f_fwd(n) * units / t_gpu / 1e12
        roof = _flop_roofline(ach, "k_density_lr", "plain-equivalent 8n^3 FLOPs per (k,iw) x Nk*2N_freq / call time "
                              "(events); the block form executes far fewer (Schur complement + rank-n_corr product "
                              "on FP64 MFMA): a speed figure, not a fraction of peak", f_fwd(n), units,
                              plain_equivalent=True, executed=_density_lr_executed(t_gpu) if nk == 32768 else None)
        if not args.no_cpu:
            def run(kk):
                O.greenftn_w(Hk[:kk], sig, mom, mu, beta, density=True, nthreads=cores)
            v, kk, tc = _cpu_sample(run, nk, lambda kk: kk * 2 * nf, budget, max(1, cores))
            cpu = dict(value=v, unit=unit, cores=cores, kind="port",
                       sample=f"oracle GreenFtn_w with densityMatDFT on {kk} of {nk} k x {2 * nf} frequencies, {tc:.1f} s")
        cfg = dict(n=n, n_corr=nc, n_k=nk, n_freq=nf, beta=beta, units_per_call=units)
    elif wname == "real_axis":
        ne, e0, de, eta = 1000, -5.0, 0.01, 0.05
        E = e0 + de * np.arange(ne)
        rng = np.random.default_rng(1)
        sig_r = np.zeros((ne, nc, nc), complex)
        for _ in range(2):
            B = rng.normal(size=(nc, nc)) + 1j * rng.normal(size=(nc, nc))
            sig_r += (B @ B.conj().T / nc)[None] / (E - rng.uniform(-1, 1) + 0.3j)[:, None, None]
        sr = G.to_dev_cm(sig_r)
        g = torch.empty((ne, nc, nc), dtype=torch.complex128, device="cuda")
        tr = torch.empty((ne,), dtype=torch.complex128, device="cuda")
        units = nk * ne
        per = {}
        for name in ("schur_f64", "schur_f32r", "plain_f64"):
            m = methods[name]
            fn = lambda: lat.check(lat.lib.dmftgpu_greenftn_real(lat.h, m, ne, e0, de, eta, G._ptr(sr), mu,  # noqa: E731
                                                                 G._ptr(g), G._ptr(tr)))
            log(f"timing real axis {name}")
            t_gpu, t_wall = _events(torch, fn, args.steps if name != "plain_f64" else 2, 1)
            per[name] = dict(value=units / t_wall, ms_per_call=t_wall * 1e3, gpu_ms=t_gpu * 1e3)
        head = args.method
        metric, unit, value = "real-axis (k,E) inversions/s (k-summed G^R)", "(k,E) inversions/s", per[head]["value"]
        dtype = "f32+f64" if head == "schur_f32r" else "f64"
        ach = f_fwd(n) * units / (per[head]["gpu_ms"] / 1e3) / 1e12
        roof = _flop_roofline(ach, "k_lattice_schur_cols<FWD> (real axis)" if head != "plain_f64" else
                              "k_lattice_plain<FWD> (real axis)", "algorithmic 8n^3 FLOPs per (k,E) x Nk*n_E / call "
                              "time (events)", f_fwd(n), units, plain_equivalent=head != "plain_f64")
        extra["methods"] = per
        extra["speedup_vs_plain_fp64"] = per[head]["value"] / per["plain_f64"]["value"]
        if not args.no_cpu:
            def run(kk):
                O.greenftn_real(Hk[:kk], sig_r, mu, e0, de, eta, nthreads=cores)
            v, kk, tc = _cpu_sample(run, nk, lambda kk: kk * ne, budget, max(1, cores // 4))
            cpu = dict(value=v, unit=unit, cores=cores, kind="port",
                       sample=f"oracle retarded_GreenFtn2 k sums on {kk} of {nk} k x {ne} energies, {tc:.1f} s")
        cfg = dict(n=n, n_corr=nc, n_k=nk, n_e=ne, eta=eta, method=head, units_per_call=units)
    elif wname == "ft":
        log("building the large32 self-energy")
        sig, _ = L.pole_sigma(10, 4096, 100.0, 1, block=5)
        beta, nf, ntau, d = 100.0, 4096, 8192, 10
        h = G.Handle()
        mom = h.get_asymto_moments(sig, beta)
        f = G.to_dev_cm(sig)
        m = G.to_dev_cm(mom)
        gt = torch.empty((ntau + 1, d, d), dtype=torch.complex128, device="cuda")
        fw = torch.empty((nf, d, d), dtype=torch.complex128, device="cuda")
        pt, vi = capi.C.c_int(0), capi.C.c_int(0)

        def fn():
            h.check(h.lib.dmftgpu_ft_w2tau(h.h, d, nf, ntau, beta, G._ptr(f), G._ptr(m), G._ptr(gt), capi.C.byref(pt)))
            h.check(h.lib.dmftgpu_ft_tau2w(h.h, d, nf, ntau, beta, G._ptr(gt), G._ptr(fw), capi.C.byref(vi)))
        log("timing iw -> tau -> iw")
        t_gpu, t_wall = _events(torch, fn, args.steps, args.warmup)
        pairs = (ntau + 1) * nf
        units = 2 * pairs
        metric, unit, value = "(tau,iw) pairs/s, iw->tau->iw round trip", "(tau,w) pairs/s", units / t_wall
        ach = 8.0 * d * d * units / t_gpu / 1e12
        roof = _flop_roofline(ach, "k_dft_mfma<W2Tau> + k_dft_mfma<Tau2W>",
                              "algorithmic 8 d^2 FLOPs (one complex MAC per element) per (tau,iw) pair and direction "
                              "/ call time (events, both calls: moment-free transform kernels, tail, causality "
                              "eigensolves); the two transforms are FP64 MFMA complex GEMMs (4 real "
                              "v_mfma_f64_16x16x4 per complex step, d^2 padded to 16-element tiles)",
                              8.0 * d * d, units)
        roof["bound"] = "mfma"
        if not args.no_cpu:
            sig_np, mom_np = sig, mom

            def run(nt):
                gts, _ = O.ft_w2tau(sig_np, mom_np, beta, nt, nthreads=cores)
                O.ft_tau2w(gts, beta, nf)
            v, nt, tc = _cpu_sample(run, ntau, lambda nt: 2 * (nt + 1) * nf, budget, 16)
            cpu = dict(value=v, unit=unit, cores=cores, kind="port",
                       sample=f"oracle FourierTransform_ (OpenMP over tau) + FT_t_to_w at N_tau={nt} of {ntau}, "
                              f"N_freq={nf}, {tc:.1f} s")
        cfg = dict(d=d, n_freq=nf, n_tau=ntau, beta=beta, units_per_call=units)
        lat = h
    elif wname == "hk":
        n, side = 32, 40
        Rs, HR = L.random_tb_hr(n, 1)
        hops = L.hopping_list(Rs, HR)
        kp_np = L.kmesh_phases(side)
        nk = kp_np.shape[0]
        h = G.Handle()
        idx = np.ascontiguousarray(np.array([x[:5] for x in hops], dtype=np.int32).reshape(-1, 5))
        tt = np.ascontiguousarray(np.array([complex(x[5]) for x in hops], dtype=np.complex128))
        kp = torch.from_numpy(np.ascontiguousarray(kp_np)).to("cuda")
        out = torch.empty((nk, n, n), dtype=torch.complex128, device="cuda")

        def fn():
            h.check(h.lib.dmftgpu_hk_from_hr(h.h, n, len(hops), idx.ctypes.data_as(capi.C.POINTER(capi.C.c_int)),
                                             tt.ctypes.data_as(capi.C.POINTER(capi.C.c_double)), nk, G._ptr(kp),
                                             G._ptr(out)))
        log("timing H(R) -> H(k)")
        t_gpu, t_wall = _events(torch, fn, args.steps, args.warmup)
        units = nk
        metric, unit, value = "H(k) matrices/s", "k points/s", units / t_wall
        nbytes = nk * n * n * 16.0
        ach = nbytes / t_gpu / 1e9
        roof = dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS,
                    traffic=None, kernel="k_hk_from_hr", bytes_per_unit=n * n * 16.0, units_per_launch=units,
                    hops=len(hops),
                    note="algorithmic bytes = the H(k) written once (n^2 x 16 B per k) / call time (events, incl. "
                         "the host grouping of the hops); each element also takes one FP64 sincos per hop, "
                         "which sets the time")
        if not args.no_cpu:
            def run(kk):
                O.hk_from_hr(n, hops, kp_np[:kk])
            v, kk, tc = _cpu_sample(run, nk, lambda kk: kk, budget, 16)
            cpu = dict(value=v, unit=unit, cores=1, kind="port",
                       sample=f"oracle H(k) sum on {kk} of {nk} k points, {len(hops)} hops, {tc:.1f} s")
        cfg = dict(n=n, n_k=nk, hops=len(hops), units_per_call=units)
        lat = h
    elif wname == "upfold":
        n, no, side = 14, 28, 32
        log("building the upfolding inputs")
        c = L.upfold_case(n, no, side, 1)
        nk = c["kphase"].shape[0]
        h = G.Handle()
        X, Cb, T, Ko, S = (G.to_dev_cm(c[k]) for k in ("dm_model", "corr_base", "transform", "ks_ortho", "overlap"))
        eps = torch.from_numpy(np.ascontiguousarray(c["ks_energy"])).to("cuda")
        kp = torch.from_numpy(np.ascontiguousarray(c["kphase"])).to("cuda")
        v = np.ascontiguousarray(c["val_to_ks"].astype(np.int32))
        ri = np.ascontiguousarray(c["r_idx"].astype(np.int32))
        nr = ri.shape[0]
        D = torch.empty((nk, no, no), dtype=torch.complex128, device="cuda")
        R = torch.empty((3, nr), dtype=torch.complex128, device="cuda")
        PI_ = capi.C.POINTER(capi.C.c_int)

        def fn():
            h.check(h.lib.dmftgpu_upfold_density(h.h, n, no, nk, nk, 40.0, 0.2, G._ptr(X), G._ptr(Cb), G._ptr(eps),
                                                 v.ctypes.data_as(PI_), G._ptr(T), G._ptr(Ko), G._ptr(S), nr,
                                                 ri.ctypes.data_as(PI_), G._ptr(kp), G._ptr(D), G._ptr(R)))
        log("timing upfolding")
        t_gpu, t_wall = _events(torch, fn, args.steps, args.warmup)
        units = nk
        fl = 8.0 * (7 * no ** 3 + 2 * n ** 3) + 24.0 * nr
        metric, unit, value = "upfolded k points/s (D_k and the three R sums)", "k points/s", units / t_wall
        ach = fl * units / t_gpu / 1e12
        roof = _flop_roofline(ach, "k_zgemm_sq / k_gj_inverse / k_dm_r", "algorithmic FLOPs per k: 8(7 n_orb^3 + "
                              "2 n^3) for the products and inverses + 24 n_R for the R sums (sincos not counted) / "
                              "call time (events)", fl, units)
        if not args.no_cpu:
            def run(kk):
                O.upfold_density(c["dm_model"][:kk], c["corr_base"][:kk], c["ks_energy"][:kk], c["val_to_ks"][:kk],
                                 c["transform"][:kk], c["ks_ortho"][:kk], c["overlap"][:kk], 40.0, 0.2, c["r_idx"],
                                 c["kphase"][:kk], n_k_global=nk)
            v_, kk, tc = _cpu_sample(run, nk, lambda kk: kk, budget, 8)
            cpu = dict(value=v_, unit=unit, cores=1, kind="port",
                       sample=f"oracle upfolding_density on {kk} of {nk} k points, {nr} R entries, {tc:.1f} s")
        cfg = dict(n=n, n_orb=no, n_k=nk, n_r=nr, units_per_call=units)
        lat = h
    elif wname == "aux":
        h = G.Handle()
        f = G.to_dev_cm(sig)
        mo = torch.empty((4, nc, nc), dtype=torch.complex128, device="cuda")
        fm = lambda: h.check(h.lib.dmftgpu_moments(h.h, nc, nf, beta, G._ptr(f), 0, G._ptr(mo), None))  # noqa: E731
        log("timing moments")
        tm_gpu, tm = _events(torch, fm, args.steps, args.warmup)
        gl, _ = lat.greenftn_w_dev(s_dev, mu, G.SCHUR_F64)
        himp = G.to_dev_cm(np.mean(Hk, axis=0)[:nc, :nc][None])[0].contiguous()
        orb = (capi.C.c_int * nc)(*range(nc))
        dl = torch.empty((nf, nc, nc), dtype=torch.complex128, device="cuda")
        fh = lambda: h.check(h.lib.dmftgpu_hyb_delta(h.h, nc, nf, beta, mu, G._ptr(gl), G._ptr(s_dev),  # noqa: E731
                                                     G._ptr(himp), nc, orb, G._ptr(dl)))
        log("timing hybridisation")
        th_gpu, th = _events(torch, fh, args.steps, args.warmup)
        mu_out, nev = capi.C.c_double(0), capi.C.c_int(0)
        fmu = lambda: lat.check(lat.lib.dmftgpu_adjust_mu(lat.h, G.SCHUR_F64, G._ptr(s_dev), 0.0, w["n_target"],  # noqa: E731
                                                          1e-10, capi.C.byref(mu_out), capi.C.byref(nev)))
        log("timing mu search")
        tu_gpu, tu = _events(torch, fmu, max(1, args.steps // 2), 1)
        units = nk * 2 * nf * nev.value
        metric, unit, value = "GreenFtn_w_adjust_mu (k,iw) units/s", "(k,w) units/s", units / tu
        ach = 8.0 * n ** 3 * units / tu_gpu / 1e12
        roof = _flop_roofline(ach, "k_lattice_schur<DENS>", "algorithmic 8n^3 FLOPs per (k,iw) per N(mu) evaluation "
                              "x evaluations / call time (events)", f_fwd(n), units, plain_equivalent=True)
        extra["moments_ms"] = tm * 1e3
        extra["hyb_delta_ms"] = th * 1e3
        extra["adjust_mu"] = dict(ms=tu * 1e3, n_eval=nev.value, mu=mu_out.value)
        if not args.no_cpu:
            t0 = time.perf_counter()
            O.asymto_moments(sig, beta)
            t_om = time.perf_counter() - t0
            gl_np = G.from_dev_cm(gl)
            t0 = time.perf_counter()
            O.hyb_delta(gl_np, sig, np.mean(Hk, axis=0)[:nc, :nc], list(range(nc)), beta, mu)
            t_oh = time.perf_counter() - t0

            def run(kk):
                O.giw_to_nele(Hk[:kk], sig, mom, mu, beta, nthreads=cores)
            v_, kk, tc = _cpu_sample(run, nk, lambda kk: kk * 2 * nf, budget, max(1, cores))
            cpu = dict(value=v_, unit=unit, cores=cores, kind="port",
                       sample=f"oracle GiwToNele (one N(mu) evaluation, eigenvalue form) on {kk} of {nk} k x "
                              f"{2 * nf} frequencies, {tc:.1f} s; moments {t_om * 1e3:.1f} ms and hybridisation "
                              f"{t_oh * 1e3:.1f} ms on 1 thread",
                       moments_ms=t_om * 1e3, hyb_delta_ms=t_oh * 1e3)
        cfg = dict(n=n, n_corr=nc, n_k=nk, n_freq=nf, beta=beta, mu_tol=1e-10, units_per_call=units)
        h.close()
    cfg["workload"] = PATHS[wname]
    line = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": units / value * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": cfg, "roofline": roof,
            "cpu_baseline": cpu}
    line.update(extra)
    print(json.dumps(line), flush=True)
    lat.close()
    calibration_copy(args, torch)


if __name__ == "__main__":
    main()
