#!/usr/bin/env python3
"""Throughput of the batched band-solve path on one MI355X.

Metric: converged k-point solves per second on the res-64 sweep (the seeded
synthetic geometry generator fixed at 64 x 64 and 12 bands, 2048 geometries as
in the batched-sweep configuration, TM and TE -> 4096 band structures, 31-point
closed paths -> 30 solves per band structure, LOBPCG tolerance 1e-7).  One
step = one batched pass over the whole sweep with its inputs resident on the
device; the clock brackets exactly --steps passes after --warmup untimed ones.

Also reported on the same line:
  roofline      one kernel class (--timed-kernel: 3 = the Rayleigh-Ritz solve,
                the dominant kernel; 1 = the Theta operator; 2 = the dual
                Rayleigh-Ritz Gram; 4 = the Ritz update; 5 = the deflation)
                timed with HIP events on the
                stream each launch runs on, during the timed steps themselves:
                algorithmic bytes (and FLOPs) per launch over the average
                launch time, against 8 TB/s HBM (and 78.6 TFLOP/s FP64);
  cpu_baseline  the FP64 C++ oracle (a restatement of the reference CPU path)
                on a bounded sample of the same sweep, one band structure per
                host thread;
  accuracy      max relative frequency error of the GPU results against the
                oracle on that sample;
  plain_gpu     the same FP64 solver run one band structure at a time;
  single_solve_hex_te  one hex TE 64x64 band structure (12 bands, 61
                k-points) through the FP64 path and the FP32-preconditioner
                path, with the per-band frequency difference between them.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged k-point solves/sec, res64 12 bands, ≤1e-6 rel freq err vs FP64"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# HBM bytes per rr_kernel launch on the default workload (2048 geometries,
# slots 2048, groups 4, two engines), FETCH_SIZE x2 + WRITE_SIZE x1 from two rocprofv3 --pmc
# passes scaled by a plain copy (profiles/rr_kernel_traffic.md); reported only
# for that configuration.
RR_TRAFFIC_BYTES = 4.14e6
RR_TRAFFIC_CONFIG = (2048, 4, 2048, 2)  # slots, groups, geometries, engines the counters were read on
TIMED_NAMES = {1: "op_kernel THETA_TM/TE (64x64 fused 2D FFT operator)",
               2: "gram2_kernel (dual Rayleigh-Ritz Gram Z^H BZ, Z^H AZ)",
               3: "rr_kernel (Cholesky/SVQB + projection + Jacobi Rayleigh-Ritz, FP64)",
               4: "lincomb_kernel (Ritz update [X'|W'] = Z C, A Z C)",
               5: "deflate_kernel (residual / direction deflation against locked bands)",
               6: "transition_kernel (k-point transition: tracking, prediction, start block)"}


def solves_of(jobs):
    # a copied closing Gamma is not a solve
    return sum(j.k_path.shape[0] - 1 for j in jobs)


def rel_err(gpu_bands, ref_bands):
    got = np.sort(gpu_bands, axis=1)
    ref = np.sort(ref_bands, axis=1)
    m = ref > 1e-8
    e = np.abs(got - ref)
    e[m] /= ref[m]
    return float(e.max())


def worst_diff(res_a, res_b):
    """Where two runs' sorted frequencies differ most (relative), over k-points
    converged in both: job, k-point, band, omega of the second run, |difference|."""
    best = (0.0, None)
    for jb, (a, b) in enumerate(zip(res_a, res_b)):
        for k in range(a.bands.shape[0]):
            if not (a.converged[k] and b.converged[k]):
                continue
            x, y = np.sort(a.bands[k]), np.sort(b.bands[k])
            for i in range(len(y)):
                if y[i] > 1e-8 and abs(x[i] - y[i]) / y[i] > best[0]:
                    best = (abs(x[i] - y[i]) / y[i], (jb, k, i, float(y[i]), float(abs(x[i] - y[i]))))
    if best[1] is None:
        return None
    jb, k, i, om, d = best[1]
    return {"job": jb, "k_index": k, "band": i, "omega": om, "abs_diff": d, "rel_diff": best[0]}


def oracle_job(j, exit_rule, tol=None):
    from oracle import oracle as O
    lat = "square" if abs(j.geom.lattice.a2[0]) < 1e-12 else "hex"
    return O.Job(lattice=lat, eps_bg=j.geom.eps_bg,
                 atoms=[(a.pos[0], a.pos[1], a.radius, a.eps_inside) for a in j.geom.atoms],
                 nx=j.nx, ny=j.ny, pol=j.pol, n_bands=j.n_bands, tol=j.tol if tol is None else tol, kpath=j.k_path,
                 exit_rule=exit_rule)


def oracle_runs(jobs, exit_rule, threads, tol=None):
    """The FP64 oracle on each job (one per host thread); the ctypes calls release the GIL."""
    from oracle import oracle as O
    with cf.ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda j: O.bandstructure(oracle_job(j, exit_rule, tol)), jobs))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def large_case(be, seed=None):
    # the large combined case: 256 x 256, 32 bands, TE + TM, batched over the
    # 63 distinct k-points of the 64-point Gamma-X-M-Gamma path (126 solves,
    # each from the deterministic start), all bands converged; then the
    # 512 x 512 TM 16-band solve at M.  Timed twice: the multi-pass FFT
    # operator class (Theta) and the Rayleigh-Ritz Gram class, each with HIP
    # events on its stream; frequencies at M checked against the oracle
    # fixture.
    import json as _json
    from blaze2d_amd.bandstructure import RunOptions, run_batch
    from blaze2d_amd.sweep import large_case_jobs, large_case_m_point
    lj = large_case_jobs()
    lopts = RunOptions(exit_rule=1)
    run_batch(be, lj[:2], lopts, slots=2)  # first launches of the 256 x 256 kernels
    large = {}
    for cls in (1, 2):
        t = time.perf_counter()
        lres, lst = run_batch(be, lj, lopts, slots=min(64, len(lj)), timed_kernel=cls)
        dt = time.perf_counter() - t
        large[cls] = (lres, lst, dt)
    lres, lst, dt = large[1]
    conv_l = sum(int(r.converged.sum()) for r in lres)
    ths = large[1][1]
    grs = large[2][1]
    th_avg = ths.timed_ms * 1e-3 / max(ths.timed_launches, 1)
    gr_avg = grs.timed_ms * 1e-3 / max(grs.timed_launches, 1)
    with open(os.path.join(ROOT, "tests", "golden", "cfg5_oracle.json")) as f:
        fx = _json.load(f)
    errs = []
    for c in fx["cases"]:
        if c["grid"] == 256 and c["k"] == [0.5, 0.5]:
            i = next(i for i, j in enumerate(lj) if j.pol == c["pol"] and
                     abs(j.k_path[0][0] - 0.5) < 1e-12 and abs(j.k_path[0][1] - 0.5) < 1e-12)
            errs.append(rel_err(lres[i].bands, np.array([c["omegas"]])))
    mj = large_case_m_point()
    run_batch(be, [mj], lopts, slots=1)
    t = time.perf_counter()
    (mres,), mst = run_batch(be, [mj], lopts, slots=1, timed_kernel=1)
    dtm5 = time.perf_counter() - t
    mref = next(c for c in fx["cases"] if c["grid"] == 512)
    result = {
        "workload": "square eps_bg 12, eps 1 rods r 0.2 at (0,0) and r 0.12 at (1/2,1/2); 256x256, 32 bands, "
                    "TE+TM, 63 distinct k-points of the 64-point Gamma-X-M-Gamma path batched over k, "
                    "tol 1e-7, all bands converged, FP64",
        "value": round(conv_l / dt, 3), "unit": "k-point solves/s", "solves": len(lj),
        "converged": conv_l, "seconds": round(dt, 3),
        "lobpcg_iterations_per_solve": round(lst.iterations / max(len(lj), 1), 3),
        "max_rel_freq_err_vs_fp64_oracle_at_M": float(max(errs)) if errs else None,
        "theta_multipass": {
            "avg_launch_ms": round(th_avg * 1e3, 4), "launches": ths.timed_launches,
            "algorithmic_bytes_per_launch": int(ths.timed_bytes / max(ths.timed_launches, 1)),
            "achieved_gbs": round(ths.timed_bytes / max(ths.timed_ms * 1e-3, 1e-12) / 1e9, 1),
            "frac_hbm": round(ths.timed_bytes / max(ths.timed_ms * 1e-3, 1e-12) / 1e9 / HBM_PEAK_GBS, 4)},
        "rayleigh_ritz_gram": {
            "avg_launch_ms": round(gr_avg * 1e3, 4), "launches": grs.timed_launches,
            "algorithmic_flops_per_launch": int(grs.timed_flops / max(grs.timed_launches, 1)),
            "achieved_tflops_fp64": round(grs.timed_flops / max(grs.timed_ms * 1e-3, 1e-12) / 1e12, 3),
            "frac_fp64_peak": round(grs.timed_flops / max(grs.timed_ms * 1e-3, 1e-12) / 1e12
                                    / FP64_PEAK_TFLOPS, 4)},
        "m_point_512": {
            "workload": "512x512, TM, 16 bands, k = M, one solve",
            "seconds": round(dtm5, 3), "iterations": int(mres.iterations[0]),
            "converged": int(mres.converged[0]),
            "max_rel_freq_err_vs_fp64_oracle": rel_err(mres.bands, np.array([mref["omegas"]])),
            "theta_achieved_gbs": round(mst.timed_bytes / max(mst.timed_ms * 1e-3, 1e-12) / 1e9, 1)}}
    if seed is not None:
        result["learned_start"] = large_learned(be, seed, lj, lopts, fx)
    return result


def large_learned(be, seed, lj, lopts, fx):
    """The learned start on the large case: a 32-vector network trained at
    64 x 64 (sweep generator, 32 bands, 16,020 converged subspaces), its input
    eps evaluated on the network's 64 x 64 grid (the large grid's eps
    downsampled), its prediction synthesised on the 256 x 256 grid (prolonged)
    as the first 32 columns of each block; the 126 solves again, timed end to
    end (inference, probe iteration, fallbacks)."""
    from blaze2d_amd import learned as LS
    t = time.perf_counter()
    data = LS.training_set(be, 2670, seed=seed + 3, n_bands=32, k_per_path=3)
    t_data = time.perf_counter() - t
    net = LS.CnnNet(be, 64, 64, 32, 4, max_batch=128, seed=1)
    epochs = 20
    t = time.perf_counter()
    hist = LS.train(net, data, epochs, batch=64, lr=1e-3)
    t_train = time.perf_counter() - t
    from blaze2d_amd.bandstructure import run_batch
    t = time.perf_counter()
    _, st_default = run_batch(be, lj, lopts, slots=min(64, len(lj)))  # default start, untimed kernels
    t_default = time.perf_counter() - t
    sol = LS.solve(be, net, lj, lopts, slots=min(64, len(lj)), groups=4)
    nop = LS.solve(be, net, lj, lopts, slots=min(64, len(lj)), groups=4, fallback=False)
    net.close()
    errs = []
    for c in fx["cases"]:
        if c["grid"] == 256 and c["k"] == [0.5, 0.5]:
            i = next(i for i, j in enumerate(lj) if j.pol == c["pol"] and
                     abs(j.k_path[0][0] - 0.5) < 1e-12 and abs(j.k_path[0][1] - 0.5) < 1e-12)
            errs.append(rel_err(sol.results[i].bands, np.array([c["omegas"]])))
    n = len(lj)
    conv = sum(int(r.converged[0]) for r in sol.results)
    it_l = float(np.mean([r.iterations[0] for r in sol.results]))
    it_d = st_default.iterations / max(n, 1)
    return {
        "training": {"samples": int(data.embed.shape[0]), "n_vectors": 32, "epochs": epochs,
                     "data_seconds": round(t_data, 3), "train_seconds": round(t_train, 3),
                     "loss_first_epoch": round(hist[0], 5), "loss_last_epoch": round(hist[-1], 5)},
        "value": round(conv / sol.seconds, 3), "unit": "k-point solves/s", "converged": conv,
        "seconds": round(sol.seconds, 4), "inference_seconds": round(sol.inference_seconds, 4),
        "lobpcg_iterations_per_solve": round(it_l, 3),
        "iterations_spent_per_solve_incl_probes": round(sol.iterations_spent / n, 3),
        "fallback_rate": round(1.0 - float(sol.used_prediction.mean()), 5),
        "without_fallback_probe": {"value": round(sum(int(r.converged[0]) for r in nop.results) / nop.seconds, 3),
                                   "seconds": round(nop.seconds, 4)},
        "default_start": {"value": round(n / t_default, 3), "lobpcg_iterations_per_solve": round(it_d, 3)},
        "iteration_ratio": round(it_l / it_d, 4),
        "solves_per_s_ratio": round((conv / sol.seconds) / (n / t_default), 4),
        "max_rel_freq_err_vs_fp64_oracle_at_M": float(max(errs)) if errs else None}


def learned_case(be, seed):
    # the learned start: a network trained on the GPU on 16k converged
    # subspaces of the sweep generator (64 x 64, 12 bands; 2670 geometries x
    # TM+TE x 3 k-points), then 1024 held-out geometries x TM+TE, one
    # k-point each, solved from the default start and from the predicted
    # start with the fallback test (one probe iteration from the default
    # start per solve; solves whose predicted start is worse are solved
    # again from the default start); both timed end to end
    import dataclasses as _dc
    from blaze2d_amd import learned as LS
    from blaze2d_amd.bandstructure import RunOptions, run_batch
    t = time.perf_counter()
    ldata = LS.training_set(be, 2670, seed=seed + 1, k_per_path=3)
    t_data = time.perf_counter() - t
    lnet = LS.CnnNet(be, 64, 64, 12, 4, max_batch=256, seed=1)
    epochs = 20
    t = time.perf_counter()
    hist = LS.train(lnet, ldata, epochs, batch=64, lr=1e-3)
    t_train = time.perf_counter() - t
    held = LS.held_out_jobs(1024, seed=seed + 2)
    lro = RunOptions(exit_rule=1)
    LS.solve(be, lnet, held[:64], lro, slots=64)  # first launches
    t = time.perf_counter()
    hbase, hbst = run_batch(be, held, lro, slots=512, groups=2)
    t_def = time.perf_counter() - t
    lsol = LS.solve(be, lnet, held, lro, slots=512, groups=2)
    lnop = LS.solve(be, lnet, held, lro, slots=512, groups=2, fallback=False)
    # untimed: projections of the default-start solutions (held-out
    # loss) and a tight solve (accuracy of both starts)
    hproj, _ = run_batch(be, held, lro, slots=512, groups=2, pw_half=4)
    hg, he = LS.network_inputs(be, held)
    hloss = LS.mean_loss(lnet, hg, he, np.array([r.pw_coefs[0] for r in hproj]))
    htight, _ = run_batch(be, [_dc.replace(j, tol=1e-10) for j in held], lro, slots=512, groups=2)

    def _errs(rs):
        out_e = []
        for r, q in zip(rs, htight):
            a, b = np.sort(r.bands[0]), np.sort(q.bands[0])
            m = b > 1e-3  # the Gamma zero mode has no relative error
            out_e.append(float(np.max(np.abs(a[m] - b[m]) / b[m])))
        return np.array(out_e)

    nh = len(held)
    it_def = float(np.mean([r.iterations[0] for r in hbase]))
    it_l = float(np.mean([r.iterations[0] for r in lsol.results]))
    conv_def = sum(int(r.converged[0]) for r in hbase)
    conv_l = sum(int(r.converged[0]) for r in lsol.results)
    result = {
        "workload": "sweep generator at 64x64, 12 bands, tol 1e-7; network: 4 periodic 3x3 conv layers x 32 "
                    "channels (strides 1,2,2,2) + MLP embedding of (k1, k2, pol, lattice), linear head to "
                    "12 x 81 complex plane-wave coefficients (|n1|,|n2| <= 4), BF16 weights, FP32 "
                    "accumulation, projector loss, Adam; held out: 1024 geometries x TM+TE, one k-point "
                    "each drawn from the closed path (2048 one-k-point solves), 512 slots",
        "training": {"samples": int(ldata.embed.shape[0]), "geometries": 2670, "epochs": epochs,
                     "data_seconds": round(t_data, 3), "train_seconds": round(t_train, 3),
                     "train_samples_per_s": round(epochs * ldata.embed.shape[0] / t_train, 1),
                     "loss_first_epoch": round(hist[0], 5), "loss_last_epoch": round(hist[-1], 5),
                     "held_out_loss": round(hloss, 5)},
        "default_start": {"value": round(conv_def / t_def, 3), "unit": "k-point solves/s",
                          "lobpcg_iterations_per_solve": round(it_def, 3), "seconds": round(t_def, 4),
                          "converged": conv_def},
        "learned_start": {"value": round(conv_l / lsol.seconds, 3), "unit": "k-point solves/s",
                          "lobpcg_iterations_per_solve": round(it_l, 3),
                          "iterations_spent_per_solve_incl_probes": round(lsol.iterations_spent / nh, 3),
                          "seconds": round(lsol.seconds, 4),
                          "inference_seconds": round(lsol.inference_seconds, 4),
                          "fallback_rate": round(1.0 - float(lsol.used_prediction.mean()), 5),
                          "converged": conv_l},
        "learned_start_without_fallback_probe": {
            "value": round(sum(int(r.converged[0]) for r in lnop.results) / lnop.seconds, 3),
            "unit": "k-point solves/s", "seconds": round(lnop.seconds, 4)},
        "iteration_ratio": round(it_l / it_def, 4),
        "solves_per_s_ratio": round((conv_l / lsol.seconds) / (conv_def / t_def), 4),
        # against a tight solve: both starts stop on the reference's
        # relative eigenvalue-change test at 1e-7, which a start close
        # to the answer can pass early on a slowly moving top band
        "rel_freq_err_vs_tol_1e-10": {
            "default_start": {"max": float(_errs(hbase).max()),
                              "solves_above_1e-6": int((_errs(hbase) > 1e-6).sum())},
            "learned_start": {"max": float(_errs(lsol.results).max()),
                              "solves_above_1e-6": int((_errs(lsol.results) > 1e-6).sum())}}}
    lnet.close()
    return result

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--geoms", type=int, default=2048, help="geometries in the sweep (x2 polarisations)")
    ap.add_argument("--slots", type=int, default=2048, help="concurrent band structures (0 = all)")
    ap.add_argument("--groups", type=int, default=4, help="slot groups pipelined on separate streams")
    ap.add_argument("--engines", type=int, default=2,
                    help="engines run side by side, each with its share of the jobs and slots")
    ap.add_argument("--precision", type=int, default=0, choices=(0, 1, 2),
                    help="0 FP64; 1 FP32 fields; 2 FP64 fields with FP32 preconditioners")
    ap.add_argument("--seed", type=int, default=0xB1A2E2D)
    ap.add_argument("--cpu-sample", type=int, default=16,
                    help="band structures timed on the CPU oracle, one per thread (16 threads at most)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip roofline / plain-GPU extra passes")
    ap.add_argument("--no-large", action="store_true", help="skip the large combined case (256x256 x 32 bands, 512x512)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned start (network training and the held-out evaluation)")
    ap.add_argument("--large-only", action="store_true",
                    help="only run the large combined case and print its figures (for profiling)")
    ap.add_argument("--learned-only", action="store_true",
                    help="only run the learned start (training and held-out evaluation) and print its figures")
    ap.add_argument("--timed-kernel", type=int, default=3, choices=(1, 2, 3, 4, 5, 6),
                    help="kernel class timed for the roofline: 1 Theta operator, 2 dual Rayleigh-Ritz Gram, "
                         "3 Rayleigh-Ritz solve (the dominant kernel), 4 Ritz-update linear combinations")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 512 MiB device-to-device copy five times (scale for PMC byte counters)")
    ap.add_argument("--exit-rule", type=int, default=1,
                    help="1: a solve ends when every band is soft-locked (all bands converged); "
                         "0: the reference solve() exit accounting")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda:0")  # 512 MiB
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": src.numel() * 8, "copies": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (band structures are "
              "independent, so a multi-GPU run would give each rank its share of the sweep)", file=sys.stderr)
        sys.exit(2)

    from blaze2d_amd import HipBackend
    from blaze2d_amd.bandstructure import RunOptions, run_batch
    from blaze2d_amd.sweep import res64_sweep

    if args.large_only or args.learned_only:
        be = HipBackend(0)
        part = {}
        if args.large_only:
            part["large_case"] = large_case(be, None if args.no_learned else args.seed)
        if args.learned_only:
            part["learned_start"] = learned_case(be, args.seed)
        print(json.dumps(part))
        be.close()
        return

    ropts = RunOptions(exit_rule=args.exit_rule)

    jobs = res64_sweep(args.geoms, seed=args.seed, tol=1e-7)
    slots = min(args.slots, len(jobs)) if args.slots > 0 else len(jobs)
    be = HipBackend(0)
    print(f"[bench] sweep: {args.geoms} geometries x 2 pols = {len(jobs)} band structures, "
          f"{solves_of(jobs)} solves per step, slots={slots}", file=sys.stderr, flush=True)

    for w in range(args.warmup):
        t = time.perf_counter()
        res, st = run_batch(be, jobs, ropts, slots=slots, precision=args.precision, groups=args.groups,
                            engines=args.engines)
        print(f"[bench] warmup {w}: {time.perf_counter() - t:.2f}s", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    conv = 0
    tot_iters = 0
    last = None
    timed = {"ms": 0.0, "launches": 0, "bytes": 0.0, "flops": 0.0, "flops_executed": 0.0}
    rr_stats = {"problems": 0, "svqb_path": 0, "rank_drops": 0}
    alg_bytes = 0.0
    for k in range(args.steps):
        res, st = run_batch(be, jobs, ropts, slots=slots, precision=args.precision, groups=args.groups,
                            timed_kernel=args.timed_kernel, engines=args.engines)
        conv += sum(int(r.converged[:-1].sum()) for r in res)
        tot_iters += st.iterations
        timed["ms"] += st.timed_ms
        timed["launches"] += st.timed_launches
        timed["bytes"] += st.timed_bytes
        timed["flops"] += st.timed_flops
        timed["flops_executed"] += st.timed_flops_executed
        rr_stats["problems"] += st.rr_problems
        rr_stats["svqb_path"] += st.svqb_fallbacks
        rr_stats["rank_drops"] += st.rank_drops
        alg_bytes += st.alg_bytes
        last = res
        print(f"[bench] step {k}: engine {st.seconds:.2f}s, {st.steps} engine steps, host plan "
              f"{st.host_plan_seconds:.2f}s, host wait {st.host_wait_seconds:.2f}s", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    expected = args.steps * solves_of(jobs)
    value = conv / elapsed
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "k-point solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / max(args.steps, 1), 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "f64", 1: "f32 fields, f64 accumulation", 2: "f64, f32 preconditioners"}[args.precision],
        "data": "synthetic (seeded geometry sweep, no external data)",
        "config": {
            "workload": f"res64 sweep: {args.geoms} seeded geometries (square/hex, eps_bg~U[2,16], 1-2 inclusions) "
                        f"x TM+TE, 64x64 grid, 12 bands, 31-point closed path (30 solves each), tol 1e-7",
            "band_structures": len(jobs),
            "solves_per_step": solves_of(jobs),
            "converged_fraction": round(conv / max(expected, 1), 6),
            "lobpcg_iterations_per_solve": round(tot_iters / max(args.steps * solves_of(jobs), 1), 3),
            "slots": slots,
            "groups": args.groups,
            "engines": args.engines,
            "seed": args.seed,
            "exit_rule": "all bands soft-locked" if args.exit_rule == 1 else "reference solve() accounting",
            "rayleigh_ritz_problems": rr_stats["problems"],
            "rayleigh_ritz_svqb_path": rr_stats["svqb_path"],
            "rayleigh_ritz_rank_drops": rr_stats["rank_drops"],
        },
    }

    if timed["launches"] > 0 and timed["ms"] > 0:
        avg_s = timed["ms"] * 1e-3 / timed["launches"]
        bytes_per = timed["bytes"] / timed["launches"]
        flops_per = timed["flops"] / timed["launches"]
        gbs = bytes_per / avg_s / 1e9
        tfs = flops_per / avg_s / 1e12
        intensity = flops_per / max(bytes_per, 1.0)
        ridge = FP64_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9)
        hbm = intensity < ridge
        # FP64 peak: MI355X's dense FP64 rate (matrix and vector alike); the
        # timed kernels issue FP64 vector FMAs, no MFMA
        out["roofline"] = {
            "bound": "hbm" if hbm else "fp64",
            "achieved": round(gbs if hbm else tfs, 3),
            "peak": HBM_PEAK_GBS if hbm else FP64_PEAK_TFLOPS,
            "unit": "GB/s" if hbm else "TFLOP/s",
            "frac": round(gbs / HBM_PEAK_GBS if hbm else tfs / FP64_PEAK_TFLOPS, 5),
            "traffic": (RR_TRAFFIC_BYTES if (args.timed_kernel == 3 and args.slots == RR_TRAFFIC_CONFIG[0]
                                              and args.groups == RR_TRAFFIC_CONFIG[1]
                                              and args.geoms == RR_TRAFFIC_CONFIG[2]
                                              and args.engines == RR_TRAFFIC_CONFIG[3] and args.precision == 0)
                        else None),
            "kernel": TIMED_NAMES[args.timed_kernel],
            "avg_launch_ms": round(avg_s * 1e3, 4),
            "launches": timed["launches"],
            "algorithmic_bytes_per_launch": int(bytes_per),
            "algorithmic_flops_per_launch": int(flops_per),
            "achieved_tflops_fp64": round(tfs, 3),
            "frac_fp64_peak": round(tfs / FP64_PEAK_TFLOPS, 5),
            "flop_model": ("method-independent: Cholesky, inverse and projection 4n^3/3 complex multiply-adds, "
                           "Hermitian eigendecomposition with vectors 30n^3, back-transform 4n^2k"
                           if args.timed_kernel == 3 else "algorithmic, per kernel class"),
            "executed_flops_per_launch": int(timed["flops_executed"] / timed["launches"]),
            "executed_tflops_fp64": round(timed["flops_executed"] / timed["launches"] / avg_s / 1e12, 3),
        }
    if alg_bytes > 0:
        # the whole solve path against HBM: algorithmic bytes of every launch of
        # the timed steps (each column read or written once per launch) over
        # their wall time
        out["engine_hbm"] = {
            "algorithmic_bytes_per_step": int(alg_bytes / max(args.steps, 1)),
            "achieved_gbs": round(alg_bytes / elapsed / 1e9, 1),
            "peak_gbs": HBM_PEAK_GBS,
            "frac": round(alg_bytes / elapsed / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_per_lobpcg_iteration": int(alg_bytes / max(tot_iters, 1)),
        }
    if not args.no_extra:
        # dielectric tables of the whole sweep on the GPU (one launch, event-timed
        # kernel) against the host path on a sample, one thread
        from blaze2d_amd.bandstructure import dielectric, dielectric_batch
        geoms = [j.geom for j in jobs[::2]]
        dielectric_batch(be, geoms[:8], 64, 64, True)  # first-launch cost out of the way
        _, _, _, kms = dielectric_batch(be, geoms, 64, 64, True)
        t = time.perf_counter()
        for g in geoms[:32]:
            dielectric(g, 64, 64, True)
        host_s = (time.perf_counter() - t) / 32
        out["dielectric_gpu"] = {
            "value": round(len(geoms) * 4096 / (kms * 1e-3), 1), "unit": "smoothed nodes/s",
            "kernel_ms": round(kms, 4), "geometries": len(geoms), "grid": "64x64",
            "host_one_thread_nodes_per_s": round(4096 / host_s, 1)}
        # the reference's one-time per-job operator diagnostics (kappa estimates),
        # batched over the whole sweep; with_diagnostics adds their wall time to
        # one step of the headline, as the reference's session totals include them
        from blaze2d_amd.bandstructure import operator_diagnostics
        operator_diagnostics(be, jobs[:8])
        t = time.perf_counter()
        diags, dms = operator_diagnostics(be, jobs)
        dwall = time.perf_counter() - t
        step_s = elapsed / max(args.steps, 1)
        out["operator_diagnostics_gpu"] = {
            "value": round(len(jobs) / dwall, 1), "unit": "jobs/s", "jobs": len(jobs),
            "device_ms": round(dms, 2), "wall_ms": round(1000 * dwall, 2),
            "power_iterations": 15,
            "headline_with_diagnostics": round(value * step_s / (step_s + dwall), 3),
            "max_self_adjoint_err": max(d.self_adjoint_err for d in diags)}
        # the whole synthetic-sweep configuration: grid drawn from {32, 48, 64} and
        # bands from {8, 12, 16} per geometry, in one call: one engine per grid
        # size (2 slot groups each) side by side, 3072 slots shared by work
        from blaze2d_amd.sweep import sweep_jobs
        mixed = sweep_jobs(args.geoms, seed=args.seed, tol=1e-7)
        by_grid = {}
        for i, j in enumerate(mixed):
            by_grid.setdefault(j.nx, []).append(i)
        for g, idx in sorted(by_grid.items()):  # first launches of each grid size out of the way
            run_batch(be, [mixed[i] for i in idx[:2]], ropts, slots=2, precision=args.precision)
        t = time.perf_counter()
        rmix, _ = run_batch(be, mixed, ropts, slots=3072, precision=args.precision, groups=2, engines=len(by_grid))
        dtm = time.perf_counter() - t
        conv_m = sum(int(x.converged[:-1].sum()) for x in rmix)
        mixed_res = {g: [rmix[i] for i in idx] for g, idx in by_grid.items()}
        by_grid = {g: [mixed[i] for i in idx] for g, idx in by_grid.items()}
        out["mixed_sweep"] = {
            "value": round(conv_m / dtm, 3), "unit": "k-point solves/s", "band_structures": len(mixed),
            "solves": solves_of(mixed), "converged_fraction": round(conv_m / max(solves_of(mixed), 1), 6),
            "band_structures_per_grid": {str(g): len(js) for g, js in sorted(by_grid.items())},
            "bands": "8/12/16 drawn per geometry", "seconds": round(dtm, 3),
            "engines": "one per grid size, 2 slot groups each, 3072 slots shared by work"}
        if not args.no_cpu:
            # accuracy on a sample: the first four band structures of each grid
            # size (both polarisations of two geometries, their band counts as drawn)
            samp = [(g, i) for g in sorted(by_grid) for i in range(min(4, len(by_grid[g])))]
            refs = oracle_runs([by_grid[g][i] for g, i in samp], args.exit_rule, min(16, os.cpu_count() or 1))
            errs = [rel_err(mixed_res[g][i].bands, r[0]) for (g, i), r in zip(samp, refs)]
            out["mixed_sweep"]["max_rel_freq_err_vs_fp64_oracle"] = float(max(errs))
            out["mixed_sweep"]["accuracy_sample"] = (
                f"{len(samp)} band structures (first 4 of each grid), bands "
                f"{sorted(set(by_grid[g][i].n_bands for g, i in samp))}, oracle under the same exit rule")
        # plain FP64 path, one band structure at a time (the batching baseline)
        pj = jobs[:2]
        t = time.perf_counter()
        run_batch(be, pj, ropts, slots=1, precision=0)
        dtp = time.perf_counter() - t
        out["plain_gpu"] = {"value": round(solves_of(pj) / dtp, 3), "unit": "k-point solves/s",
                            "sample": "first 2 band structures of the sweep, slots=1, FP64"}
        # one band structure on its own: hex lattice, eps_bg 13 with an eps 1 hole
        # of r = 0.3, TE, 64x64, 12 bands, Gamma-M-K-Gamma at 61 k-points, tol 1e-7;
        # the FP64 path against the FP32-preconditioner path, per band
        from blaze2d_amd.bandstructure import (BandStructureJob, BasisAtom, Geometry2D, Lattice2D,
                                               k_path as make_k_path)
        hj = BandStructureJob(Geometry2D(Lattice2D.hexagonal(), 13.0, [BasisAtom((0.5, 0.5), 0.3, 1.0)]), 64, 64,
                              "TE", make_k_path("hex", 20), n_bands=12, tol=1e-7)
        single = {}
        for prec in (0, 2):
            run_batch(be, [hj], ropts, slots=1, precision=prec)  # warm-up
            t = time.perf_counter()
            (hr,), hst = run_batch(be, [hj], ropts, slots=1, precision=prec)
            single[prec] = (hr, time.perf_counter() - t, hst.iterations)
        b64, b32 = np.sort(single[0][0].bands, axis=1), np.sort(single[2][0].bands, axis=1)
        pos = b64 > 1e-8
        per_band = [float(np.max(np.abs(b32[:, i] - b64[:, i])[pos[:, i]] / b64[:, i][pos[:, i]]))
                    if pos[:, i].any() else 0.0 for i in range(b64.shape[1])]
        out["single_solve_hex_te"] = {
            "workload": "hex eps_bg 13, eps 1 hole r 0.3, TE, 64x64, 12 bands, Gamma-M-K-Gamma 61 k-points, tol 1e-7, "
                        "one slot",
            "exit_rule": out["config"]["exit_rule"],
            "fp64": {"seconds": round(single[0][1], 4), "k_points_per_s": round(61 / single[0][1], 2),
                     "lobpcg_iterations": single[0][2], "converged": int(single[0][0].converged.sum())},
            "fp32_preconditioners": {"seconds": round(single[2][1], 4), "k_points_per_s": round(61 / single[2][1], 2),
                                     "lobpcg_iterations": single[2][2],
                                     "converged": int(single[2][0].converged.sum())},
            "max_rel_freq_diff_per_band": [float(f"{e:.3e}") for e in per_band]}
        if not args.no_large:
            out["large_case"] = large_case(be, None if args.no_learned else args.seed)
        if not args.no_learned:
            out["learned_start"] = learned_case(be, args.seed)
        # the reference's own exit accounting on the same sweep (one pass)
        if args.exit_rule != 0:
            t = time.perf_counter()
            resf, stf = run_batch(be, jobs, RunOptions(exit_rule=0), slots=slots, precision=args.precision,
                                  groups=args.groups, engines=args.engines)
            dtf = time.perf_counter() - t
            out["solve_exit_mode"] = {
                "value": round(sum(int(r.converged[:-1].sum()) for r in resf) / dtf, 3),
                "unit": "k-point solves/s",
                "lobpcg_iterations_per_solve": round(stf.iterations / max(solves_of(jobs), 1), 3)}
            if not args.no_cpu:
                # against the oracle under the same (reference) exit accounting, and
                # against a tightly converged oracle run (tol 1e-11, all bands),
                # which shows how far that accounting leaves the top bands
                ns = min(8, len(jobs))
                nt = min(16, os.cpu_count() or 1)
                r0 = oracle_runs(jobs[:ns], 0, nt)
                rt = oracle_runs(jobs[:ns], 1, nt, tol=1e-11)
                e_same = max(rel_err(resf[i].bands, r0[i][0]) for i in range(ns))
                e_conv = max(rel_err(resf[i].bands, rt[i][0]) for i in range(ns))
                e_orc = max(rel_err(r0[i][0], rt[i][0]) for i in range(ns))
                # bands the accounting left settled (oracle within 1e-7 of converged)
                settled = []
                for i in range(ns):
                    a = np.sort(resf[i].bands, axis=1)
                    o = np.sort(r0[i][0], axis=1)
                    c = np.sort(rt[i][0], axis=1)
                    m = (c > 1e-8) & (np.abs(o - c) <= 1e-7 * c)
                    settled.append(float((np.abs(a - o)[m] / o[m]).max()) if m.any() else 0.0)
                out["solve_exit_mode"].update({
                    "max_rel_freq_err_vs_fp64_oracle_same_rule": e_same,
                    "max_rel_freq_err_vs_fp64_oracle_same_rule_settled_bands": max(settled),
                    "max_rel_freq_err_vs_converged_fp64": e_conv,
                    "oracle_same_rule_vs_converged_fp64": e_orc,
                    "accuracy_sample": f"first {ns} band structures of the sweep"})
        # FP64 fields with the FD / TP preconditioners in FP32 (precision 2), one pass
        if args.precision == 0:
            t = time.perf_counter()
            resm, stm = run_batch(be, jobs, ropts, slots=slots, precision=2, groups=args.groups,
                                  engines=args.engines)
            dtm = time.perf_counter() - t
            out["fp32_preconditioners"] = {
                "value": round(sum(int(r.converged[:-1].sum()) for r in resm) / dtm, 3),
                "unit": "k-point solves/s",
                "lobpcg_iterations_per_solve": round(stm.iterations / max(solves_of(jobs), 1), 3),
                # over the k-points both runs converged (the unconverged ones, at the
                # iteration cap, are not counted in either throughput)
                "max_rel_freq_diff_vs_fp64_run": rel_err(
                    np.concatenate([r.bands[(r.converged != 0) & (q.converged != 0)] for r, q in zip(resm, last)]),
                    np.concatenate([q.bands[(r.converged != 0) & (q.converged != 0)] for r, q in zip(resm, last)])),
                "k_points_converged_in_both": int(sum(((r.converged != 0) & (q.converged != 0)).sum()
                                                      for r, q in zip(resm, last))),
                "worst": worst_diff(resm, last)}

    if not args.no_cpu and args.cpu_sample > 0:
        from oracle import oracle as O

        sample = jobs[: args.cpu_sample]

        def run_one(j):
            lat = "square" if abs(j.geom.lattice.a2[0]) < 1e-12 else "hex"
            oj = O.Job(lattice=lat, eps_bg=j.geom.eps_bg,
                       atoms=[(a.pos[0], a.pos[1], a.radius, a.eps_inside) for a in j.geom.atoms],
                       nx=j.nx, ny=j.ny, pol=j.pol, n_bands=j.n_bands, tol=j.tol, kpath=j.k_path,
                       exit_rule=args.exit_rule)
            return O.bandstructure(oj)

        threads = min(len(sample), os.cpu_count() or 1, 16)
        t = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            refs = list(ex.map(run_one, sample))
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {
            "value": round(solves_of(sample) / dt, 3),
            "unit": "k-point solves/s",
            "cores": threads,
            "kind": "port",
            "sample": f"first {len(sample)} band structures of the sweep ({solves_of(sample)} solves), "
                      f"FP64 C++ oracle, one band structure per thread",
        }
        # the same oracle on one host core: its first two band structures (TM, TE) in turn
        one = sample[:2]
        t = time.perf_counter()
        for j in one:
            run_one(j)
        dt1 = time.perf_counter() - t
        out["cpu_baseline"].update({
            "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count()},
            "one_core": {"value": round(solves_of(one) / dt1, 3), "unit": "k-point solves/s", "cores": 1,
                         "sample": f"first {len(one)} band structures of the sweep ({solves_of(one)} solves)"},
            "reference_published_context": "reference CPU path, one core, f64, tol 1e-6, res 64, 12 bands: "
                                           "~40.8 TM / ~17.1 TE k-point solves/s (hardware unstated)"})
        errs = [rel_err(last[i].bands, refs[i][0]) for i in range(len(sample))]
        out["accuracy"] = {"max_rel_freq_err_vs_fp64_oracle": float(max(errs)),
                           "sample_band_structures": len(sample)}

    print(json.dumps(out))


if __name__ == "__main__":
    main()
