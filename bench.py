"""Benchmark: batched MNDO SCF throughput on one MI355X.

Workload: 4096 water clusters (H2O)_n, n ~ U{2..64} from seed 20240607
(12..384 basis functions each), SCF to |dE| < 1e-6 eV.  One step = the whole
hot path over the batch with the geometries already resident in HBM:
two-center integrals, Hcore, core-core, guess, and every SCF iteration of
every molecule until all have converged.  value = molecules converged within
1e-4 eV of the batched FP64 result / step time.

Printed JSON (one line): the driver's contract fields plus
  roofline          the dominant kernel (Householder tridiagonalisation, FP64
                    launch) rate: algorithmic flops of its FP64 molecules per
                    launch / average launch time, the time from HIP events around
                    each launch on the library's stream during the timed steps
  tridiag_fp32_roofline  the same for the FP32 launch of the mixed policy
  fock_roofline     the Fock build's HBM rate (algorithmic bytes / kernel time)
  density_roofline  the density product's flop rate (algorithmic flops / time)
  plain_fp64        one molecule at a time through the same kernels (subsample)
  serving           continuous batching under Poisson arrivals: p50/p99
                    arrival-to-result latency and molecules/s
  cpu_baseline      the oracle in reference-algorithm mode (O(N^4) Fock, dense
                    eigensolver) timed on this host on a bounded sample and
                    extrapolated to the batch
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The serving loop runs each size class on its own stream.  HIP maps a
# process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default) and
# streams sharing a queue run in sequence: after the batch run (graph capture
# and replay) 4 queues left two of the three classes on one queue (their p50
# 3.7 s instead of 0.5 s).  8 gives every stream its own queue.  Set before
# the HIP runtime starts (the first HIP call).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"


from semigpu_amd import api, geometry  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 matrix
METRIC = "SCF molecules/s batched mixed vs plain FP64 at |dE|<=1e-4 eV; p50/p99 latency"
PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_tridiag_water4096_pmc.md")


def workload(count: int, seed: int):
    return geometry.water_batch(count, seed=seed)


def pair_block_doubles(mol) -> int:
    """Stored pair-integral doubles I = 100 n_XX + 10 n_XH + n_HH over i >= j."""
    heavy = sum(1 for a in mol if a[0] > 1)
    h = len(mol) - heavy
    return (100 * heavy * (heavy + 1) // 2 + 10 * heavy * h + h * (h + 1) // 2)


def n_occ(mol) -> int:
    return geometry.n_electrons(mol) // 2


def run_step(b: api.Batch, threshold: float, policy: str = "mixed"):
    b.two_center()
    b.hcore()
    b.core_core()
    return b.scf(threshold=threshold, max_steps=100, policy=policy)


def pmc_traffic(key: str = "traffic_bytes_per_launch"):
    """HBM bytes per tridiagonalisation launch (FP64 launch, or key
    traffic_bytes_per_launch_fp32 for the FP32 one) from the committed PMC
    summary."""
    try:
        with open(PMC_SUMMARY) as f:
            m = re.search(key + r":\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def cpu_baseline(mols, steps, budget_s: float = 25.0):
    """Oracle in reference-algorithm mode on a bounded sample, extrapolated.

    Full SCFs (O(N^4) element-wise Fock as twoelectronmatrix.cpp:14-35, dense
    Householder + QR eigensolver) of (H2O)_n for growing n until the budget is
    spent; the per-Fock-build constant c = t / ((steps + 1) N^4) of the largest
    runs extrapolates to every molecule of the batch with its GPU step count.
    """
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    t_start = time.time()
    runs = []
    for n in (2, 4, 6, 8, 10, 12, 14, 16, 18, 20):
        mol = geometry.water_cluster(n, geometry.SplitMix64(1000 + n))
        N = geometry.n_basis(mol)
        est = 1.5 * 13 * (runs[-1][3] / ((runs[-1][1] + 1) * runs[-1][0] ** 4)) * N ** 4 if runs else 0
        if time.time() - t_start + est > budget_s:
            break
        t0 = time.perf_counter()
        r = po.Molecule(mol).scf(threshold=1e-6, mode="ref")
        runs.append((N, r["steps"], n, time.perf_counter() - t0))
    big = [x for x in runs if x[0] >= 48] or runs[-1:]
    c = float(np.median([t / ((st + 1) * N ** 4) for N, st, _, t in big]))
    total = sum((st + 1) * c * geometry.n_basis(m) ** 4 for m, st in zip(mols, steps))
    return {
        "value": round(len(mols) / total, 6),
        "unit": "molecules/s",
        "cores": 1,
        "kind": "port",
        "sample": ("oracle C restatement in reference-algorithm mode (O(N^4) element-wise Fock, "
                   "twoelectronmatrix.cpp:14-35; dense Householder+QR eigensolver), 1 core: full "
                   "SCFs to 1e-6 eV of (H2O)_n n=%s (N=%s, %.1f s of CPU work); c=%.3e s per N^4 "
                   "Fock build, extrapolated to the %d-cluster batch with the GPU step counts"
                   % ([x[2] for x in runs], [x[0] for x in runs], sum(x[3] for x in runs), c,
                      len(mols))),
        "sample_mol_per_s": round(len(runs) / sum(x[3] for x in runs), 4),
    }


# basis-size bounds of the server's size classes (each its own stream and pace)
SERVE_CLASSES = (300, 700)


def serve(eng, rate: float, seconds: float, threshold: float, policy: str, seed: int = 7,
          limit_s: float = 90.0, diis: int = 0):
    """Poisson arrivals of alkanes C_nH_2n+2 (n <= 60) and water clusters (n <= 200)."""
    reqs = geometry.serving_stream(rate, seconds, seed)
    srv = eng.server(hbm_bytes=64e9, max_slots=16384, max_atoms=1 << 21, threshold=threshold,
                     max_steps=100, policy=policy, classes=SERVE_CLASSES)
    if diis:
        srv.set_diis(diis)
    done = []
    i = 0
    t0 = srv.clock()
    steps = 0
    while True:
        now = srv.clock() - t0
        while i < len(reqs) and reqs[i][0] <= now:
            srv.submit(i, reqs[i][2])
            i += 1
        n_in, n_q = srv.step()
        steps += 1
        done.extend(srv.poll())
        if i == len(reqs) and n_in == 0 and n_q == 0:
            break
        if now > limit_s:
            break
        if n_in == 0 and n_q == 0 and i < len(reqs):
            time.sleep(max(0.0, reqs[i][0] - (srv.clock() - t0)))
    srv.close()
    lat = np.array([r["t_done"] - t0 - reqs[r["id"]][0] for r in done])
    t_end = max(r["t_done"] for r in done) - t0 if done else 0.0
    conv = sum(1 for r in done if r["converged"])
    edges = (0,) + tuple(SERVE_CLASSES) + (1 << 30,)
    by_class = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        lc = np.array([r["t_done"] - t0 - reqs[r["id"]][0] for r in done if lo <= r["n_ao"] < hi])
        by_class.append({"n_ao": [lo, min(hi, 1280)], "completed": int(len(lc)),
                         "p50_ms": round(float(np.percentile(lc, 50)) * 1e3, 2) if len(lc) else None,
                         "p99_ms": round(float(np.percentile(lc, 99)) * 1e3, 2) if len(lc) else None})
    return {"rate_per_s": rate, "window_s": seconds, "requests": len(reqs), "completed": len(done),
            "converged": conv, "server_steps": steps,
            "value": round(len(done) / max(t_end, 1e-9), 3), "unit": "molecules/s",
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2) if len(lat) else None,
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2) if len(lat) else None,
            "size_classes": list(SERVE_CLASSES), "by_class": by_class,
            "basis": "arrival time to result on the server clock; alkanes n~U{1..60} / water "
                     "clusters n~U{2..200} 50/50, seed %d" % seed}


def calib_copy(eng, nbytes: int = 1 << 30):
    """One plain device copy kernel of known size, the scale for the PMC byte counters."""
    eng.calibration_copy(nbytes)
    print("k_calibration_copy: %d bytes read, %d bytes written" % (nbytes, nbytes),
          file=sys.stderr, flush=True)


def water_box(args):
    """The large single system: 1000 H2O (N = 6000) on the multi-workgroup
    eigensolver, FP64 and mixed SCF from the same resident batch."""
    eng = api.Engine(0)
    mol = geometry.water_box()
    b = eng.batch([mol])
    out = {"workload": "water box, 1000 H2O on a 10x10x10 lattice (3.1 A, +-0.2 A jitter), "
                       "N = %d, SCF to %g eV" % (b.n_ao(0), args.threshold)}
    eng.timing(True)
    for pol in ("fp64", "mixed"):
        eng.timing_reset()
        t = time.perf_counter()
        r = b.scf(threshold=args.threshold, max_steps=args.box_max_steps, policy=pol)[0]
        dt = time.perf_counter() - t
        st = eng.timing_read()
        out[pol] = {"seconds": round(dt, 3), "steps": r["steps"], "converged": r["converged"],
                    "e_tot_eV": r["e_tot"],
                    "stage_ms": {k: round(v[0], 2) for k, v in st.items()},
                    "tridiag_ms_per_iteration": round(st["tridiag"][0] / max(1, st["tridiag"][1]), 2)}
        print("water box %s: %d steps, %.2f s" % (pol, r["steps"], dt), file=sys.stderr, flush=True)
    out["mixed_minus_fp64_eV"] = out["mixed"]["e_tot_eV"] - out["fp64"]["e_tot_eV"]
    if args.diis:
        b.set_diis(args.diis)
        for pol in ("fp64", "mixed"):
            eng.timing_reset()
            t = time.perf_counter()
            r = b.scf(threshold=args.threshold, max_steps=args.box_max_steps, policy=pol)[0]
            dt = time.perf_counter() - t
            key = pol + "_diis%d" % args.diis
            out[key] = {"seconds": round(dt, 3), "steps": r["steps"], "converged": r["converged"],
                        "e_tot_eV": r["e_tot"],
                        "minus_fp64_eV": r["e_tot"] - out["fp64"]["e_tot_eV"]}
            print("water box %s: %d steps, %.2f s" % (key, r["steps"], dt), file=sys.stderr,
                  flush=True)
    eng.timing(False)
    b.close()
    eng.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=20240607)
    ap.add_argument("--threshold", type=float, default=1e-6)
    ap.add_argument("--plain-sample", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--serve-extra", default="",
                    help="more arrival rates, RATE:SECONDS[,RATE:SECONDS...], reported under "
                         "serving_rates (e.g. 1000:1,5000:1)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 1 GiB device copy (PMC byte-counter calibration)")
    ap.add_argument("--policy", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--water-box", action="store_true",
                    help="instead: the 1000-water box (N = 6000), FP64 and mixed SCF, one JSON line")
    ap.add_argument("--box-max-steps", type=int, default=60)
    ap.add_argument("--diis", type=int, default=6,
                    help="also time the batch and the serving stream with Pulay DIIS of this "
                         "depth (an SCF aid the reference does not have: reported beside value, "
                         "never as it; 0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(molecules of a batch are independent and would be split across GPUs).",
              file=sys.stderr)
        sys.exit(2)

    if args.water_box:
        water_box(args)
        return

    mols = workload(args.count, args.seed)
    sizes = np.array([geometry.n_basis(m) for m in mols])
    occ = np.array([n_occ(m) for m in mols])
    eng = api.Engine(0)
    if args.calib_copy:
        calib_copy(eng)
    t0 = time.time()
    b = eng.batch(mols)
    upload_s = time.time() - t0
    print("batch of %d molecules resident (sum N=%d, max N=%d) in %.1f s"
          % (len(mols), sizes.sum(), sizes.max(), upload_s), file=sys.stderr, flush=True)

    for w in range(args.warmup):
        t = time.time()
        run_step(b, args.threshold, args.policy)
        print("warmup %d: %.3f s" % (w, time.time() - t), file=sys.stderr, flush=True)

    eng.timing(True)
    eng.timing_reset()
    times = []
    res = None
    for k in range(args.steps):
        t = time.perf_counter()
        res = run_step(b, args.threshold, args.policy)
        times.append(time.perf_counter() - t)
        print("step %d: %.3f s" % (k, times[-1]), file=sys.stderr, flush=True)
    stage = eng.timing_read()
    eng.timing(False)
    step_s = float(np.mean(times))
    value = len(mols) / step_s
    steps = np.array([r["steps"] for r in res])
    conv = np.array([r["converged"] for r in res])

    # Algorithmic work per launch = work of one timed step / launches of one step.
    # The FP64 reduction launch takes each SCF step's FP64 molecules (FP32 steps
    # of the mixed policy go to their own launch, timed as tridiag_fp32, unless
    # the batch is small enough for one launch to take both).
    K = max(1, args.steps)
    N3 = sizes.astype(np.float64) ** 3
    s32 = b.fp32_steps().astype(np.float64)
    tri_ms, tri_cnt = stage["tridiag"]
    t32_ms, t32_cnt = stage["tridiag_fp32"]
    steps64 = steps - s32 if t32_cnt else steps
    tri_flops_step = float(np.sum(steps64 * (4.0 / 3.0) * N3))
    tri_avg_ms = tri_ms / max(1, tri_cnt)
    tri_flops_launch = tri_flops_step * K / max(1, tri_cnt)
    tri_tflops = tri_flops_launch / (tri_avg_ms * 1e-3) / 1e12
    traffic = pmc_traffic()
    roofline = {"kernel": "k_tridiag<6, double>", "bound": "mfma", "achieved": round(tri_tflops, 4),
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tri_tflops / FP64_PEAK_TFLOPS, 6), "traffic": traffic,
                "avg_launch_ms": round(tri_avg_ms, 4),
                "flops_per_launch": tri_flops_launch,
                "work": "(4/3) N^3 flops per molecule in an FP64 step, per launch",
                "traffic_rate_GBs": (round(traffic / (tri_avg_ms * 1e-3) / 1e9, 1)
                                     if traffic else None)}
    tri32_roof = None
    if t32_cnt:
        t32_avg_ms = t32_ms / t32_cnt
        t32_flops_launch = float(np.sum(s32 * (4.0 / 3.0) * N3)) * K / t32_cnt
        t32_tflops = t32_flops_launch / (t32_avg_ms * 1e-3) / 1e12
        traffic32 = pmc_traffic("traffic_bytes_per_launch_fp32")
        tri32_roof = {"kernel": "k_tridiag<6, float>", "bound": "mfma",
                      "achieved": round(t32_tflops, 4), "peak": FP32_PEAK_TFLOPS,
                      "unit": "TFLOP/s", "frac": round(t32_tflops / FP32_PEAK_TFLOPS, 6),
                      "traffic": traffic32,
                      "traffic_rate_GBs": (round(traffic32 / (t32_avg_ms * 1e-3) / 1e9, 1)
                                           if traffic32 else None),
                      "avg_launch_ms": round(t32_avg_ms, 4),
                      "flops_per_launch": t32_flops_launch,
                      "fp32_steps_mean": round(float(s32.mean()), 3),
                      "work": "(4/3) N^3 flops per molecule in an FP32 step, per launch"}
    T = sizes * (sizes + 1) / 2
    I = np.array([pair_block_doubles(m) for m in mols], dtype=np.float64)
    fock_ms, fock_cnt = stage["fock"]
    # B_F = s_I I + 3 s_M T per molecule (read integrals, P, H; write F); guess Fock + one per step
    fp32 = args.policy == "mixed"
    bf = 8.0 * I + 3 * 8.0 * T
    fock_bytes_step = float(np.sum((steps + 1) * bf))
    fock_avg_ms = fock_ms / max(1, fock_cnt)
    fock_gbs = fock_bytes_step * K / max(1, fock_cnt) / (fock_avg_ms * 1e-3) / 1e9
    fock_roof = {"kernel": "k_wdiag + k_fock", "bound": "hbm", "achieved": round(fock_gbs, 2),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(fock_gbs / HBM_PEAK_GBS, 5),
                 "traffic": None, "avg_launch_ms": round(fock_avg_ms, 4),
                 "work": "8 I + 24 T bytes per molecule (FP64 sizes also for FP32 steps)"}
    den_ms, den_cnt = stage["density"]
    den_flops_step = float(np.sum(steps * sizes * (sizes + 1.0) * occ))
    den_avg_ms = den_ms / max(1, den_cnt)
    den_tflops = den_flops_step * K / max(1, den_cnt) / (den_avg_ms * 1e-3) / 1e12
    den_peak = FP32_PEAK_TFLOPS if fp32 else FP64_PEAK_TFLOPS
    den_roof = {"kernel": "k_density", "bound": "mfma", "achieved": round(den_tflops, 4),
                "peak": den_peak, "unit": "TFLOP/s", "frac": round(den_tflops / den_peak, 6),
                "avg_launch_ms": round(den_avg_ms, 4),
                "work": "N(N+1) n_occ flops per molecule (P = 2 C_occ C_occ^T, lower triangle)"}
    stage_ms = {k: round(v[0] / K, 3) for k, v in stage.items()}
    stage_ms["eigh"] = round(sum(stage_ms[k] for k in api.EIGH_STAGES), 3)

    # per-molecule completion inside a batch step (iteration of completion x mean iteration)
    it_ms = step_s * 1e3 * (steps + 1) / (steps.max() + 1)
    batch_latency = {"p50_ms": round(float(np.percentile(it_ms, 50)), 2),
                     "p99_ms": round(float(np.percentile(it_ms, 99)), 2),
                     "basis": "completion iteration x mean iteration time within the batch step"}

    # batched FP64 (the reference precision) on the same batch: |dE| per molecule
    t = time.perf_counter()
    res64 = run_step(b, args.threshold, "fp64")
    fp64_s = time.perf_counter() - t
    dE = np.array([abs(r["e_tot"] - q["e_tot"]) for r, q in zip(res, res64)])
    ok = int(np.sum((dE <= 1e-4) & conv))
    diis = None
    if args.diis:
        b.set_diis(args.diis)
        run_step(b, args.threshold, args.policy)
        dt = []
        for k in range(args.steps):
            t = time.perf_counter()
            rd = run_step(b, args.threshold, args.policy)
            dt.append(time.perf_counter() - t)
            print("diis step %d: %.3f s" % (k, dt[-1]), file=sys.stderr, flush=True)
        d_s = float(np.mean(dt))
        dEd = np.array([abs(r["e_tot"] - q["e_tot"]) for r, q in zip(rd, res64)])
        okd = int(np.sum((dEd <= 1e-4) & np.array([r["converged"] for r in rd])))
        diis = {"depth": args.diis, "value": round(okd / d_s, 3), "unit": "molecules/s",
                "ms_per_step": round(d_s * 1e3, 3), "speedup_vs_value": round(step_s / d_s, 3),
                "scf_steps_mean": round(float(np.mean([r["steps"] for r in rd])), 3),
                "scf_steps_max": int(max(r["steps"] for r in rd)),
                "max_abs_dE_vs_fp64_eV": float(dEd.max()), "policy": args.policy}
    b.close()

    # plain FP64 path: one molecule at a time through the same kernels
    rng = np.random.default_rng(1)
    idx = rng.choice(len(mols), size=min(args.plain_sample, len(mols)), replace=False)
    plain_value = None
    max_dE_plain = None
    if len(idx):
        t = time.perf_counter()
        plain_res = []
        for i in idx:
            bb = eng.batch([mols[i]])
            plain_res.append(run_step(bb, args.threshold, "fp64")[0])
            bb.close()
        plain_s = time.perf_counter() - t
        plain_value = len(idx) / plain_s
        max_dE_plain = max(abs(p["e_tot"] - res64[i]["e_tot"]) for p, i in zip(plain_res, idx))

    out = {
        "metric": METRIC,
        "value": round(ok / step_s, 3),
        "unit": "molecules/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.policy == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": "%d water clusters (H2O)_n, n~U{2..64}, seed %d, SCF to %g eV"
                               % (len(mols), args.seed, args.threshold),
                   "molecules": len(mols), "basis_min": int(sizes.min()),
                   "basis_max": int(sizes.max()),
                   "precision_policy": ("fp64 throughout" if args.policy == "fp64" else
                                        "fp32 Fock, tridiagonal reduction, back-transform and "
                                        "density until |dE|<1e-3 eV, then fp64; integrals, "
                                        "bisection, inverse iteration, energies fp64"),
                   "scf_threshold_eV": args.threshold},
        "roofline": roofline,
        "tridiag_fp32_roofline": tri32_roof,
        "fock_roofline": fock_roof,
        "density_roofline": den_roof,
        "stage_ms_per_step": stage_ms,
        "converged": int(conv.sum()),
        "within_1e-4_eV_of_fp64": ok,
        "max_abs_dE_vs_fp64_eV": float(dE.max()),
        "batched_fp64": {"value": round(len(mols) / fp64_s, 3), "unit": "molecules/s",
                         "mixed_speedup": round(fp64_s / step_s, 3)},
        "scf_steps_max": int(steps.max()),
        "scf_steps_mean": round(float(steps.mean()), 3),
        "batch_latency": batch_latency,
    }
    if diis is not None:
        out["diis"] = diis
    if plain_value is not None:
        out["plain_fp64"] = {"value": round(plain_value, 3), "unit": "molecules/s",
                             "sample": len(idx), "max_abs_dE_vs_batched_eV": max_dE_plain,
                             "batched_mixed_speedup": round(value / plain_value, 3)}
    # central-difference gradients (sg_gradient_fd): one batched SCF of the
    # molecule and its 6 n_atoms displaced copies, DIIS, threshold 1e-9 eV
    grad = []
    for name, mol in (("pyridine", geometry.pyridine()),
                      ("(H2O)20", geometry.water_cluster(20, geometry.SplitMix64(20)))):
        eng.gradient_fd(mol)
        t = time.perf_counter()
        g, e, gconv = eng.gradient_fd(mol)
        grad.append({"molecule": name, "atoms": len(mol), "scf_batch": 1 + 6 * len(mol),
                     "ms": round((time.perf_counter() - t) * 1e3, 2), "converged": gconv,
                     "max_abs_sum_eV_per_A": float(np.abs(g.sum(axis=0)).max())})
    out["gradient"] = grad
    if not args.no_serve:
        t = time.time()
        out["serving"] = serve(eng, args.serve_rate, args.serve_seconds, args.threshold,
                               args.policy)
        out["latency"] = {"p50_ms": out["serving"]["p50_ms"], "p99_ms": out["serving"]["p99_ms"],
                          "basis": "serving stream, arrival to result"}
        if args.diis:
            out["serving_diis"] = serve(eng, args.serve_rate, args.serve_seconds, args.threshold,
                                        args.policy, diis=args.diis)
            out["serving_diis"]["diis_depth"] = args.diis
        print("serving: %.1f s" % (time.time() - t), file=sys.stderr, flush=True)
        extra = [x for x in args.serve_extra.split(",") if x]
        if extra:
            out["serving_rates"] = [out["serving"]]
            for x in extra:
                rate, secs = (float(v) for v in x.split(":"))
                t = time.time()
                out["serving_rates"].append(serve(eng, rate, secs, args.threshold, args.policy))
                print("serving at %g/s for %g s: %.1f s" % (rate, secs, time.time() - t),
                      file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mols, steps)
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
