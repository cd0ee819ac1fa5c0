"""Benchmark: power-flow solves/s at 1e-8 pu max mismatch on one MI355X.

Step = one batched solve of the uneven-scenario workload (2048 scenarios of
the 118-bus network + 128 scenarios of the 1,888-bus 16-copy network, per-bus
load factors clip(exp(0.15 N(0,1)), 0.7, 1.3), Q limits on, FP32 block LU +
FP64 refinement, flat start), launched as one captured hipGraph of persistent
kernels.  value = converged solves / second over the timed steps.

Printed beside it, in the same JSON line:
  * per-topology histograms of NR passes per scenario, and each topology's
    persistent kernel launched alone (the roofline's dominant kernel);
  * plain FP64 GPU NR on the same 118-bus scenarios, one scenario at a time
    (the FP64 NR loop as one captured graph per iteration, the sparse LU on
    the straightforward level schedule, lu_blocked = 0) and the
    batched/plain ratio;
  * the single 15,104-bus solve: that plain FP64 path against the optimised
    schedule (lu_blocked = 1) in FP64 and with the FP32 factor + FP64
    refinement, every variant one captured graph per iteration;
  * the learned warm start (944-bus network and the headline 118-bus
    scenarios): training on the solver's own answers, iterations and
    solves/s against the flat start, the warm answers checked against the
    flat-start oracle;
  * roofline of the dominant persistent kernel from its HIP-event time, and
    its HBM traffic from the committed rocprofv3 PMC summary when that was
    taken on this source stamp with these arguments;
  * the oracle (FP64 C restatement of the reference CPU path) timed on the
    host cores over the same workload, the accuracy of the GPU answers and
    their PV->PQ sets against it, and the single-precision (SPFP) reference
    point.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "experience-of-using-gpu-in-power-flow-computation_amd")
HBM_PEAK_GBS = 8000.0


def load_pkg():
    spec = importlib.util.spec_from_file_location("pfamd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    m = importlib.util.module_from_spec(spec)
    sys.modules["pfamd"] = m
    spec.loader.exec_module(m)
    import importlib as il
    return m, il.import_module("pfamd.synth")


def scenarios(synth, net, n, seed0):
    is_pv = np.zeros(net.nB, bool)
    is_pv[net.pv] = True
    return synth.scenario_batch(net.Pg, net.Pload, net.Qg, net.Qload, is_pv, np.arange(seed0, seed0 + n))


def bytes_per_iter(nB, nnzY_off, n, nnzL, nPV, s_lu, refine):
    """Algorithmic HBM bytes of one NR iteration of one scenario, per-scenario
    part and topology part (stage formulas: PQ+mismatch, J fill, LU numeric,
    triangular solves x (1 + refinement), refinement residuals, update,
    limits; int32 indices, s_lu bytes per factor value)."""
    nJ = 2 * n
    nnzJ = 4 * nnzY_off + 4 * n
    nnzLU = 4 * (n + 2 * nnzL)
    per = (36 * nB + 8 * nJ                     # PQ + mismatch
           + 32 * nB + s_lu * nnzJ              # J fill
           + s_lu * nnzJ + s_lu * nnzLU         # LU numeric
           + (1 + refine) * (s_lu * nnzLU + 24 * nJ)   # solves
           + refine * (12 * nnzJ + 24 * nJ)     # FP64 residuals
           + 8 * nJ + 32 * nB + 36 * nPV)       # update + limits
    topo = (4 * (nB + 1) + 20 * (nnzY_off + nB)) * 2 + 16 * (nnzY_off + nB) + \
           4 * nnzJ // 4 + (1 + refine) * 4 * nnzLU // 4
    return per, topo


def host_info():
    """The host the CPU baseline runs on: CPU model, logical CPUs, the CPUs
    this process may run on, OMP_NUM_THREADS, and the thread count used (the
    OMP_NUM_THREADS share when set, else every CPU of the affinity mask)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    used = aff
    if omp and omp.isdigit() and int(omp) > 0:
        used = min(aff, int(omp))
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": aff, "OMP_NUM_THREADS": omp,
            "threads_used": used}


def log(*a):
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print("[bench]", *a, file=sys.stderr, flush=True)


def source_stamp():
    """Hash of the library sources (kernels, host code, C ABI): names the
    build a profile was taken from, without needing git on the GPU host."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(PKG, "csrc")
    files = sorted(os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith((".hip", ".hpp", ".cpp")))
    for f in files + [os.path.join(ROOT, "include", "pf.h")]:
        h.update(os.path.basename(f).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def profile_args(args):
    """The bench arguments that change what the dominant kernel does."""
    return f"refine={args.refine} n118={args.n118} n1888={args.n1888}"


def measured_traffic(kernel, args):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE of the same workload, each scaled by
    what the counter reads for a plain device copy of known size).  The
    summary names the source stamp and bench arguments it was taken with;
    when they differ from this run's, no figure is quoted (None, reason)."""
    path = os.path.join(ROOT, "profiles", "k_batch_pmc_traffic.txt")
    stamp = pargs = None
    value = None
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if line.startswith("source_stamp:"):
                    stamp = line.split(":", 1)[1].strip()
                elif line.startswith("bench_args:"):
                    pargs = line.split(":", 1)[1].strip()
                elif len(parts) >= 4 and parts[0] == "TRAFFIC" and " ".join(parts[1:3]) == kernel:
                    value = int(parts[3])
    except OSError:
        return None, "no profile"
    src = os.path.relpath(path, ROOT)
    if stamp != source_stamp():
        return None, f"{src}: taken on source stamp {stamp}, this build is {source_stamp()}"
    if pargs != profile_args(args):
        return None, f"{src}: taken with '{pargs}', this run is '{profile_args(args)}'"
    if value is None:
        return None, f"{src}: no TRAFFIC line for {kernel}"
    return value, src


def warm_start_leg(pf, synth, net, opts, n_train, n_eval, epochs, seed_train, seed_eval):
    """Learned warm start on one topology, without Q limits (see below):
    solve n_train perturbed scenarios with the batch solver (flat start),
    train the MLP (bus P, Q and the PV setpoints -> |V|, theta) on the
    converged answers with the library's own kernels, then solve n_eval
    held-out scenarios from the flat start and from the predicted start
    (flat-start fallback inside the kernel).  A warm-started solve counts
    only if its answer matches the flat-start answer (|dV| <= 1e-6 pu,
    |dtheta| <= 1e-6 rad); the warm figure includes the inference time.

    Why without Q limits: with limits the reference's answer depends on the
    switching path, since a bus switched back to PV keeps the |V| it reached
    while limited (CPU/limInjReat.cpp:62-87).  The limits-on figure below
    starts every scenario from its own flat-start answer -- the best start
    any predictor could give -- and counts how many still end at the
    flat-start answer and limQ set."""
    is_pv = np.zeros(net.nB, bool)
    is_pv[net.pv] = True
    vset = np.where(is_pv, net.V, 0.0)
    vset[net.slack] = net.V[net.slack]
    o_free = pf.options(tolerancia=opts.tolerancia, no_max_iter=opts.no_max_iter, lim_inj_reat=0,
                        precision=opts.precision, refine_steps=opts.refine_steps)

    n_in, n_out = 3 * net.nB, 2 * net.nB
    n_in8, n_out8 = -(-n_in // 8) * 8, -(-n_out // 8) * 8   # the MLP takes multiples of 8

    def feats(P, Q):
        X = np.zeros((len(P), n_in8), np.float32)
        X[:, :n_in] = np.concatenate([P, Q, np.repeat(vset[None, :], len(P), 0)], 1)
        return X

    def solve(o, P, Q, L, V0=None, T0=None, reps=1):
        b = pf.Batch([net], o)
        b.set_scenarios(np.zeros(len(P), np.int32), P.ravel(), Q.ravel(), L.ravel(),
                        None if V0 is None else V0.ravel(), None if T0 is None else T0.ravel())
        b.solve()
        ms = min(b.solve() for _ in range(reps))
        V, th, it, cv, mm = b.results()
        lq, _ = b.limq()
        return V.reshape(len(P), -1), th.reshape(len(P), -1), it, cv, ms, np.array(lq)

    def agree(Va, Ta, Vb, Tb):
        return (np.abs(Va - Vb).max(1) <= 1e-6) & (np.abs(Ta - Tb).max(1) <= 1e-6)

    t = time.perf_counter()
    log(f"warm start {net.nB}-bus: solving {n_train} training scenarios")
    P, Q, L = synth.scenario_batch(net.Pg, net.Pload, net.Qg, net.Qload, is_pv,
                                   np.arange(seed_train, seed_train + n_train))[:3]
    V, th, it, cv, ms_data, _ = solve(o_free, P, Q, L)
    log(f"warm start {net.nB}-bus: training, {epochs} epochs")
    ok = cv.astype(bool)
    X = feats(P, Q)[ok]
    Y = np.zeros((int(ok.sum()), n_out8), np.float32)
    Y[:, :n_out] = np.concatenate([V, th], 1)[ok]
    t_data = time.perf_counter() - t
    w = pf.WarmStart(n_in8, n_out8, 512, 3, seed=11)
    w.set_norm(X.mean(0), X.std(0) + 1e-6, Y.mean(0), Y.std(0) + 1e-9)   # padding: mean 0, std ~1e-9
    t = time.perf_counter()
    losses = w.fit(X, Y, epochs=epochs, batch=256, lr=1e-3, weight_decay=1e-5)
    t_fit = time.perf_counter() - t
    log(f"warm start {net.nB}-bus: evaluating {n_eval} scenarios")
    Pt, Qt, Lt = synth.scenario_batch(net.Pg, net.Pload, net.Qg, net.Qload, is_pv,
                                      np.arange(seed_eval, seed_eval + n_eval))[:3]
    Xt = feats(Pt, Qt)
    w.predict(Xt[:256])
    t = time.perf_counter()
    Yt = w.predict(Xt)
    t_pred = time.perf_counter() - t
    Vf, thf, itf, cvf, msf, _ = solve(o_free, Pt, Qt, Lt, reps=3)
    Vw, thw, itw, cvw, msw, _ = solve(o_free, Pt, Qt, Lt, Yt[:, :net.nB].astype(np.float64),
                                      Yt[:, net.nB:n_out].astype(np.float64), reps=3)
    match = cvw.astype(bool) & cvf.astype(bool) & agree(Vw, thw, Vf, thf)
    n_match = int(match.sum())
    # device-resident: the features staged on the GPU once (scenario inputs
    # resident in HBM, as the headline's are), the inference written straight
    # into the batch's starting state; timed: inference + solve
    bd = pf.Batch([net], o_free)
    bd.set_scenarios(np.zeros(len(Pt), np.int32), Pt.ravel(), Qt.ravel(), Lt.ravel())
    w.stage(Xt)
    w.predict_into(bd, 0, net.nB)
    bd.solve()
    t = time.perf_counter()
    w.predict_into(bd, 0, net.nB)
    t_pred_dev = time.perf_counter() - t
    msw_dev = min(bd.solve() for _ in range(3))
    Vd, thd, itd, cvd, _ = bd.results()
    dev_same = bool(np.array_equal(Vd.reshape(len(Pt), -1), Vw) and np.array_equal(thd.reshape(len(Pt), -1), thw)
                    and np.array_equal(itd, itw))
    # limits on: start every scenario from its own flat-start answer
    n_lim = min(n_eval, 512)
    o_lim = pf.options(tolerancia=opts.tolerancia, no_max_iter=opts.no_max_iter, lim_inj_reat=1,
                       precision=opts.precision, refine_steps=opts.refine_steps)
    Vl, thl, itl, cvl, _, lql = solve(o_lim, Pt[:n_lim], Qt[:n_lim], Lt[:n_lim])
    Vl2, thl2, itl2, cvl2, _, lql2 = solve(o_lim, Pt[:n_lim], Qt[:n_lim], Lt[:n_lim], Vl, thl)
    lim_match = cvl.astype(bool) & cvl2.astype(bool) & agree(Vl2, thl2, Vl, thl) & \
        np.array([np.array_equal(a, b) for a, b in zip(lql, lql2)], bool)
    res = {"buses": net.nB, "q_limits": False, "train_scenarios": int(n_train), "train_converged": int(cv.sum()),
           "eval_scenarios": int(n_eval), "epochs": epochs,
           "mlp": f"{n_in8}->512x3 GELU->{n_out8}, bf16 weights, fp32 accumulate",
           "train_data_s": round(t_data, 2), "fit_s": round(t_fit, 2),
           "loss_first_last": [round(float(losses[0]), 5), round(float(losses[-1]), 6)],
           "iterations_flat_mean": round(float(itf.mean()), 3),
           "iterations_warm_mean": round(float(itw.mean()), 3),
           "converged_flat": int(cvf.sum()), "converged_warm": int(cvw.sum()),
           "warm_matching_flat": n_match,
           "solve_ms_flat": round(msf, 3), "solve_ms_warm": round(msw, 3),
           "predict_ms": round(1e3 * t_pred, 3),
           "solves_per_s_flat": round(float(cvf.sum()) / (msf * 1e-3), 1),
           "solves_per_s_warm_incl_predict": round(n_match / ((msw + 1e3 * t_pred) * 1e-3), 1),
           "warm_vs_flat_incl_predict": round((n_match / (msw + 1e3 * t_pred)) / (float(cvf.sum()) / msf), 3),
           "device_resident": {
               "predict_into_ms": round(1e3 * t_pred_dev, 3), "solve_ms_warm": round(msw_dev, 3),
               "same_answers_as_host_path": dev_same,
               "solves_per_s_warm_incl_predict": round(n_match / ((msw_dev + 1e3 * t_pred_dev) * 1e-3), 1),
               "warm_vs_flat_incl_predict": round((n_match / (msw_dev + 1e3 * t_pred_dev)) /
                                                  (float(cvf.sum()) / msf), 3)},
           "max_abs_dV_warm_vs_flat_pu": float(np.abs(Vw[match] - Vf[match]).max()) if match.any() else None,
           "limits_on_started_at_own_answer": {
               "scenarios": n_lim, "converged": int(cvl.sum()), "matching_flat_answer_and_limQ": int(lim_match.sum()),
               "iterations_flat_mean": round(float(itl.mean()), 3),
               "iterations_from_own_answer_mean": round(float(itl2.mean()), 3)}}
    # inputs and answers for the oracle check in the host leg
    check = {"P": Pt, "Q": Qt, "L": Lt, "Vw": Vw, "thw": thw, "cvw": cvw}
    return res, check


def single_leg(pf, synth, tmp, variants, refine):
    """The single 15,104-bus solve (flat start, Q limits, tol 1e-8), NR loop
    time median of 3 after a warm-up solve, for each requested variant:
    plain_fp64_graph (the plain FP64 GPU NR: device-resident loop, one
    captured graph per iteration, the level schedule: one launch per wide
    elimination-tree level, the narrow ones in one workgroup),
    plain_fp64_eager (the same with a host sync per iteration), fp64_eager /
    fp64_graph (subtree groups + dense separator block), mixed_eager /
    mixed_graph (the same with the FP32 factor and FP64 refinement)."""
    kinds = {"plain_fp64_graph": dict(lu_blocked=0, use_graph=1), "plain_fp64_eager": dict(lu_blocked=0),
             "fp64_eager": dict(), "fp64_graph": dict(use_graph=1),
             "mixed_eager": dict(precision=pf.MIXED, refine_steps=refine),
             "mixed_graph": dict(use_graph=1, precision=pf.MIXED, refine_steps=refine)}
    p15 = os.path.join(tmp, "c15104.cdfx")
    synth.write_cdf(synth.case15104(), p15)
    n15 = pf.Network(p15)
    res = {}
    s15 = None
    for tag in variants:
        s15 = pf.Solver(n15, pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1, **kinds[tag]))
        s15.solve()
        times = []
        for _ in range(3):
            s15.reset()
            r = s15.solve()
            times.append(r["solve_ms"])
        res[tag] = {"ms": round(float(np.median(times)), 3), "iterations": r["iterations"],
                    "converged": r["converged"]}
        log(f"single 15,104-bus {tag}: {res[tag]['ms']} ms")
    if "plain_fp64_graph" in res and "mixed_graph" in res:
        res["speedup_mixed_graph_vs_plain"] = round(res["plain_fp64_graph"]["ms"] / res["mixed_graph"]["ms"], 2)
        res["plain"] = "plain_fp64_graph"
    if "fp64_graph" in res and "mixed_graph" in res:
        res["speedup_mixed_graph_vs_fp64_graph"] = round(res["fp64_graph"]["ms"] / res["mixed_graph"]["ms"], 2)
    if s15 is not None:
        res["lu"] = s15.lu_info()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n118", type=int, default=2048)
    ap.add_argument("--n1888", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-single", action="store_true", help="skip the single 15,104-bus solve")
    ap.add_argument("--single-only", action="store_true", help="only the single 15,104-bus solve")
    ap.add_argument("--single-variants", default="plain_fp64_graph,plain_fp64_eager,fp64_graph,mixed_graph")
    ap.add_argument("--refine", type=int, default=1, help="FP64 refinement steps after each FP32 solve")
    ap.add_argument("--no-warm", action="store_true", help="skip the learned warm-start legs")
    ap.add_argument("--warm-train", type=int, default=50000)
    ap.add_argument("--warm-eval", type=int, default=4096)
    ap.add_argument("--warm-epochs", type=int, default=8)
    ap.add_argument("--tune-cache", default=os.path.expanduser("~/.cache/libpf/batch_shapes.txt"),
                    help="launch-shape cache of the batch autotuner ('' = measure, keep nothing)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 64 << 20
        a = torch.ones(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                          "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU scaling is not provided",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    pf, synth = load_pkg()
    tmp = tempfile.mkdtemp(prefix="pfbench_")
    if args.single_only:
        print(json.dumps({"single_15104": single_leg(pf, synth, tmp, args.single_variants.split(","), args.refine)}))
        return
    p118 = os.path.join(tmp, "c118.cdf")
    p1888 = os.path.join(tmp, "c1888.cdf")
    synth.write_cdf(synth.case118(), p118)
    synth.write_cdf(synth.case1888(), p1888)
    n118, n1888 = pf.Network(p118), pf.Network(p1888)
    A = scenarios(synth, n118, args.n118, 3000000)
    B = scenarios(synth, n1888, args.n1888, 3000000 + args.n118)
    nsc = args.n118 + args.n1888
    net_of = np.array([1] * args.n1888 + [0] * args.n118, np.int32)   # big ones first in the queue
    P = np.concatenate([B[0].ravel(), A[0].ravel()])
    Q = np.concatenate([B[1].ravel(), A[1].ravel()])
    L = np.concatenate([B[2].ravel(), A[2].ravel()])
    opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1, precision=pf.MIXED,
                      refine_steps=args.refine)
    cache = args.tune_cache or None
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
    batch = pf.Batch([n118, n1888], opts)
    batch.set_scenarios(net_of, P, Q, L)
    t = time.perf_counter()
    n_tuned = batch.autotune(cache, reps=1)
    tune_s = time.perf_counter() - t
    for _ in range(args.warmup):
        batch.solve()
    t0 = time.perf_counter()
    ev_ms = []
    for _ in range(args.steps):
        ev_ms.append(batch.solve())
    t1 = time.perf_counter()
    V, TH, IT, CV, MM = batch.results()
    LQ, FL = batch.limq()
    wall = t1 - t0
    n_conv = int(CV.sum())
    value = n_conv * args.steps / wall
    ms_step = 1e3 * wall / args.steps

    extras = {}
    # per-topology histogram of NR passes per scenario: a launch lasts as long
    # as its slowest scenario (the non-converging ones run to no_max_iter)
    hist = {}
    for name, its in (("1888-bus", IT[:args.n1888]), ("118-bus", IT[args.n1888:])):
        vals, cnt = np.unique(its, return_counts=True)
        hist[name] = {str(int(v)): int(c) for v, c in zip(vals, cnt)}
    extras["iteration_histogram"] = hist
    spfp_gpu = None
    # roofline of each persistent kernel from its own timed launch
    roof = None
    log("headline batch done", f"{value:.1f} solves/s")
    if not args.no_extras:
        log("per-kernel launches and plain FP64 GPU NR")
        kern = {}
        for name, net, S, k in (("k_batch 118-bus", n118, A, args.n118),
                                ("k_batch 1888-bus", n1888, B, args.n1888)):
            b1 = pf.Batch([net], opts)
            b1.set_scenarios(np.zeros(k, np.int32), S[0].ravel(), S[1].ravel(), S[2].ravel())
            b1.autotune(cache, reps=1)
            b1.solve()
            ts = [b1.solve() for _ in range(max(3, args.steps // 2))]
            _, _, it1, cv1, _ = b1.results()
            pr = pf.SymbolicProbe(net)
            rp = pr.rowptr
            coorow = np.repeat(np.arange(net.nB), np.diff(rp))
            off = int(np.sum((coorow != net.slack) & (pr.colind != net.slack) & (coorow != pr.colind)))
            per, topo = bytes_per_iter(net.nB, off, pr.n, pr.nnzL, net.nPV, 4, args.refine)
            alg = per * int(it1.sum()) + topo * int(it1.max())
            kern[name] = dict(ms=float(np.mean(ts)), bytes=alg, iters=int(it1.sum()),
                              conv=int(cv1.sum()), n=k, shape=b1.shape(0))
        dom = max(kern, key=lambda k: kern[k]["ms"])
        d = kern[dom]
        ach = d["bytes"] / (d["ms"] * 1e-3) / 1e9
        traffic, tsrc = measured_traffic(dom, args)
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": tsrc, "launch_ms": round(d["ms"], 4), "alg_bytes_per_launch": d["bytes"],
                "shape": d["shape"]}
        extras["kernels"] = {k: {"launch_ms": round(v["ms"], 4), "alg_bytes": v["bytes"],
                                 "scenario_iterations": v["iters"], "converged": v["conv"],
                                 "scenarios": v["n"], "shape": v["shape"]} for k, v in kern.items()}

        # plain FP64 GPU NR, one 118-bus scenario at a time: the device-resident
        # FP64 loop (one captured graph per iteration) with the level schedule
        sv = pf.Solver(n118, pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1, lu_blocked=0,
                                        use_graph=1))
        m = min(64, args.n118)
        sv.set_injections(A[0][0], A[1][0], A[2][0])
        sv.reset()
        sv.solve()
        t = time.perf_counter()
        pconv = 0
        for s in range(m):
            sv.set_injections(A[0][s], A[1][s], A[2][s])
            sv.reset()
            pconv += sv.solve()["converged"]
        plain = pconv / (time.perf_counter() - t)
        b118 = kern["k_batch 118-bus"]
        batched118 = b118["conv"] / (b118["ms"] * 1e-3)
        extras["plain_fp64_gpu_118"] = {"solves_per_s": round(plain, 1), "sample": m,
                                        "path": "FP64, flat start, one scenario per launch sequence, "
                                                "one captured graph per NR iteration, level-scheduled LU"}
        extras["batched_mixed_gpu_118"] = {"solves_per_s": round(batched118, 1)}
        extras["batched_vs_plain_118"] = round(batched118 / plain, 1)
        # the mixed-precision answer of the 118-bus case without Q limits, for
        # the single-precision (SPFP) reference point computed on the host
        sm = pf.Solver(n118, pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=0, precision=pf.MIXED,
                                        refine_steps=args.refine, use_graph=1))
        rm = sm.solve()
        spfp_gpu = (sm.state()[:2], rm)

    if not args.no_extras and not args.no_single:
        log("single 15,104-bus solve")
        extras["single_15104"] = single_leg(pf, synth, tmp, args.single_variants.split(","), args.refine)

    warm_checks = {}
    if not args.no_extras and not args.no_warm:
        # learned warm start: the 944-bus 8-copy network (50k training, 4096
        # held-out scenarios) and the 118-bus network of the headline batch
        p944 = os.path.join(tmp, "c944.cdf")
        synth.write_cdf(synth.case944(), p944)
        n944 = pf.Network(p944)
        extras["warm_start_944"], warm_checks[p944] = warm_start_leg(
            pf, synth, n944, opts, args.warm_train, args.warm_eval, args.warm_epochs, 4000000, 5000000)
        extras["warm_start_118"], warm_checks[p118] = warm_start_leg(
            pf, synth, n118, opts, args.warm_train, args.n118, args.warm_epochs, 6000000, 3000000)

    cpu = None
    if not args.no_cpu:
        log("oracle on the host cores")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        O.build()
        host = host_info()
        threads = host["threads_used"]
        o118, o1888 = O.Case(p118), O.Case(p1888)
        t = time.perf_counter()
        Vb, Tb, Ib, Cb, Lb = o1888.solve_batch(B[0], B[1], B[2], tol=1e-8, max_iter=60, lim=True,
                                               nthreads=threads, want_limq=True)
        Va, Ta, Ia, Ca, La = o118.solve_batch(A[0], A[1], A[2], tol=1e-8, max_iter=60, lim=True,
                                              nthreads=threads, want_limq=True)
        tc = time.perf_counter() - t
        cpu = {"value": round((int(Ca.sum()) + int(Cb.sum())) / tc, 2), "unit": "solves/s", "cores": threads,
               "kind": "port",
               "sample": f"the full workload once ({args.n118} x 118-bus + {args.n1888} x 1888-bus), "
                         f"scenario-parallel over {threads} threads, ordering + symbolic redone every "
                         f"iteration as PARDISO phases 11/22/33; {tc:.1f} s wall",
               "host": host}
        # the same port on an eighth of the workload: one thread, and the
        # reference's own shape (scenarios in sequence, OpenMP inside a solve)
        k118, k1888 = max(1, args.n118 // 8), max(1, args.n1888 // 8)
        for tag, nt, inner in (("one_thread", 1, False), ("reference_style", threads, True)):
            t = time.perf_counter()
            c1 = o1888.solve_batch(B[0][:k1888], B[1][:k1888], B[2][:k1888], tol=1e-8, max_iter=60, lim=True,
                                   nthreads=nt, inner=inner)[3]
            c2 = o118.solve_batch(A[0][:k118], A[1][:k118], A[2][:k118], tol=1e-8, max_iter=60, lim=True,
                                  nthreads=nt, inner=inner)[3]
            te = time.perf_counter() - t
            cpu[tag] = {"value": round((int(c1.sum()) + int(c2.sum())) / te, 2), "unit": "solves/s",
                        "threads": nt, "sample": f"{k118} x 118-bus + {k1888} x 1888-bus, {te:.1f} s wall"}
            log(f"oracle {tag}: {cpu[tag]['value']} solves/s")
        # accuracy of the batched GPU answers against the oracle, over the
        # scenarios both converged, and the final PV->PQ sets of every scenario
        nb = n1888.nB * args.n1888
        Vg = [V[:nb].reshape(args.n1888, -1), V[nb:].reshape(args.n118, -1)]
        Tg = [TH[:nb].reshape(args.n1888, -1), TH[nb:].reshape(args.n118, -1)]
        cg = [CV[:args.n1888], CV[args.n1888:]]
        lg = [LQ[:args.n1888], LQ[args.n1888:]]
        dV = dT = 0.0
        both = 0
        same_lq = 0
        for vg, tg, c_g, l_g, vo, to, c_o, l_o in ((Vg[0], Tg[0], cg[0], lg[0], Vb, Tb, Cb, Lb),
                                                   (Vg[1], Tg[1], cg[1], lg[1], Va, Ta, Ca, La)):
            m = c_g & c_o
            both += int(m.sum())
            same_lq += sum(int(np.array_equal(a, b)) for a, b in zip(l_g, l_o))
            if m.any():
                dV = max(dV, float(np.abs(vg[m] - vo[m]).max()))
                dT = max(dT, float(np.abs(tg[m] - to[m]).max()))
        extras["accuracy_vs_oracle"] = {"max_abs_dV_pu": dV, "max_abs_dtheta_rad": dT,
                                        "scenarios_compared": both,
                                        "gpu_converged": int(CV.sum()),
                                        "oracle_converged": int(Ca.sum() + Cb.sum()),
                                        "same_converged_set": bool(np.array_equal(np.concatenate(cg),
                                                                                  np.concatenate([Cb, Ca]))),
                                        "same_limQ_set": same_lq, "scenarios": nsc}
        # the reference's whole-program single-precision build (oracle/spfp.py)
        # as an accuracy reference point for the mixed-precision design
        if spfp_gpu is not None:
            import spfp
            o = o118.solve(tol=1e-8, max_iter=60, lim=False)
            V32, T32, it32, ok32 = spfp.solve(o118, np.float32, tol=1e-4, max_iter=60)
            (Vm, Tm), rm = spfp_gpu
            extras["spfp_reference_point"] = {
                "case": "118-bus, Q limits off, flat start",
                "spfp": {"tol": 1e-4, "converged": bool(ok32), "iterations": int(it32),
                         "max_abs_dV_vs_fp64_pu": float(np.abs(V32 - o["V"]).max()),
                         "max_abs_dtheta_vs_fp64_rad": float(np.abs(T32 - o["theta"]).max())},
                "mixed": {"tol": 1e-8, "converged": bool(rm["converged"]), "iterations": int(rm["iterations"]),
                          "max_abs_dV_vs_fp64_pu": float(np.abs(Vm - o["V"]).max()),
                          "max_abs_dtheta_vs_fp64_rad": float(np.abs(Tm - o["theta"]).max())}}
        # warm-started answers (Q limits off) against the flat-start oracle
        for path, chk in warm_checks.items():
            oc = O.Case(path)
            k = min(64, len(chk["P"]))
            Vo, To, _, Co = oc.solve_batch(chk["P"][:k], chk["Q"][:k], chk["L"][:k], tol=1e-8, max_iter=60,
                                           lim=False, nthreads=threads)
            m = Co & chk["cvw"][:k].astype(bool)
            key = "warm_start_944" if oc.nB == 944 else "warm_start_118"
            extras[key]["oracle_check"] = {
                "scenarios": k, "both_converged": int(m.sum()),
                "max_abs_dV_warm_vs_flat_oracle_pu": float(np.abs(chk["Vw"][:k][m] - Vo[m]).max()) if m.any() else None,
                "max_abs_dtheta_warm_vs_flat_oracle_rad":
                    float(np.abs(chk["thw"][:k][m] - To[m]).max()) if m.any() else None}

    out = {"metric": "power-flow solves/s at 1e-8 pu mismatch, 15k-bus + batched 118-bus vs plain NR",
           "value": round(value, 2), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64 mismatch/update, f32 LU + f64 refinement",
           "data": "synthetic (seeded CDF cases and lognormal load scenarios)",
           "config": {"workload": f"{args.n118} x 118-bus + {args.n1888} x 1888-bus scenarios in one "
                                  f"persistent launch, Q limits on, tol 1e-8, flat start",
                      "scenarios": nsc, "converged_per_step": n_conv,
                      "launch_shapes": {"118-bus": batch.shape(0), "1888-bus": batch.shape(1),
                                        "autotune_candidates_measured": n_tuned,
                                        "autotune_s": round(tune_s, 2)},
                      "event_ms_per_step": round(float(np.mean(ev_ms)), 4)},
           "roofline": roof, "cpu_baseline": cpu}
    out.update(extras)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
