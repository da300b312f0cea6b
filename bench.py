"""Benchmark: converged eigenpairs/s (FP64 true residual <= 1e-8) of the
three-fluxonia flat grid (three PhaseGrid(-6pi, 6pi, 512; order=2) modes,
N = 512^3 = 134 217 728 real states, Lanczos lowest 10) on one MI355X, plus the
fused apply's HBM bandwidth against peak.

A step is one full Lanczos solve (operator construction, the solve, the FP64
acceptance residuals).  Prints ONE JSON line (rank 0)."""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "kronfreecircuits.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


# the fused stencil's kernels by launch-shape variant (kfc_op_stencil_get)
STENCIL_KERNELS = {0: "k_stencil", 1: "k_stencil_q", 2: "k_stencil_p", 3: "k_stencil_a", 4: "k_stencil_pp",
                   5: "k_stencil_pm"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def apply_bandwidth(H, prec, reps=20):
    """Average fused-apply duration from HIP events on the library's stream."""
    import torch
    dt = torch.float64 if prec == "f64" else torch.float32
    X = torch.randn(H.N, dtype=dt, device="cuda:0")
    Y = torch.empty_like(X)
    torch.cuda.synchronize()
    s = torch.cuda.ExternalStream(H.ctx.stream)
    for _ in range(3):
        H.mul(Y, X, prec)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        H.mul(Y, X, prec)
    e1.record(s)
    e1.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    nbytes = 2 * H.N * (8 if prec == "f64" else 4)
    del X, Y
    return t, nbytes / t / 1e9


def cpu_baseline(theta, N, n_matvec_gpu, m):
    """The reference's apply algorithm (the per-term two-buffer chain,
    src/apply.jl:370-432) restated in C with OpenMP (oracle/apply_chain.c),
    timed on this host's cores on a bounded sample of full-size applies and
    converted to eigenpairs/s with the GPU solve's matvec count (the Lanczos
    orthogonalization work is left out, which favours the CPU)."""
    import numpy as np

    from oracle import c_apply
    from oracle import kfc_oracle as O
    Ho = O.three_fluxonia_operator(theta["E_C"], theta["E_L"], theta["E_J"], theta["phi_ext"], N=N, order=2)
    x = np.random.default_rng(0).standard_normal(Ho.N)
    if c_apply.available():
        f = c_apply.CApply(Ho)
        kind, threads = "C restatement (oracle/apply_chain.c, OpenMP)", int(os.environ.get("OMP_NUM_THREADS", os.cpu_count()))
    else:
        f = lambda v: O.apply_chain(Ho, v)
        kind, threads = "numpy oracle", 1
    y = f(x)
    t0 = time.perf_counter()
    napp = 0
    while napp < 20 and time.perf_counter() - t0 < 15:
        f(x)
        napp += 1
    t = (time.perf_counter() - t0) / napp
    del y
    return {"value": round(m / (n_matvec_gpu * t), 6), "unit": "eigenpairs/s", "cores": threads, "kind": "port",
            "sample": f"{napp} applies of the {kind} at N={N}^3 ({t:.3f} s each), scaled by the GPU solve's "
                      f"{n_matvec_gpu} matvecs (orthogonalization not counted)",
            "seconds_per_apply": round(t, 4)}


def measured_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE runs, each scaled by what the counter reads for a copy of known size),
    with the file and the commit it was measured at: PMC counters cannot run inside
    the timed solve, so this figure is not from this run."""
    tf = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(tf):
        return None, None
    with open(tf) as f:
        d = json.load(f)
    return d.get("bytes_per_launch"), f"profiles/{name} (measured at commit {d.get('commit', 'unrecorded')})"


def calibrate_copy(gib=1, reps=5):
    """y = x + 0 over `gib` GiB of FP64, `reps` times: each launch reads and writes
    exactly gib GiB, the yardstick that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = gib * (1 << 30) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.add(x, 0.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "elementwise copy", "bytes_read_per_launch": n * 8,
                      "bytes_written_per_launch": n * 8, "launches": reps}))


def summarize_pmc(args):
    """HBM bytes per launch of one kernel from two rocprofv3 --pmc runs of the same
    command (FETCH_SIZE, then WRITE_SIZE; counter_collection CSVs, KiB per dispatch),
    each scaled by what the counter reads for the elementwise copy of known size
    (the same two counters over `bench.py --calibrate-copy`).  Written to a profiles
    JSON stamped with the commit of this tree (the code that was measured)."""
    import csv
    import subprocess

    def kib(path, counter, want, stat="mean"):
        vals = [float(r["Counter_Value"]) for r in csv.DictReader(open(path))
                if r["Counter_Name"] == counter and want(r["Kernel_Name"])]
        if stat == "median":   # the typical launch: wider launches of the same kernel left out
            vals.sort()
            return vals[len(vals) // 2] * len(vals), len(vals)
        return sum(vals), len(vals)

    is_copy = lambda k: "add" in k.lower() and "fill" not in k.lower()   # torch.add(x, 0, out=y)
    is_kernel = lambda k: args.pmc_kernel in k
    cf, ncf = kib(args.pmc_copy_fetch, "FETCH_SIZE", is_copy)
    cw, ncw = kib(args.pmc_copy_write, "WRITE_SIZE", is_copy)
    copy_bytes = (1 << 30)
    # the copy reads and writes exactly 1 GiB per launch: bytes per KiB of counter
    fscale = copy_bytes / (cf / ncf * 1024.0)
    wscale = copy_bytes / (cw / ncw * 1024.0)
    f, nf = kib(args.pmc_fetch, "FETCH_SIZE", is_kernel, args.pmc_stat)
    w, nw = kib(args.pmc_write, "WRITE_SIZE", is_kernel, args.pmc_stat)
    fetch = f / nf * 1024.0 * fscale
    write = w / nw * 1024.0 * wscale
    try:
        commit = subprocess.run(["git", "-C", ROOT, "rev-parse", "--short", "HEAD"], capture_output=True,
                                text=True).stdout.strip() or "unknown"
    except OSError:
        commit = "unknown"
    out = {"kernel": args.pmc_kernel, "workload": args.pmc_workload, "command": args.pmc_command,
           "bytes_per_launch": int(round(fetch + write)), "fetch_bytes": int(round(fetch)),
           "write_bytes": int(round(write)), "launches_counted": nf, "fetch_scale": round(fscale, 4),
           "write_scale": round(wscale, 4), "statistic": args.pmc_stat + " launch", "commit": commit}
    with open(args.pmc_out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


def pairs_bench(args):
    """The dense kept-basis apply of the zero-pi parent [40, 40, 30] on 64 columns,
    FP64 and FP32 (the tuned launch shape per axis-pair group), timed with HIP events
    on the library's stream: the MFMA roofline of the lockstep sweep's kernels."""
    import torch

    from kfc import configs
    from kfc.sweeps import ZeroPiSweep
    sw = ZeroPiSweep(contexts=1)
    H, _ = sw.parent(configs.zero_pi_theta(50_000))
    _, flops_col = H.work()
    ncol, reps = 64, max(1, args.steps) * 20
    st = H.ctx.torch_stream()
    res = {}
    for prec, dt, peak in (("f64", torch.float64, 78.6), ("f32", torch.float32, 157.3)):
        X = torch.randn(ncol, H.N, dtype=dt, device="cuda:0")
        Y = torch.empty_like(X)
        for _ in range(max(1, args.warmup) * 3):
            H.mul(Y, X, prec)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            H.mul(Y, X, prec)
        e1.record(st)
        e1.synchronize()
        t = e0.elapsed_time(e1) / reps / 1e3
        rate = flops_col * ncol / t / 1e12
        res[prec] = {"us_per_apply": round(t * 1e6, 1), "tflops": round(rate, 2), "peak": peak,
                     "frac": round(rate / peak, 4)}
        log(f"[pairs] {prec}: {t * 1e6:.1f} us per apply of {ncol} columns, {rate:.2f} TFLOP/s")
    out = {"metric": "dense kept-basis apply TFLOP/s (zero-pi parent [40,40,30], 64 columns)",
           "value": res["f64"]["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
           "ms_per_step": round(res["f64"]["us_per_apply"] / 1e3, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded circuit parameters, random block)",
           "config": {"workload": "zero_pi_parent_pairs", "dims": [40, 40, 30], "columns": ncol,
                      "flops_per_column": flops_col},
           "roofline": {"bound": "mfma", "achieved": res["f64"]["tflops"], "peak": 78.6, "unit": "TFLOP/s",
                        "frac": res["f64"]["frac"], "traffic": None},
           "f32": res["f32"]}
    print(json.dumps(out))


def zero_pi_runs(args):
    """The zero-pi sweep (parent dims [40, 40, 30], LOBPCG lowest 16, tol 1e-8) at
    args.test_points seeded points, each path timed end to end (operators, seeds,
    solve, FP64 acceptance residuals), counting pairs whose FP64 true residual is
    <= 1e-8:

    * plain_fp64_cold: point after point, product seed (the reference's materialize);
    * plain_fp64_warm: point after point, each warm-started from the previous point's
      vectors (the reference's sweep convention), the points walked in
      nearest-neighbour order;
    * lockstep_fp64: every point's LOBPCG step in one launch per kernel;
    * lockstep_mixed_learned: lockstep, FP32 bulk and FP64 refinement, started from
      the warm-start network's prediction (trained here on args.train_points
      points' plain-solver eigenpairs);
    * lockstep_fp64_subspace / lockstep_mixed_subspace: lockstep, every point
      started from the lowest Ritz vectors of its operator on the learned subspace
      (the leading args.subspace_basis principal directions of the same training
      eigenvectors), the Rayleigh-Ritz step on the device, timed with the solve.

    Returns the record and the lockstep run's outputs."""
    import numpy as np

    from kfc import configs
    from kfc.sweeps import LearnedSubspace, WarmStart, ZeroPiSweep, path_order, subspace_capture

    sw = ZeroPiSweep(contexts=args.contexts)   # lockstep runs on the first; the training solves use them all
    m = sw.m
    test = [configs.zero_pi_theta(50_000 + i) for i in range(args.test_points)]
    train = [configs.zero_pi_theta(i) for i in range(args.train_points)]

    def counted(out):
        return sum(m for o in out if o[2]["max_residual"] <= 1e-8)

    def summary(out, wall):
        return {"seconds": round(wall, 3), "eigenpairs_per_s": round(counted(out) / wall, 2),
                "mean_iterations": round(float(np.mean([o[2]["iterations"] for o in out])), 1),
                "matvec_columns": int(sum(o[2]["n_matvec"] for o in out)),
                "max_residual": max(o[2]["max_residual"] for o in out)}

    # warm-up: every path once at the sizes it runs at (compiles, tunes launch shapes)
    sw.run_plain(test[:1])
    sw.run_batched(test, mode="f64", lockstep=True)
    sw.run_batched(test, mode="mixed", lockstep=True)
    res = {}
    cold, t = sw.run_plain(test)
    res["plain_fp64_cold"] = summary(cold, t)
    log(f"[zero-pi] plain FP64 cold: {t:.2f} s, {res['plain_fp64_cold']['eigenpairs_per_s']} pairs/s")
    order = path_order(test)
    warm, t = sw.run_plain_warm([test[i] for i in order])
    res["plain_fp64_warm"] = summary(warm, t)
    log(f"[zero-pi] plain FP64 warm from the previous point: {t:.2f} s, "
        f"{res['plain_fp64_warm']['eigenpairs_per_s']} pairs/s")
    lock, t = sw.run_batched(test, mode="f64", lockstep=True)
    res["lockstep_fp64"] = summary(lock, t)
    log(f"[zero-pi] lockstep FP64: {t:.2f} s, {res['lockstep_fp64']['eigenpairs_per_s']} pairs/s")
    # the same sweep with each step's kernels launched one by one instead of replayed
    # from its graph (the applies are timed with HIP events only here)
    os.environ["KFC_GRAPHS"] = "0"
    try:
        sw.run_batched(test[:4], mode="f64", lockstep=True)
        lock_l, t = sw.run_batched(test, mode="f64", lockstep=True)
    finally:
        del os.environ["KFC_GRAPHS"]
    res["lockstep_fp64_launches"] = summary(lock_l, t)
    res["lockstep_fp64_launches"]["seconds_apply_gpu"] = round(float(sum(o[2]["seconds_apply_gpu"] for o in lock_l)), 4)
    log(f"[zero-pi] lockstep FP64, launches instead of step graphs: {t:.2f} s, "
        f"{res['lockstep_fp64_launches']['eigenpairs_per_s']} pairs/s")
    # the warm-start network, trained on the plain solver's lowest m + g pairs (the
    # lockstep solver with one guard column, so that m + g pairs fit its block)
    nev = sw.width(40 * 40 * 30)
    t0 = time.perf_counter()
    gen, _ = sw.run_batched(train, mode="f64", nev=nev, lockstep=True, guard=1)
    t_gen = time.perf_counter() - t0
    frames = [sw.to_frame(V, sw.frame_index(kids)) for E, V, st, kids in gen]
    ws = WarmStart(sw, nbasis=args.nbasis, seed=1)
    loss = ws.train(train, frames, steps=args.train_steps, lr=1e-3)
    t_train = time.perf_counter() - t0
    log(f"[zero-pi] warm start: {len(train) * nev} training pairs ({t_gen:.1f} s), loss {loss[0]:.4g} -> {loss[1]:.4g}, "
        f"{t_train:.1f} s")
    t0 = time.perf_counter()
    sub = LearnedSubspace(sw, nbasis=args.subspace_basis).fit(frames)
    t_sub = time.perf_counter() - t0
    sw.run_batched(test[:2], mode="f64", subspace=sub)    # warm-up at the subspace width
    out, t = sw.run_batched(test, mode="f64", subspace=sub)
    res["lockstep_fp64_subspace"] = summary(out, t)
    msub, t = sw.run_batched(test, mode="mixed", subspace=sub)
    res["lockstep_mixed_subspace"] = summary(msub, t)
    # the two effects apart: FP64 from the prediction, mixed from the product seed
    out, t = sw.run_batched(test, mode="f64", x0_fn=ws.x0_fn, lockstep=True)
    res["lockstep_fp64_learned"] = summary(out, t)
    out, t = sw.run_batched(test, mode="mixed", lockstep=True)
    res["lockstep_mixed"] = summary(out, t)
    learned, t = sw.run_batched(test, mode="mixed", x0_fn=ws.x0_fn, lockstep=True)
    res["lockstep_mixed_learned"] = summary(learned, t)
    for k in ("lockstep_fp64_learned", "lockstep_mixed", "lockstep_mixed_learned", "lockstep_fp64_subspace",
              "lockstep_mixed_subspace"):
        log(f"[zero-pi] {k}: {res[k]['seconds']:.2f} s, {res[k]['eigenpairs_per_s']} pairs/s, "
            f"mean iterations {res[k]['mean_iterations']}")
    base = res["plain_fp64_cold"]["eigenpairs_per_s"]
    # the mixed paths' error against the plain FP64 solve, per eigenvalue (max over points)
    rel = np.array([np.abs(a[0] - b[0]) / np.abs(b[0]) for a, b in zip(msub, cold)])
    rel_net = np.array([np.abs(a[0] - b[0]) / np.abs(b[0]) for a, b in zip(learned, cold)])
    caps = [float(np.mean(subspace_capture(sw.from_frame(xf, sw.frame_index(kids), V.shape[0]), V[:, :m])))
            for xf, (E, V, st, kids) in zip(ws.predict_frames(test[:8]), lock[:8])]
    H0, _ = sw.parent(test[0])
    _, flops_col = H0.work()
    bsec, bcols = res["lockstep_fp64_launches"]["seconds_apply_gpu"], res["lockstep_fp64_launches"]["matvec_columns"]
    rate = flops_col * bcols / bsec / 1e12 if bsec > 0 else None
    rec = {"workload": "zero_pi_sweep", "dims": [40, 40, 30], "m": m, "tol": 1e-8, "points": len(test),
           **res,
           "speedup_vs_plain_fp64_cold": {k: round(v["eigenpairs_per_s"] / base, 2) for k, v in res.items()},
           "best": max(res, key=lambda k: res[k]["eigenpairs_per_s"]),
           "target": "lockstep_mixed_subspace (batched, mixed precision, learned start) >= 5x plain_fp64_cold",
           "target_met": bool(res["lockstep_mixed_subspace"]["eigenpairs_per_s"] >= 5 * base),
           "mixed_rel_dE_per_eigenvalue_vs_fp64": [float(f"{x:.3e}") for x in rel.max(axis=0)],
           "mixed_max_rel_dE_vs_fp64": float(rel.max()),
           "mixed_rel_dE_path": "lockstep_mixed_subspace against plain_fp64_cold (max over points per eigenvalue)",
           "mixed_learned_max_rel_dE_vs_fp64": float(rel_net.max()),
           "warm_start": {"training_points": len(train), "training_eigenpairs": len(train) * nev,
                          "outputs": f"{ws.nb} x {ws.m}", "initial_loss": float(loss[0]), "final_loss": float(loss[1]),
                          "seconds_data_and_training": round(t_train, 2),
                          "mean_subspace_capture": round(float(np.mean(caps)), 4)},
           "learned_subspace": {"basis": args.subspace_basis, "frame_rows": int(sub.B.shape[0]),
                                "training_eigenpairs": len(train) * nev, "seconds_fit": round(t_sub, 2),
                                "singular_value_ratio_last": float(sub.singular_values[-1] / sub.singular_values[0])},
           "roofline_dense_pairs": {"bound": "mfma",
                                    "kernel": "dense kept-basis pairs (lockstep FP64 applies, step launches)",
                                    "achieved": round(rate, 3) if rate else None, "peak": 78.6, "unit": "TFLOP/s",
                                    "frac": round(rate / 78.6, 4) if rate else None,
                                    "flops_per_column": flops_col}}
    return rec, lock


def sweep_bench(args):
    """The zero-pi sweep on its own line: the value is lockstep mixed precision
    started from the learned subspace."""
    rec, _ = zero_pi_runs(args)
    r = rec["lockstep_mixed_subspace"]
    out = {"metric": "eigenpairs/s at resid<=1e-8 (zero-pi sweep, parent [40,40,30], LOBPCG lowest 16)",
           "value": r["eigenpairs_per_s"], "unit": "eigenpairs/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": round(1e3 * r["seconds"], 1), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (seeded circuit parameters)",
           "config": {"workload": "zero_pi_sweep", "dims": [40, 40, 30], "m": 16, "tol": 1e-8,
                      "test_points": args.test_points, "train_points": args.train_points},
           "roofline": rec["roofline_dense_pairs"], "zero_pi_sweep": rec}
    print(json.dumps(out))


def ports_bench(args):
    """Port formation V' O V on a materialized zero-pi parent (48 000 states, 16
    kept): the host formula (numpy, what _port_matrix does on the CPU) against the
    device path (kfc_sandwich), per port, for single and joint ports."""
    import numpy as np

    from kfc import configs, materialize
    from kfc.tree import joint_port, port_matrix

    th = configs.zero_pi_theta(8800)
    spec, E_Cz = configs.zero_pi_spec(th["E_CS"], th["n_g"])
    mn = materialize(configs.zero_pi_tree(E_Cz), spec).root
    singles = [("theta", ("n", 0)), ("theta", ("expiphi", 1)), ("phi", ("phi", 1)), ("phi", ("n", 0)),
               ("zeta", ("n", 0)), ("zeta", ("phi2", 0))]
    joints = [[("theta", ("expiphi", 1)), ("phi", ("expiphi", 1))], [("theta", ("n", 0)), ("zeta", ("n", 0))]]
    out = {}
    for dev in (False, True):
        port_matrix(mn, *singles[0], device=dev)       # warm-up (library load, tuner)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for mode, port in singles:
                port_matrix(mn, mode, port, device=dev)
            for f in joints:
                joint_port(mn, f, device=dev)
        out["device" if dev else "host"] = (time.perf_counter() - t0) / (args.steps * (len(singles) + len(joints)))
    err = max(float(np.max(np.abs(port_matrix(mn, m, p, device=True) - port_matrix(mn, m, p, device=False))))
              for m, p in singles)
    res = {"metric": "port matrices/s (zero-pi parent, 48 000 states x 16 kept)", "value": round(1.0 / out["device"], 2),
           "unit": "ports/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": round(1e3 * out["device"] * (len(singles) + len(joints)), 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded circuit parameters)",
           "config": {"workload": "ports", "dims": list(mn.dims), "kept": int(mn.vectors.shape[1]),
                      "ports": len(singles), "joint_ports": len(joints)},
           "host_seconds_per_port": round(out["host"], 5), "device_seconds_per_port": round(out["device"], 5),
           "speedup": round(out["host"] / out["device"], 2), "max_abs_diff_device_vs_host": err}
    print(json.dumps(res))


def seeds_bench(args):
    """Cold-start inputs of the block solvers at 256^3 (16.8 M states): the full
    diagonal and the product seed (k = 16) of the three-fluxonia operator, and the
    unit-vector seed of an operator whose axes are eigenbases (radix select of the
    20 smallest of 16.8 M diagonal entries), each formed in device memory
    (kfc_op_diagonal_device, kfc_op_product_seed_device) against the host forms
    (kfc_op_diagonal, kfc_op_product_seed) that fill host memory."""
    import numpy as np
    import torch

    import kfc
    from kfc import configs

    n = args.seed_n
    H = configs.three_fluxonia_operator(configs.three_fluxonia_theta(args.seed), N=n)
    rng = np.random.default_rng(1)
    C0 = rng.standard_normal((n, n))
    np.fill_diagonal(C0, 0.0)
    Hu = kfc.KronFreeOperator([n] * 3, [
        kfc.AxisTerm(1.0, [(a, kfc.DiagonalOp(rng.uniform(1, 2, n))) for a in range(3)]),
        kfc.AxisTerm(0.1, [(0, kfc.BandedOp(n, [-1, 1], [np.ones(n), np.ones(n)])),
                           (2, kfc.BandedOp(n, [-1, 1], [np.ones(n), np.ones(n)]))])])
    N = H.N

    def dev_time(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, out

    def host_time(fn):
        t0 = time.perf_counter()
        out = fn()
        return time.perf_counter() - t0, out

    k, ku = 16, 20
    t_dd, Dd = dev_time(H.diagonal_device, args.steps)
    t_dh, Dh = host_time(H.diagonal)
    t_sd, Pd = dev_time(lambda: H.product_seed_device(k), args.steps)
    t_sh, Ph = host_time(lambda: H.product_seed(k))
    t_ud, Ud = dev_time(lambda: Hu.product_seed_device(ku), args.steps)
    t_uh, Uh = host_time(lambda: Hu.product_seed(ku))
    same = bool(np.array_equal(Dd.cpu().numpy(), Dh) and np.array_equal(Pd.cpu().numpy(), Ph)
                and np.array_equal(Ud.cpu().numpy(), Uh))
    log(f"[seeds] diagonal {t_dd * 1e3:.2f} ms (host {t_dh * 1e3:.0f} ms) | product seed k={k} {t_sd * 1e3:.2f} ms "
        f"(host {t_sh * 1e3:.0f} ms) | unit seed k={ku} {t_ud * 1e3:.2f} ms (host {t_uh * 1e3:.0f} ms) | equal {same}")
    res = {"metric": f"product seeds/s (three fluxonia {n}^3, k = {k}, formed in device memory)",
           "value": round(1.0 / t_sd, 2), "unit": "seeds/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": round(1e3 * t_sd, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (seeded circuit parameters)",
           "config": {"workload": "seeds", "dims": [n] * 3, "k": k, "k_unit": ku},
           "roofline": {"bound": "hbm", "achieved": round(N * k * 8 / t_sd / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(N * k * 8 / t_sd / 1e9 / 8000.0, 4), "traffic": None, "kernel": "k_seed_product",
                        "algorithmic_bytes_per_launch": N * k * 8},
           "device_ms": {"diagonal": round(t_dd * 1e3, 3), "product_seed": round(t_sd * 1e3, 3),
                         "unit_seed": round(t_ud * 1e3, 3)},
           "host_ms": {"diagonal": round(t_dh * 1e3, 1), "product_seed": round(t_sh * 1e3, 1),
                       "unit_seed": round(t_uh * 1e3, 1)},
           "diagonal_gbs": round(N * 8 / t_dd / 1e9, 1), "device_equals_host": same}
    print(json.dumps(res))


def drive_bench(args):
    """The matrix-free apply of time evolution (KronFreeApply, kfc_apply_affine): w =
    alpha 2 pi (H - E0) v + w for complex kets on the real three-fluxonia operator
    (--seed-n points per mode), one ket and a block of 4.  Algorithmic bytes per ket:
    read v and w, write w (48 N)."""
    import math

    import torch

    import kfc
    from kfc import configs

    n = args.seed_n
    H = configs.three_fluxonia_operator(configs.three_fluxonia_theta(args.seed), N=n)
    N = H.N
    A = kfc.KronFreeApply(H, 2 * math.pi, -2.5)
    out = {}
    for b in (1, 4):
        shape = (N,) if b == 1 else (b, N)
        v = torch.randn(shape, dtype=torch.complex128, device="cuda")
        w = torch.randn(shape, dtype=torch.complex128, device="cuda")
        A(w, v, -1j, 1.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            A(w, v, -1j, 1.0)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        out[b] = (dt, 48.0 * N * b / dt / 1e9)
        log(f"[drive] {b} ket(s): {dt * 1e3:.3f} ms, {out[b][1]:.0f} GB/s algorithmic")
    dt, gbs = out[4]
    res = {"metric": f"kets/s of w = alpha 2pi (H - E0) v + w (three fluxonia {n}^3, real H, complex kets)",
           "value": round(4 / dt, 2), "unit": "kets/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "c128", "data": "synthetic (seeded circuit parameters, random kets)",
           "config": {"workload": "drive", "dims": [n] * 3, "block": 4},
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(gbs / 8000.0, 4), "traffic": None, "kernel": "kfc_apply_affine (pack, stencil, combine)",
                        "algorithmic_bytes_per_launch": 48 * N * 4},
           "single_ket_ms": round(out[1][0] * 1e3, 3), "single_ket_gbs": round(out[1][1], 1)}
    print(json.dumps(res))


def sternheimer_bench(args):
    """The derivatives' linear solve on the complement (kfc_complement_solve,
    src/resolvent.jl:146-256): the three-fluxonia node at --seed-n^3 keeps its lowest
    10 states (Lanczos), then d psi_n / d E_J1 for all 10 kept states in one lockstep
    solve (shift E_n, right-hand side -Q dH psi_n, Jacobi preconditioner, tol 1e-10)."""
    import numpy as np
    import torch

    import kfc
    from kfc import configs
    from kfc.resolvent import Complement, Resolvent, apply_resolvent

    n = args.seed_n
    th = configs.three_fluxonia_theta(args.seed)
    H = configs.three_fluxonia_operator(th, N=n)
    h = 1e-3
    thp, thm = dict(th), dict(th)
    thp["E_J"] = [th["E_J"][0] + h] + list(th["E_J"][1:])
    thm["E_J"] = [th["E_J"][0] - h] + list(th["E_J"][1:])
    Hp, Hm = configs.three_fluxonia_operator(thp, N=n), configs.three_fluxonia_operator(thm, N=n)
    m = 10
    t0 = time.perf_counter()
    E, V = kfc.solve(kfc.Lanczos(tol=1e-10), H, m)
    t_solve = time.perf_counter() - t0
    Vt = torch.from_numpy(np.ascontiguousarray(V.T)).to("cuda")
    R = ((Hp @ Vt) - (Hm @ Vt)) / (2 * h)      # dH/dE_J1 psi_n (H is linear in E_J1)
    Hp.sync(), Hm.sync()
    R = R.cpu().numpy().T
    D = H.diagonal_device()
    # the complement solve alone, its right-hand sides -Q R already in device memory
    Rc = R - V @ (V.T @ R)
    B = torch.from_numpy(np.ascontiguousarray(-Rc.T)).to("cuda")
    scale = np.linalg.norm(R, axis=0)
    kfc.complement_solve(H, D, Vt, B[:2], E[:2], 1e-10, 20000, scale=scale[:2])   # warm-up (tuner, buffers)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    X, it, resid, fs, af, fr = kfc.complement_solve(H, D, Vt, B, E, 1e-10, 20000, scale=scale)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # end to end as _apply_resolvent: the host sums over the kept states around it
    log_ = []
    res = Resolvent(np.asarray(E), V, Complement(H, D, Vt, 1e-10, 20000, log_), 1e-9)
    t0 = time.perf_counter()
    W, clashes = apply_resolvent(res, R, E, [[j] for j in range(m)])
    dt_e2e = time.perf_counter() - t0
    # per CG iteration 25 passes over the N x 10 block: the 10 shifted applies (2), two
    # Grams against the basis (2 each), two projection updates that also reduce a
    # column dot (4 each), the X/R update with the residual norms (6), the direction
    # update (3) and the preconditioner (2)
    N = H.N
    gbs = 25 * N * 8 * m * it / dt / 1e9
    log(f"[sternheimer] {n}^3: Lanczos {t_solve:.2f} s; complement solve of 10 columns {dt:.2f} s, {it} CG "
        f"iterations, {dt / it * 1e3:.2f} ms per iteration (~{gbs:.0f} GB/s), residual {np.max(resid):.2e}; "
        f"apply_resolvent end to end {dt_e2e:.2f} s")
    out = {"metric": f"derivative columns/s (three fluxonia {n}^3, d psi_n / d E_J1 of the 10 kept states)",
           "value": round(m / dt, 3), "unit": "columns/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": round(dt * 1e3, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (seeded circuit parameters)",
           "config": {"workload": "sternheimer", "dims": [n] * 3, "kept": m, "tol": 1e-10},
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(gbs / 8000.0, 4), "traffic": None,
                        "kernel": "complement solve (25 block passes per CG iteration)"},
           "cg_iterations": it, "ms_per_iteration": round(dt / it * 1e3, 3),
           "max_true_residual": float(np.max(resid)), "lanczos_seconds": round(t_solve, 2),
           "apply_resolvent_seconds": round(dt_e2e, 3), "from_start": fs, "at_floor": af,
           "finite": bool(np.all(np.isfinite(W)))}
    print(json.dumps(out))


def transmon_bench(args):
    """512 coupled-transmon instances on periodic grids ({256..768} points per
    mode, complex leaves keeping 8, tops keeping 20 = 10 240 eigenpairs): plain
    instance after instance against all leaves batched."""
    import numpy as np

    from kfc import transmon_sweep as T

    insts = T.draw(args.instances, seed=args.seed)
    sw = T.TransmonSweep(contexts=args.contexts)
    mode, method = args.leaf_mode, args.leaf_method
    sw.run_batched(insts[:2], mode=mode, method=method)      # warm-up
    sw.run_plain(insts[:1])
    plain, t_plain = sw.run_plain(insts[:args.plain_points])
    n_plain = sum(len(t.energies) for t in plain
                  if all(l.stats["max_residual"] <= 1e-8 for l in t.children))
    rate_plain = n_plain / t_plain
    log(f"[transmons] plain f64: {args.plain_points} instances {t_plain:.2f} s -> {rate_plain:.1f} pairs/s")
    tops, wall, t_leaves = sw.run_batched(insts, mode=mode, method=method)
    ok = [t for t in tops if all(l.stats["max_residual"] <= 1e-8 for l in t.children)]
    value = sum(len(t.energies) for t in ok) / wall
    log(f"[transmons] batched {method} {mode}: {len(insts)} instances {wall:.2f} s (leaves {t_leaves:.2f} s) "
        f"-> {value:.1f} pairs/s")
    dE = max(float(np.max(np.abs(a.energies - b.energies) / np.abs(a.energies))) for a, b in zip(plain, tops))
    out = {"metric": "eigenpairs/s at resid<=1e-8 (512 coupled transmons on periodic grids, tops keep 20)",
           "value": round(value, 2), "unit": "eigenpairs/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": round(1e3 * wall, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64" if mode == "mixed" else "f64",
           "data": "synthetic (seeded circuit parameters)",
           "config": {"workload": "transmon_sweep", "instances": len(insts), "grid_sizes": list(T.GRID_SIZES),
                      "order": 8, "k_leaf": 8, "k_top": 20, "contexts": args.contexts, "mode": mode,
                      "leaf_solver": ("batched ChebyshevFilter degree 60, extra 8, tol 1e-9" if method == "chebyshev"
                                      else "Lanczos krylovdim 48, tol 1e-9, one host thread per context")},
           "plain_fp64": {"instances": args.plain_points, "seconds": round(t_plain, 3),
                          "eigenpairs_per_s": round(rate_plain, 2)},
           "speedup_vs_plain": round(value / rate_plain, 2), "seconds_leaves": round(t_leaves, 3),
           "max_rel_dE_batched_vs_plain": dE}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--mode", default="f64", choices=["f64", "mixed"])
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-sweep", action="store_true", help="leave the zero-pi sweep sub-record out")
    ap.add_argument("--verbosity", type=int, default=20, help="Lanczos progress line every this many cycles")
    ap.add_argument("--workload", default="three_fluxonia",
                    choices=["three_fluxonia", "zero_pi_sweep", "transmon_sweep", "ports", "seeds", "drive",
                             "sternheimer", "pairs"])
    ap.add_argument("--seed-n", type=int, default=256, help="seeds, drive and sternheimer workloads: points per mode")
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only an elementwise copy of known size (the reference for HBM byte counters)")
    ap.add_argument("--contexts", type=int, default=16, help="device contexts (streams) of a batched sweep")
    ap.add_argument("--contexts-only", action="store_true",
                    help="zero-pi sweep: batched solves as independent solves on contexts instead of lockstep")
    ap.add_argument("--test-points", type=int, default=64)
    ap.add_argument("--plain-points", type=int, default=8)
    ap.add_argument("--train-points", type=int, default=1000)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--nbasis", type=int, default=64, help="frame basis size of the warm start")
    ap.add_argument("--subspace-basis", type=int, default=256, help="directions of the learned subspace start")
    ap.add_argument("--sweep-mode", default="mixed", choices=["f64", "mixed"])
    ap.add_argument("--leaf-mode", default="mixed", choices=["f64", "mixed"], help="transmon sweep leaf solves")
    ap.add_argument("--leaf-method", default="chebyshev", choices=["chebyshev", "lanczos"])
    ap.add_argument("--summarize-pmc", action="store_true",
                    help="write a profiles JSON of one kernel's HBM bytes per launch from rocprofv3 CSVs (no GPU)")
    ap.add_argument("--pmc-kernel", default="k_tsmm")
    ap.add_argument("--pmc-workload", default="")
    ap.add_argument("--pmc-command", default="")
    ap.add_argument("--pmc-fetch")
    ap.add_argument("--pmc-write")
    ap.add_argument("--pmc-copy-fetch")
    ap.add_argument("--pmc-copy-write")
    ap.add_argument("--pmc-out")
    ap.add_argument("--pmc-stat", default="mean", choices=["mean", "median"])
    args = ap.parse_args()
    if args.summarize_pmc:
        return summarize_pmc(args)
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        return calibrate_copy()
    if args.workload == "zero_pi_sweep":
        return sweep_bench(args)
    if args.workload == "ports":
        return ports_bench(args)
    if args.workload == "seeds":
        return seeds_bench(args)
    if args.workload == "drive":
        return drive_bench(args)
    if args.workload == "sternheimer":
        return sternheimer_bench(args)
    if args.workload == "transmon_sweep":
        return transmon_bench(args)
    if args.workload == "pairs":
        return pairs_bench(args)
    import numpy as np
    import torch

    import kfc
    from kfc import configs

    theta = configs.three_fluxonia_theta(args.seed)
    m = 10
    log(f"[bench] three fluxonia N={args.N}^3 theta={ {k: [round(float(x), 4) for x in v] for k, v in theta.items()} }")

    shapes = []   # the fused stencil's launch shape each solve used (the tuner's pick)

    def step(mode):
        H = configs.three_fluxonia_operator(theta, N=args.N)
        E, V, st = kfc.solve_stats(kfc.Lanczos(tol=args.tol, mode=mode, verbosity=args.verbosity), H, m)
        shapes.append(H.stencil)
        return H, E, st

    # apply bandwidth (the dominant kernel)
    H = configs.three_fluxonia_operator(theta, N=args.N)
    assert H.fused
    t64, gbs64 = apply_bandwidth(H, "f64")
    t32, gbs32 = apply_bandwidth(H, "f32")
    log(f"[bench] fused apply: f64 {t64 * 1e6:.1f} us {gbs64:.0f} GB/s | f32 {t32 * 1e6:.1f} us {gbs32:.0f} GB/s")
    del H
    torch.cuda.empty_cache()

    for i in range(args.warmup):
        _, E, st = step(args.mode)
        log(f"[bench] warmup {i}: {st['seconds_total']:.2f} s, matvec {st['n_matvec']}, resid {st['max_residual']:.2e}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    counted, stats = 0, []
    for i in range(args.steps):
        _, E, st = step(args.mode)
        counted += m if st["max_residual"] <= 1e-8 else 0
        stats.append(st)
        log(f"[bench] step {i}: {st['seconds_total']:.2f} s, matvec {st['n_matvec']} (f32 {st['n_matvec_f32']}), "
            f"resid {st['max_residual']:.2e}, E0 {E[0]:.10f}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = counted / wall
    st = stats[-1]
    # roofline of the solve's apply kernel (the fused stencil): HIP-event time of every
    # apply on the library's stream over the timed solve, per column.  Nearly all
    # are Chebyshev steps y = a Hx + b x + g z: algorithmic bytes 3 N s (read x and
    # z, write y); the explicit check's block apply and the acceptance residuals
    # are a few dozen columns of 2 N s, counted at the same rate.
    cols = st["n_matvec"] + m
    avg_col = st["seconds_apply_gpu"] / max(1, cols)
    s_bytes = 8 if args.mode == "f64" else 4
    bytes_per_launch = 3 * (args.N ** 3) * s_bytes
    achieved = bytes_per_launch / avg_col / 1e9
    traffic, traffic_src = measured_traffic(f"stencil_traffic_{args.N}.json") if args.mode == "f64" else (None, None)
    roof_apply = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                  "kernel": STENCIL_KERNELS.get(shapes[-1][0], f"variant {shapes[-1][0]}"),
                  "launch_shape": list(shapes[-1]),
                  "algorithmic_bytes_per_launch": bytes_per_launch, "avg_launch_us": round(avg_col * 1e6, 1),
                  "seconds": round(st["seconds_apply_gpu"], 3)}
    # roofline of the block passes (k_tsmm: Gram-only first Gram-Schmidt pass,
    # chained Cholesky scaling, thick restart): HIP-event time of every launch over
    # the timed solve and their algorithmic bytes (columns read + columns written,
    # counted by the library per launch); per launch = totals / launches
    roof = roof_apply
    if st.get("n_ortho_calls", 0) > 0:
        no = st["n_ortho_calls"]
        b_o, t_o = st["bytes_ortho"] / no, st["seconds_ortho_gpu"] / no
        a_o = b_o / t_o / 1e9
        traffic_o, src_o = measured_traffic(f"tsmm_traffic_{args.N}.json") if args.mode == "f64" else (None, None)
        roof_ortho = {"bound": "hbm", "achieved": round(a_o, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(a_o / HBM_PEAK_GBS, 4), "traffic": traffic_o, "traffic_source": src_o,
                      "kernel": "k_tsmm",
                      "algorithmic_bytes_per_launch": round(b_o), "avg_launch_us": round(t_o * 1e6, 1),
                      "launches": no, "seconds": round(st["seconds_ortho_gpu"], 3)}
        # the line's roofline is the kernel that holds the GPU longer in the solve
        if st["seconds_ortho_gpu"] > st["seconds_apply_gpu"]:
            roof, other = roof_ortho, {"roofline_apply": roof_apply}
        else:
            other = {"roofline_ortho": roof_ortho}
    else:
        other = {}
    extra = {}
    if not args.skip_plain and args.mode != "f64":
        _, E64, st64 = step("f64")
        extra["plain_fp64"] = {"seconds": round(st64["seconds_total"], 3), "n_matvec": st64["n_matvec"],
                               "max_residual": st64["max_residual"],
                               "eigenpairs_per_s": round(m / st64["seconds_total"], 4),
                               "max_abs_dE_vs_mixed": float(np.max(np.abs(E64 - E)))}
        log(f"[bench] plain f64: {st64['seconds_total']:.2f} s matvec {st64['n_matvec']}")
    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline(theta, args.N, st["n_matvec"], m)
    if not args.skip_sweep:
        extra["zero_pi_sweep"], _ = zero_pi_runs(args)
    out = {"metric": "eigenpairs/s at resid<=1e-8 (three fluxonia 512^3, Lanczos lowest 10)",
           "value": round(value, 5), "unit": "eigenpairs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if args.mode == "mixed" else "f64",
           "data": "synthetic (seeded circuit parameters)",
           "config": {"workload": "three_fluxonia_flat", "dims": [args.N] * 3, "order": 2, "m": m,
                      "solver": "Lanczos", "mode": args.mode, "tol": args.tol, "seed": args.seed},
           "roofline": roof, **other, "cpu_baseline": cpu,
           "apply": {"f64_us": round(t64 * 1e6, 1), "f64_gbs": round(gbs64, 1), "f32_us": round(t32 * 1e6, 1),
                     "f32_gbs": round(gbs32, 1)},
           "solve": {"n_matvec": st["n_matvec"], "n_matvec_f32": st["n_matvec_f32"], "restarts": st["iterations"],
                     "seconds_apply_gpu": round(st["seconds_apply_gpu"], 3),
                     "seconds_ortho_gpu": round(st.get("seconds_ortho_gpu", 0.0), 3), "max_residual": st["max_residual"]},
           **extra}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
