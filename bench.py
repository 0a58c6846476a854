"""Throughput benchmark of the circulant-embedding engine on one MI355X.

Workload (one step): 64 realizations of a 3D Gaussian random field on 256^3
cells of [0,1]^3, Matern nu = 1.5 (matern32), variance 2, axiparallel
correlation lengths (0.1, 0.1, 0.05), embedding factor 2 -> 512^3 extended
grid, DCT matrix (FP64 eigenvalues) + R2C field, FP32 realizations from the
device Philox4x32-10 stream (seed 7), log-normal value transform, output
resident in HBM. value = fields per second over the timed steps.

Also reported: per-kernel device times (HIP events around every launch of the
timed steps, on the stream the launch runs on), the
roofline of the dominant kernel against 8 TB/s and against the VALU issue
peak (its launch time without the two-stream overlap), the FP64 per-realization GPU
path on the same grid, and the FP64 CPU oracle on a bounded sample (one
realization) as the CPU baseline.

The other configurations run after the headline workload, each in a child
process (`--leg NAME`), and land under `other_configs`: the 128^3 covariance
inverse (plain CG against embedding, spectral and learned PCG), the 1024^3
DCT/DST field (8 realizations, sign transform) and the uneven batch of 1000
2D fields (one captured graph). `--no-legs` skips them.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parafields-core_amd"))
import pf_binding as pfb  # noqa: E402
import workloads as wl  # noqa: E402

METRIC = json.load(open(os.path.join(ROOT, "BASELINE.json")))["metric"] if os.path.exists(
    os.path.join(ROOT, "BASELINE.json")) else "fields/s at 512^3 embedding"
HBM_PEAK = 8.0e12

CELLS = wl.ENSEMBLE_CELLS
KW = wl.ENSEMBLE_KW
SEED = wl.ENSEMBLE_SEED
BATCH = 64


def alg_bytes(cells, batch, s_c=8, s_r=4):
    """Algorithmic HBM bytes of the pruned 3-pass schedule per field and per
    kernel (R2C field, 3D): pass z writes H*N1*n2 complex; pass y reads that
    and writes H*n1*n2 complex; pass x reads that and writes n0*n1*n2 reals;
    FP64 DCT-half eigenvalues read once per batch."""
    n0, n1, n2 = cells
    N0, N1, N2 = (2 * c for c in cells)
    H = N0 // 2 + 1
    P = (N0 // 2 + 1) * (N1 // 2 + 1) * (N2 // 2 + 1)
    k = {
        "gen_last": s_c * H * N1 * n2 + 8 * P / batch,
        "gen_mid": s_c * H * N1 * n2 + s_c * H * n1 * n2,
        "gen_c2r_x": s_c * H * n1 * n2 + s_r * n0 * n1 * n2,
    }
    return sum(k.values()), k


def timings(plan):
    names = ctypes.create_string_buffer(4096)
    ms = (ctypes.c_double * 64)()
    n = ctypes.c_int32()
    lib = pfb.load()
    lib.pf_last_timings.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32,
                                    ctypes.POINTER(ctypes.c_double), ctypes.c_int32,
                                    ctypes.POINTER(ctypes.c_int32)]
    lib.pf_set_profiling.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    pfb.check(lib.pf_last_timings(plan.h, names, 4096, ms, 64, ctypes.byref(n)))
    labels = names.value.decode().split(";")[: n.value]
    return list(zip(labels, [ms[i] for i in range(min(n.value, 64))]))


def measured_traffic(kernel, fields_per_launch):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes of this same workload, stored per field (each counter
    scaled by what it reads for a 1 GiB device copy) in
    profiles/pmc_traffic_config2.json; None when that record is absent."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_config2.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get("per_field_bytes", {}).get(kernel)
    if rec is None:
        return None
    return int((rec["fetch"] + rec["write"]) * fields_per_launch)


# VALU issue peak of MI355X: gfx950 issues wave64 FP32 add / mul / FMA
# instructions (not packed) at two per SIMD every 4 cycles (1.80 measured for
# independent FMA chains): 256 CUs x 4 SIMDs x 2 / 4 cycles x 2.4 GHz. The
# transcendentals (v_log, v_sin, v_cos, v_sqrt) take about 4 and the 64-bit
# integer multiply-add of Philox about 3 of those half-slots.
VALU_PEAK = 256 * 4 * 2 / 4 * 2.4e9


def valu_per_field(kernel):
    """wave64 VALU instructions per field of `kernel` (SQ_INSTS_VALU of a
    rocprofv3 --pmc pass over this workload, profiles/pmc_valu_config2.json);
    None when that record is absent."""
    path = os.path.join(ROOT, "profiles", "pmc_valu_config2.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path)).get("per_field", {}).get(kernel)
    return None if rec is None else float(rec["SQ_INSTS_VALU"])


def set_profiling(plan, on):
    lib = pfb.load()
    lib.pf_set_profiling.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    pfb.check(lib.pf_set_profiling(plan.h, int(on)))


def run_gpu(steps, warmup, batch):
    """The timed region: `steps` calls of 64 realizations each. Every launch
    is bracketed by HIP events on the stream it runs on (the generation pass
    and the y/x passes run on two streams, chunk by chunk), so the per-kernel
    durations come from the timed steps themselves."""
    import torch
    torch.cuda.set_device(0)
    d = pfb.make_desc(3, CELLS, [1.0, 1.0, 1.0], precision="fp32", **KW)
    plan = pfb.Plan(d)
    t0 = time.time()
    st = plan.eigs()
    eig_s = time.time() - t0
    out = torch.empty((batch, plan.domain), dtype=torch.float32, device="cuda")
    # the autotuner picks the launch choices of this batch (outside the timed
    # region; PF_TUNE_CACHE names an optional cache file of earlier results)
    tune = plan.autotune(batch, out, seed=SEED, transform="logNormal",
                         cache=os.environ.get("PF_TUNE_CACHE"))
    print(f"bench: autotune {tune}", file=sys.stderr, flush=True)
    set_profiling(plan, True)
    for w in range(warmup):
        plan.generate(batch, seed=SEED, first=w * batch, transform="logNormal", out=out)
    torch.cuda.synchronize()
    per_kernel = {}
    t0 = time.perf_counter()
    for s in range(steps):
        plan.generate(batch, seed=SEED, first=(warmup + s) * batch, transform="logNormal",
                      out=out)
        for name, ms in timings(plan):
            per_kernel.setdefault(name, []).append(ms)
        print(f"bench: step {s + 1}/{steps} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # realization 0 of the last timed step (Philox realization r0) as the
    # Gaussian field, i.e. the log of its log-normal output, for the error
    # against the FP64 oracle on identical noise
    r0 = (warmup + steps - 1) * batch
    field0 = out[0].double().log().cpu().numpy()
    # after the timed steps: one step with the kernels run one after another
    # (no two-stream overlap), for the launch times of each kernel alone
    tuned = plan.tuning()
    plan.set_tuning(overlap=0)
    alone = {}
    plan.generate(batch, seed=SEED, first=(warmup + steps) * batch, transform="logNormal", out=out)
    for name, ms in timings(plan):
        alone.setdefault(name, []).append(ms)
    plan.set_tuning(**tuned)
    set_profiling(plan, False)
    stats = {"mean_log": float(field0.mean()), "var_log": float(field0.var()), "realization": r0}
    # PCIe-inclusive rate: 16 realizations into a host buffer (the library
    # stages through its own device buffers and copies back over the link)
    host = np.empty((16, plan.domain), dtype=np.float32)
    plan.generate(16, seed=SEED, first=0, transform="logNormal", out=host)
    t0 = time.perf_counter()
    plan.generate(16, seed=SEED, first=16, transform="logNormal", out=host)
    pcie_fps = 16 / (time.perf_counter() - t0)
    del host
    plan.close()
    del out
    torch.cuda.empty_cache()
    stats["pcie_inclusive_fields_per_s"] = round(pcie_fps, 1)
    stats["alone"] = alone
    stats["tuning"] = tune
    return elapsed, per_kernel, st, eig_s, stats, field0


def run_gpu_fp64_single(reps, r0):
    """Plain FP64 per-realization path on the same grid (batch 1)."""
    import torch
    d = pfb.make_desc(3, CELLS, [1.0, 1.0, 1.0], precision="fp64", **KW)
    plan = pfb.Plan(d)
    plan.eigs()
    out = torch.empty((1, plan.domain), dtype=torch.float64, device="cuda")
    plan.generate(1, seed=SEED, transform="logNormal", out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(reps):
        plan.generate(1, seed=SEED, first=r, transform="logNormal", out=out)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    field0 = plan.generate(1, seed=SEED, first=r0)[0]
    plan.close()
    del out
    torch.cuda.empty_cache()
    return reps / el, field0


def field_error(got, want):
    """max |got - want| relative to std(want) and to max |want|, and
    ||got - want|| / ||want|| (in FP64)."""
    g = np.asarray(got, dtype=np.float64)
    w = np.asarray(want, dtype=np.float64)
    e = np.abs(g - w).max()
    return {"max_abs_over_std": float(e / w.std()), "max_abs_over_max": float(e / np.abs(w).max()),
            "rel_l2": float(np.linalg.norm(g - w) / np.linalg.norm(w))}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(r0, gpu_fields=None, reps=3, single_thread=True):
    """FP64 CPU oracle (restatement of the reference loops, threaded over
    transform lines and noise samples) on the same workload: Philox noise of
    one realization + generate + log-normal transform per field; eigenvalue
    setup excluded (timed separately). All host cores (at most 16): one warm-up
    field, then the median of `reps` fields. One thread: one field. The
    oracle's Gaussian field of realization r0 is then the reference for the
    GPU fields of the same noise (`gpu_fields`: name -> field)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as orc
    nproc = os.cpu_count() or 1
    cores = min(nproc, 16)
    d = pfb.make_desc(3, CELLS, [1.0, 1.0, 1.0], precision="fp64", **KW)
    n = int(orc.sizes(d).noise_per_field)

    def one(real):
        t0 = time.perf_counter()
        nz = orc.noise_philox(SEED, real, 0, n)
        f = orc.generate(d, nz)
        del nz
        g = f.copy()
        orc.value_transform(g, "logNormal")
        return time.perf_counter() - t0, f

    orc.set_threads(cores)
    t0 = time.perf_counter()
    orc.eigs(d, allow_negative=True)
    eig_s = time.perf_counter() - t0
    print("bench: cpu baseline warm-up", file=sys.stderr, flush=True)
    _, f0 = one(r0)
    times = []
    for i in range(reps):
        el, _ = one(r0 + 1 + i)
        times.append(el)
        print(f"bench: cpu baseline {cores} threads field {i + 1}/{reps}: {el:.2f} s",
              file=sys.stderr, flush=True)
    med = float(np.median(times))
    single = None
    if single_thread:
        orc.set_threads(1)
        el1, _ = one(r0 + 1 + reps)
        print(f"bench: cpu baseline 1 thread: {el1:.2f} s", file=sys.stderr, flush=True)
        single = {"value": round(1.0 / el1, 5), "unit": "fields/s", "cores": 1,
                  "s_per_field": round(el1, 2), "fields": 1}
        orc.set_threads(cores)
    acc = {k: field_error(v, f0) for k, v in (gpu_fields or {}).items() if v is not None}
    return {"value": round(1.0 / med, 5), "unit": "fields/s", "cores": cores, "kind": "port",
            "accuracy_vs_oracle": acc,
            "s_per_field_median": round(med, 3), "s_per_field_all": [round(t, 3) for t in times],
            "single_thread": single, "nproc": nproc, "cpu_model": cpu_model(),
            "eig_setup_s": round(eig_s, 2),
            "sample": f"FP64 CPU oracle on the same 256^3/512^3 workload, one realization per "
                      f"field (Philox noise, generation, log-normal transform), threaded over "
                      f"noise samples and transform lines: {cores} threads, 1 warm-up field "
                      f"then the median of {reps} fields; 1 thread: one field; eigenvalue "
                      f"setup ({eig_s:.1f} s) excluded"}


def run_uneven(nfields=1000, seed=3, reps=3, branches=16):
    """Config 3: nfields 2D fields, sides log-uniform on {64..1024} (2^k and
    3*2^k), Gaussian covariance with full geometric anisotropy, R2C matrix
    with per-field FP64 eigenvalues, FP32 realizations, whole batch captured
    in one hipGraph on `branches` forked streams and replayed (one capture;
    the branch-count autotuner is not run here)."""
    import torch
    sides = wl.UNEVEN_SIDES
    descs, shapes, falg = wl.uneven_batch(nfields, seed)
    alg = sum(falg)
    t0 = time.perf_counter()
    b = pfb.Batch(descs)
    b.set_branches(branches)
    out = torch.empty(int(b.offsets[-1]), dtype=torch.float32, device="cuda")
    b.generate(seed=11, out=out, use_graph=True)   # eager preparation + capture + first launch
    setup = time.perf_counter() - t0
    torch.cuda.synchronize()
    print("bench: uneven batch captured", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        b.generate(seed=11, out=out, use_graph=True)
    el = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    b.generate(seed=11, out=out, use_graph=False)
    eager = time.perf_counter() - t0
    # per-size throughput: device time of every field in one timed eager
    # (serial) pass, summed over the fields whose larger side is the key
    ms = b.field_times(seed=11, out=out)
    per_size = {}
    for side in sides:
        idx = [i for i, (n0, n1) in enumerate(shapes) if max(n0, n1) == side]
        if idx:
            t = float(ms[idx].sum()) * 1e-3
            per_size[str(side)] = {"fields": len(idx), "device_ms": round(t * 1e3, 3),
                                   "fields_per_s": round(len(idx) / t, 1),
                                   "alg_GBps": round(sum(falg[i] for i in idx) / t / 1e9, 1)}
    b.close()
    del out
    torch.cuda.empty_cache()
    return {"fields": nfields, "fields_per_s": round(nfields / el, 1),
            "ms_per_batch": round(el * 1e3, 2), "eager_ms_per_batch": round(eager * 1e3, 2),
            "branches": branches, "alg_GBps": round(alg / el / 1e9, 1),
            "frac_of_hbm": round(alg / el / HBM_PEAK, 4), "setup_s": round(setup, 2),
            "per_size_class": per_size,
            "note": "per-field FP64 eigenvalues + one FP32 realization each, one graph launch; "
                    "per_size_class: fields keyed by their larger side, device time of each "
                    "field in one eager serial pass (HIP events around every field)"}


def run_cg(ntrain=2000, cells=128, held=(0.1, 0.25), max_it=20000):
    """Config 4: 128^3 exponential + nugget 1e-3; learned MLP trained on ntrain
    method-generated parameter sets; plain CG vs learned PCG to 1e-8 (FP64
    true residual), RHS N(0,1) seed 11."""
    sys.path.insert(0, os.path.join(ROOT, "parafields-core_amd"))
    import learned_precond as lp
    t0 = time.perf_counter()
    cases = lp.sample_cases(ntrain, seed=4, cells=cells)
    x, y = lp.build_training_set(cases)
    data_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    model, losses = lp.train(x, y, steps=3000, lr=2e-3)
    train_s = time.perf_counter() - t0
    b = np.random.default_rng(11).standard_normal(cells ** 3)
    # algorithmic HBM bytes of one iteration. A product with the 2x embedding
    # (N = 2 cells, H = cells + 1 kept x frequencies) is five passes: x R2C
    # reads the vector (V) and writes the x half spectrum (A = cells^2 H
    # complex); the forward y pass reads A and writes B = H N cells complex;
    # the fused z pass reads B, the folded DCT eigenvalue table ((cells+1)^3
    # doubles) and writes B; the inverse y pass reads B and writes A; the x
    # C2R pass reads A and writes V. The CG vector work is p.Ap (2V), the x/r
    # update with r.r (6V) and p = z + beta p (3V); a preconditioned
    # iteration adds a second product and r.z (2V).
    V = 8 * cells ** 3
    A = 16 * cells ** 2 * (cells + 1)
    B = 16 * (cells + 1) * 2 * cells * cells
    product = 2 * V + 4 * A + 4 * B + 8 * (cells + 1) ** 3
    alg_iter = {"plain": product + 11 * V}
    for mode in ("embed_pcg", "spectral_pcg", "learned_pcg"):
        alg_iter[mode] = 2 * product + 13 * V
    res = []
    modes = ("plain", "embed_pcg", "spectral_pcg", "learned_pcg")
    for l in held:
        case = lp.CovCase(variance=1.0, corr_length=[l], cells=cells)
        row = {"corr_length": l}
        for mode in modes:
            _, st = lp.solve(case, b, mode, model if mode == "learned_pcg" else None,
                             max_iterations=max_it)
            row[mode] = {k: (float(f"{v:.4g}") if k == "rel_residual" else round(v, 6))
                         if isinstance(v, float) else v for k, v in st.items()}
            per_it = st["seconds"] / max(1, st["iterations"])
            row[mode]["us_per_iteration"] = round(per_it * 1e6, 1)
            row[mode]["alg_MB_per_iteration"] = round(alg_iter[mode] / 1e6, 1)
            row[mode]["frac_of_hbm"] = round(alg_iter[mode] / per_it / HBM_PEAK, 4)
        sp, sl = row["plain"], row["learned_pcg"]
        row["iteration_ratio"] = round(sp["iterations"] / max(1, sl["iterations"]), 2)
        row["time_ratio"] = round(sp["seconds"] / max(1e-9, sl["seconds"]), 2)
        res.append(row)
    return {"cells": cells, "train_sets": ntrain, "data_s": round(data_s, 1),
            "train_s": round(train_s, 1), "final_train_mse": round(float(losses[-1]), 5),
            "modes": {"plain": "unpreconditioned CG", "embed_pcg": "embedding-inverse PCG "
                      "(the reference's preconditioner)", "spectral_pcg": "radial spectral "
                      "PCG with the exact bin targets of the case", "learned_pcg": "radial "
                      "spectral PCG with the network's predicted multipliers"},
            "cases": res}


def run_dctdst(batch=8, steps=1):
    """Config 5: 1024^3 cells, DCT/DST field backend (2048^3 even-symmetric
    embedding), Matern nu=1.5 with a random rotation and lengths in
    [0.01, 0.05] (geometric anisotropy evaluated by the DCT matrix), FP64
    eigenvalues in HBM, 8 FP32 realizations per batch, sign transform at 0.5."""
    import torch
    d = wl.dctdst_1024("fp32")
    plan = pfb.Plan(d)
    t0 = time.perf_counter()
    st = plan.eigs()
    eig_s = time.perf_counter() - t0
    out = torch.empty((batch, plan.domain), dtype=torch.float32, device="cuda")
    plan.generate(1, seed=13, transform="sign", param=0.5, out=out)
    torch.cuda.synchronize()
    plan.set_profiling(True)
    per_kernel = {}
    t0 = time.perf_counter()
    for s in range(steps):
        plan.generate(batch, seed=13, first=s * batch, transform="sign", param=0.5, out=out)
        for name, ms in plan.last_timings():
            per_kernel[name] = per_kernel.get(name, 0.0) + ms
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    plan.set_profiling(False)
    P = 1025 ** 3
    alg = 4 * 12 * P + 4 * 1024 ** 3 + 8 * P / batch
    plan.close()
    del out
    torch.cuda.empty_cache()
    return {"batch": batch, "fields_per_s": round(batch / el, 3), "s_per_batch": round(el, 3),
            "alg_GB_per_field": round(alg / 1e9, 2),
            "alg_GBps": round(alg * batch / el / 1e9, 1),
            "frac_of_hbm": round(alg * batch / el / HBM_PEAK, 4), "eig_setup_s": round(eig_s, 2),
            "kernel_ms_per_field": {k: round(v / (steps * batch), 3) for k, v in per_kernel.items()},
            "eig_negative": st["negative"]}


LEGS = {"uneven_batch_2d": run_uneven, "cg_vs_learned_pcg": run_cg, "dctdst_1024": run_dctdst}


def run_leg_child(name, timeout_s):
    """One leg of another configuration in a child process (`bench.py --leg
    name`), so that whatever happens to it leaves the headline line intact;
    the child's last stdout line is its JSON result."""
    import subprocess
    print(f"bench: {name} leg (child process)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--leg", name],
                           stdout=subprocess.PIPE, timeout=timeout_s, text=True)
    except subprocess.TimeoutExpired:
        return {"error": f"leg did not finish in {timeout_s} s"}
    lines = [x for x in r.stdout.splitlines() if x.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"leg exited with status {r.returncode}"}
    res = json.loads(lines[-1])
    res["leg_wall_s"] = round(time.perf_counter() - t0, 1)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-single", action="store_true",
                    help="skip the one-thread CPU baseline field (about a minute)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run a 1 GiB device-to-device copy of known size (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters of a --pmc run)")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the other configurations (uneven batch, CG, 1024^3 DCT/DST), "
                         "which otherwise run after the headline workload, each in a child "
                         "process")
    ap.add_argument("--leg", choices=sorted(LEGS),
                    help="run one other configuration in this process and print its JSON "
                         "result (what the child processes of the default run do)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine holds the whole extended grid on one "
              "device and has no multi-GPU path (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("bench: copy probe of 1073741824 bytes x3 done", file=sys.stderr, flush=True)
        del src, dst
        torch.cuda.empty_cache()
    if a.leg:
        print(json.dumps(LEGS[a.leg]()), flush=True)
        return
    elapsed, per_kernel, st, eig_s, stats, field32 = run_gpu(a.steps, a.warmup, a.batch)
    fields = a.steps * a.batch
    fps = fields / elapsed
    total_b, kb = alg_bytes(CELLS, a.batch)
    # per launch: one launch of each kernel processes batch / launches-per-step fields
    avg = {k: float(np.mean(v)) for k, v in per_kernel.items()}
    per_step = {k: len(v) / a.steps for k, v in per_kernel.items()}
    fields_per_launch = {k: a.batch / per_step[k] for k in per_kernel}
    total_ms = {k: avg[k] * per_step[k] for k in avg}
    dom = max(total_ms, key=total_ms.get)
    achieved = kb[dom] * fields_per_launch[dom] / (avg[dom] * 1e-3)
    kernels = {k: {"avg_ms_per_launch": round(avg[k], 4), "launches_per_step": per_step[k],
                   "fields_per_launch": fields_per_launch[k],
                   "alg_GBps": round(kb[k] * fields_per_launch[k] / (avg[k] * 1e-3) / 1e9, 1)}
               for k in avg if k in kb}
    traffic = measured_traffic(dom, fields_per_launch[dom])
    # VALU roofline of the same kernel: its instructions per launch over its
    # launch time without the overlap (one launch of all batch fields)
    alone = stats.pop("alone")
    vpf = valu_per_field(dom)
    valu_roof = None
    if vpf is not None and dom in alone:
        t_alone = float(np.mean(alone[dom])) * 1e-3 / a.batch  # seconds per field
        valu_roof = {"bound": "valu", "kernel": dom, "achieved": round(vpf / t_alone / 1e9, 1),
                     "peak": VALU_PEAK / 1e9, "unit": "G wave64 VALU instructions/s",
                     "frac": round(vpf / t_alone / VALU_PEAK, 4),
                     "valu_insts_per_field": int(vpf), "ms_per_field_alone": round(t_alone * 1e3, 4),
                     "source": "SQ_INSTS_VALU per field from profiles/pmc_valu_config2.json; "
                               "launch time from one non-overlapped step after the timed steps"}
    for k in kernels:
        if k in alone:
            kernels[k]["ms_per_field_alone"] = round(float(np.mean(alone[k])) / a.batch, 4)
    r0 = stats.pop("realization")
    fp64_fps, field64 = (None, None) if a.no_fp64 else run_gpu_fp64_single(3, r0)
    extra = {}
    if not a.no_legs:
        import torch
        torch.cuda.empty_cache()  # the child processes take the device memory
        for name, limit in (("cg_vs_learned_pcg", 420), ("dctdst_1024", 420),
                            ("uneven_batch_2d", 300)):
            extra[name] = run_leg_child(name, limit)
    cpu = None if a.no_cpu else run_cpu_baseline(r0, {"fp32_batched": field32,
                                                      "fp64_per_realization": field64},
                                                 single_thread=not a.no_cpu_single)
    accuracy = None if cpu is None else cpu.pop("accuracy_vs_oracle")
    line = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (device Philox4x32-10 noise, seed 7)",
        "config": {
            "workload": "3D 256^3 cells on [0,1]^3, 512^3 embedding, Matern nu=1.5 (matern32), "
                        "variance 2, axiparallel l=(0.1,0.1,0.05), DCT matrix FP64 eigenvalues, "
                        "R2C field FP32, logNormal, batch of 64 realizations per step",
            "batch": a.batch,
            "cells": CELLS,
            "embedding": [2 * c for c in CELLS],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved / 1e9, 1),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4),
            "traffic": traffic,
        },
        "valu_roofline": valu_roof,
        "pipeline_alg_GBps": round(total_b * fps / 1e9, 1),
        "pipeline_frac_of_hbm": round(total_b * fps / HBM_PEAK, 4),
        "alg_bytes_per_field": int(total_b),
        "kernels": kernels,
        "fp64_per_realization_fields_per_s": None if fp64_fps is None else round(fp64_fps, 2),
        "fp32_batched_over_fp64_single": None if fp64_fps is None else round(fps / fp64_fps, 2),
        "eig_setup_s": round(eig_s, 3),
        "eig_stats": st,
        "pcie_inclusive_fields_per_s": stats.pop("pcie_inclusive_fields_per_s"),
        "tuning": stats.pop("tuning"),
        "field_log_stats": stats,
        "cpu_baseline": cpu,
        "accuracy_vs_fp64_oracle": None if accuracy is None else dict(
            accuracy, realization=r0,
            note="realization 0 of the last timed batched step (Philox realization r0) as a "
                 "Gaussian field (log of its log-normal output) and the FP64 per-realization "
                 "path's field of the same realization, against the FP64 CPU oracle on "
                 "identical Philox noise; bars: FP32 max_abs_over_std <= 2e-5, FP64 "
                 "max_abs_over_max <= 1e-11"),
        "other_configs": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
