#!/usr/bin/env python3
"""RK3 cell-updates/s on one MI355X: plain FP64 against the mixed-precision
and batched forms, with their rel-L2 error against FP64.

Workload of `value` (default): 2048 x 2048 periodic box [0, 2pi)^2, 8 fields,
FP64, seeded broadband Alfvenic turbulence on rho = p = 1, B0 = (1, 0, 0)
(1 <= |k| <= 16, |k|^(-5/3), rms 0.1, dB from A_z), gamma = 5/3, cfl = 0.3,
Hall CFL term with d_i = 0.02, 200 RK3 steps.  A step is one full RK3 step
(3 stages with derivatives, RHS, update, x and y filters, and the CFL dt).

Beside it (object "forms"): the mixed form and FP32 on the same workload;
the batched ensemble of 64 linear waves on 512 x 512 grids with d_i drawn per
instance; and the uneven ensemble of 96 turbulence runs on four grid sizes;
each in FP64, mixed and FP32, with its speed-up over FP64 and its per-field
rel-L2 error against an FP64 run of the same steps.

    python bench.py [--gpus 1] [--steps 200] [--warmup 5] [--no-cpu] [--no-forms]

Prints ONE JSON line.  The state is resident in HBM before the timed region;
time is taken with hipEvents on the stream the kernels run on, bracketed by
device synchronisation on both sides.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, B/s
# algorithmic bytes per cell-update of the canonical 2-sweep schedule
# (70 words per stage, 3 stages): all FP64, all FP32, and the mixed form
# (32 state words per stage at 8 B, 38 derivative / R_prev words at 4 B)
B_ALG = {"fp64": 1680.0, "fp32": 840.0, "mixed": 3 * (32 * 8 + 38 * 4)}


def kernel_alg_bytes(nx, ny, elem=8, delem=None, batch=1, divb=False):
    """Algorithmic HBM bytes per launch of each kernel class (inputs read once,
    outputs written once), per cell times the cells of one launch.  elem:
    bytes of a state (and E) value, delem: of a derivative / R_prev value.
    divb: dBx/dx and dBy/dy are formed (div B cleaning or an open x
    boundary); the ideal periodic RHS reads neither."""
    n = nx * ny * batch
    s, d = elem, (elem if delem is None else delem)
    k = 1 if divb else 0
    # x chunk of the solves (rows per thread): two carries per chunk and field
    cr = 18.0 / min(32, max(8, nx // 64))
    return {
        # y-sweep: read u (8), write filtered u (8), Ey and Ez (2), dudy (7,
        # all but dBy/dy, which the kernel stores all the same: it comes out
        # of the (Bx, By) transform pair), dEx/dy and dEz/dy (2)
        "ysweep": n * (18 * s + (9 + k) * d),
        # x-derivative: read u (7, all but Bx), Ey and Ez (2), write their
        # d/dx (9)
        "xderiv": n * ((9 + k) * s + (9 + k) * d),
        # RHS + RK, average of the three stages: read u (8) and dudx, dudy
        # (18), R_prev in stages 2-3 (8); write u (8), R_prev in stages 1-2 (8)
        "rhs_rk": n * (16 * s + (18 + 2 * k) * d + (16.0 / 3.0) * 2 * d),
        # x-filter in place: read 8, write 8
        "xfilter": n * 16 * s,
        # fused path: the carry solve reads the 9 fields the x-derivative
        # reads and writes 2 values per chunk and field; the chunk kernel
        # reads u (8), Ey and Ez (2), dudy (9), R_prev in stages 2-3 (8) and
        # the carries, writes u (8) and R_prev in stages 1-2 (8)
        "xcarry": n * (9 * s + cr * d),
        "xrhs_rk": n * (18 * s + 9 * d + cr * d + (16.0 / 3.0) * 2 * d),
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


LSCPU_KEYS = ("Model name", "Vendor ID", "CPU family", "Model", "CPU(s)", "Socket(s)",
              "Core(s) per socket", "Thread(s) per core", "CPU max MHz", "L3 cache")


def _lscpu():
    """the host CPU as lscpu reports it (the fields in LSCPU_KEYS), or None"""
    try:
        r = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20)
    except (OSError, subprocess.SubprocessError):
        return None
    if r.returncode != 0:
        return None
    out = {}
    for line in r.stdout.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in LSCPU_KEYS:
            out[k.strip()] = v.strip()
    return out or None


def cpu_baseline(nx, ny, u0, hall, di, c1_steps=200, c1_steps_1t=100):
    """The FP64 C restatement (the checker) timed on this host's cores:

    * the Orszag-Tang workload (256 x 256, Hall CFL term with d_i = 0.05) with
      OpenMP over independent lines and cells, n threads, 200 RK3 steps, and
      single-threaded for a bounded 100 steps (the same code; results equal
      bit for bit);
    * one RK3 step of the timed GPU workload (nx x ny turbulence), n threads.
    n = OMP_NUM_THREADS (at most 16, this host's share), else the CPU count."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    n = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    n = max(1, min(n, 16))
    import mhd2dcfdfftw_open_amd as m

    def timed(lib_threads, gx, gy, state, steps, **kw):
        orc.lib().orc_set_threads(lib_threads)
        s = orc.Sim(gx, gy, **kw)
        s.set_state(state)
        t0 = time.perf_counter()
        s.run(steps)
        dt = time.perf_counter() - t0
        del s
        return gx * gy * steps / dt, dt

    ot = m.orszag_tang(256, 256)
    with orc.threaded():
        v1, t1 = timed(1, 256, 256, ot, c1_steps_1t, hall=1, di=0.05)
        vn, tn = timed(n, 256, 256, ot, c1_steps, hall=1, di=0.05)
        vb, tb = timed(n, nx, ny, u0, 1, hall=hall, di=di)
        threads = orc.lib().orc_threads()
    return {"value": vn, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"256x256 Orszag-Tang (d_i = 0.05 Hall CFL), {c1_steps} RK3 steps of the "
                      f"FP64 oracle restatement, OpenMP over lines and cells, {threads} threads "
                      f"({tn:.1f} s)",
            "cpu_model": _cpu_model(), "lscpu": _lscpu(), "cores_visible": os.cpu_count(),
            "threads_1": {"value": v1, "steps": c1_steps_1t, "seconds": t1,
                          "workload": "256x256 Orszag-Tang, 1 thread"},
            "threads_n": {"value": vn, "threads": threads, "steps": c1_steps, "seconds": tn,
                          "workload": "256x256 Orszag-Tang"},
            "gpu_workload_1_step": {"value": vb, "threads": threads, "seconds": tb,
                                    "workload": f"{nx}x{ny} turbulence, 1 RK3 step"}}


COPY_BYTES = 1 << 30
# every context first runs the x-solve shape autotuner (mhd2d_autotune)
AUTOTUNE = True
# fused x-derivative + RHS: None = the context's own choice (the autotuner's
# where it runs), 0 / 1 = forced where it applies
XFUSE = None


def _apply_xfuse(ctx):
    if XFUSE is not None:
        try:
            ctx.set_xfuse(XFUSE)
        except Exception:
            pass  # the path does not apply to this context


def _pmc_rows(path):
    """rows of a rocprofv3 counter_collection.csv: (kernel name, counter,
    value), of each kernel's dispatches with its largest grid only (the
    one-field x-solves of the diagnostics share the step's kernel)"""
    import csv
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r.get("Kernel_Name") or r.get("Kernel-Name") or ""
            ctr = r.get("Counter_Name") or r.get("Counter-Name") or ""
            val = r.get("Counter_Value") or r.get("Counter-Value") or "0"
            grid = int(float(r.get("Grid_Size") or r.get("Grid-Size") or 0))
            rows.append((name, ctr, float(val), grid))
    gmax = {}
    for name, _, _, grid in rows:
        gmax[name] = max(gmax.get(name, 0), grid)
    return [(name, ctr, val) for name, ctr, val, grid in rows if grid == gmax[name]]


PMC_HEADER = """# HBM traffic per launch from rocprofv3 PMC counters, {nx}x{nx} FP64 turbulence
# commands (GPU host, repo root $REPO; cd /tmp && export TMPDIR=/tmp first):
#   rocprofv3 --pmc FETCH_SIZE --output-format csv -d $REPO/gpurun_out/pmc_fetch -o run -- python3 $REPO/bench.py --steps 2 --warmup 1 --no-cpu --no-forms --profile-steps 1 --graph 0 --no-autotune --xfuse 1
#   rocprofv3 --pmc WRITE_SIZE --output-format csv -d $REPO/gpurun_out/pmc_write -o run -- python3 $REPO/bench.py --steps 2 --warmup 1 --no-cpu --no-forms --profile-steps 1 --graph 0 --no-autotune --xfuse 1
#   rocprofv3 --pmc FETCH_SIZE --output-format csv -d $REPO/gpurun_out/pmc_copy_fetch -o run -- python3 $REPO/bench.py --copy-probe
#   rocprofv3 --pmc WRITE_SIZE --output-format csv -d $REPO/gpurun_out/pmc_copy_write -o run -- python3 $REPO/bench.py --copy-probe
#   the first two again with --xfuse 0 into pmc_fetch_unfused, pmc_write_unfused
#   python bench.py --summarize-pmc <the four run_counter_collection.csv> <the two unfused ones>
# calibration: the 1 GiB copy kernel's counters scale FETCH by {fs:.3f}, WRITE by {ws:.3f}
# per launch: median over a kernel's full-size dispatches; the RHS + RK and
# the fused chunk kernel: mean over whole steps (their R_prev traffic
# differs by stage).  Fused path (the default in FP64): ysweep, xcarry,
# xrhs_rk, xfilter (its chunk edge rows from the fused kernel's buffer);
# unfused path: xderiv, rhs_rk, xfilter_unfused
"""


def summarize_pmc(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, nx, ny,
                  fetch_unfused=None, write_unfused=None):
    """HBM bytes per launch of each kernel class from two --pmc passes
    (FETCH_SIZE, WRITE_SIZE in KiB per dispatch), scaled by what the same
    counters read for the copy kernel of known size (COPY_BYTES read and
    COPY_BYTES written per launch); the fused path's classes from the first
    pair of passes, the unfused path's from the optional second pair."""
    def per_kernel(path, counter):
        acc = {}
        for name, ctr, val in _pmc_rows(path):
            if ctr != counter:
                continue
            acc.setdefault(name, []).append(val * 1024.0)
        # median over dispatches: drops the odd launches of the same kernel
        # (first-stage y-sweep without filter); the RHS + RK moves different
        # bytes per stage (R_prev is not read in stage 1 nor written in
        # stage 3), so it takes the mean over its launches, whole steps of
        # three each
        out = {}
        for k, v in acc.items():
            if "k_rhs_rk" in k or "k_xrhs_rk" in k:
                v = v[len(v) % 3:]
                out[k] = float(np.mean(v))
            else:
                out[k] = float(np.median(v))
        return out

    cf = per_kernel(copy_fetch_csv, "FETCH_SIZE")
    cw = per_kernel(copy_write_csv, "WRITE_SIZE")
    ck = [k for k in cf if "k_copy" in k]
    fscale = COPY_BYTES / cf[ck[0]] if ck and cf[ck[0]] > 0 else 1.0
    ck = [k for k in cw if "k_copy" in k]
    wscale = COPY_BYTES / cw[ck[0]] if ck and cw[ck[0]] > 0 else 1.0
    alg = kernel_alg_bytes(nx, ny)
    alg["xfilter_unfused"] = alg["xfilter"]
    # class: (kernel name part, k_xsolve KIND or None); KIND 0 derivative,
    # 2 filter, 3 carry solve, 4 filter with edge rows
    fused = {"ysweep": ("k_ysweep", None), "xcarry": ("k_xsolve", 3),
             "xrhs_rk": ("k_xrhs_rk", None), "xfilter": ("k_xsolve", 4)}
    unfused = {"xderiv": ("k_xsolve", 0), "rhs_rk": ("k_rhs_rk", None),
               "xfilter_unfused": ("k_xsolve", 2)}
    passes = [(fused, fetch_csv, write_csv)]
    if fetch_unfused and write_unfused:
        passes.append((unfused, fetch_unfused, write_unfused))
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    rows = []
    for classes, fcsv, wcsv in passes:
        fetch = per_kernel(fcsv, "FETCH_SIZE")
        write = per_kernel(wcsv, "WRITE_SIZE")
        for cls, (key, kind) in classes.items():
            names = [k for k in fetch if key + "<" in k]
            if kind is not None:
                names = [k for k in names if f", {kind}, false>(" in k]
            if not names:
                continue
            k = names[0]
            rd = fetch[k] * fscale
            wr = write.get(k, 0.0) * wscale
            rec = {"kernel": k, "workload": f"{nx}x{ny} FP64 turbulence",
                   "fetch_bytes_per_launch": rd, "write_bytes_per_launch": wr,
                   "bytes_per_launch": rd + wr, "alg_bytes_per_launch": alg[cls],
                   "traffic_over_alg": (rd + wr) / alg[cls],
                   "copy_scale": {"fetch": fscale, "write": wscale}}
            with open(os.path.join(ROOT, "profiles", f"traffic_{cls}_{nx}.json"), "w") as f:
                json.dump(rec, f, indent=1)
            print(json.dumps(rec))
            rows.append(f"{cls} | {rd:.3e} | {wr:.3e} | {rd + wr:.3e} | {alg[cls]:.3e} | "
                        f"{(rd + wr) / alg[cls]:.3f}")
    with open(os.path.join(ROOT, "profiles", f"traffic_fp64_{nx}.txt"), "w") as f:
        f.write(PMC_HEADER.format(nx=nx, fs=fscale, ws=wscale))
        f.write("# kernel | fetch B/launch | write B/launch | total | algorithmic | total/alg\n")
        f.write("\n".join(rows) + "\n")


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary, if one exists (profiles/traffic_<kernel>_2048.json)."""
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel}_2048.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def _precision(m, name):
    return {"fp64": m.FP64, "fp32": m.FP32, "mixed": m.MIXED}[name]


def _bytes(name):
    return {"fp64": (8, 8), "fp32": (4, 4), "mixed": (8, 4)}[name]


def run_single(m, u0, nx, ny, prec, steps, warmup, graph, hall, di, **dissipation):
    """(cell-updates/s, ms, context) of one run; the context holds the state
    after warmup + steps"""
    ctx = m.Context(nx, ny, hall=hall, di=di, precision=_precision(m, prec), **dissipation)
    ctx.set_state(u0)
    ctx.x_shapes = ctx.autotune() if AUTOTUNE else None
    _apply_xfuse(ctx)
    ctx.run(warmup, graph=graph)
    ctx.sync()
    ms = ctx.time_run(steps, graph=graph)
    ctx.sync()
    return nx * ny * steps / (ms * 1e-3), ms, ctx


def ensemble_states(m, nx, ny, count, seed):
    """64-instance ensemble: rho = p = 1, B0 = (1, 0, 0) with one seeded
    linear wave (integer kx, ky in [1, 8], amplitude 1e-4, random phase) per
    instance; d_i uniform in [0, 0.2]"""
    rng = np.random.default_rng(seed)
    dis = rng.uniform(0.0, 0.2, count)
    return [m.linear_wave(nx, ny, seed=seed + b, amp=1e-4) for b in range(count)], dis


def uneven_runs(m, count, seed):
    """uneven ensemble: grid sizes drawn from {256x256, 512x256, 1024x1024,
    2048x512}, turbulence with rms ~ U[0.05, 0.2], d_i ~ U[0, 0.1]"""
    rng = np.random.default_rng(seed)
    sizes = [(256, 256), (512, 256), (1024, 1024), (2048, 512)]
    runs, states = [], []
    for k in range(count):
        nx, ny = sizes[rng.integers(0, 4)]
        runs.append(dict(nx=nx, ny=ny, hall=1, di=float(rng.uniform(0, 0.1))))
        states.append(m.turbulence(nx, ny, seed=seed + k, rms=float(rng.uniform(0.05, 0.2)),
                                   kmax=16))
    return runs, states


def run_batch(m, states, dis, nx, ny, prec, steps, warmup, hall=1, graph=True):
    B = len(states)
    ctx = m.Context(nx, ny, hall=hall, precision=_precision(m, prec), batch=B)
    for b in range(B):
        ctx.set_di(b, float(dis[b]))
        ctx.set_state(states[b], inst=b)
    ctx.x_shapes = ctx.autotune() if AUTOTUNE else None
    ctx.run(warmup, graph=graph)
    ctx.sync()
    ms = ctx.time_run(steps, graph=graph)
    ctx.sync()
    return B * nx * ny * steps / (ms * 1e-3), ms, ctx


def advance_to(obj, done, target, reset, graph=True):
    """step obj (a context or an ensemble) until `target` steps have been
    taken since its initial state; `reset` restores that state when more
    than `target` steps are done already"""
    if done > target:
        reset()
        done = 0
    if target > done:
        obj.run(target - done, graph=graph)
    obj.sync()
    return target


# steps after which each workload's error against FP64 is taken: the
# 2048^2 and uneven-ensemble runs are 200 steps, the 64 x 512^2 batch 500
ERR_STEPS = {"single": 200, "batched": 500, "uneven": 200}


def field_errors(a, b):
    """per-field rel-L2 of a against b (fields of b below 1e-6 of its rms
    measured in absolute L2 over that rms)"""
    scale = np.sqrt(np.mean(b ** 2)) or 1.0
    out = []
    for v in range(b.shape[0]):
        nb = np.sqrt(np.mean(b[v] ** 2))
        d = np.sqrt(np.mean((a[v] - b[v]) ** 2))
        out.append(float(d / scale if nb < 1e-6 * scale else d / nb))
    return out


def sweep_alg_bytes(nx, ny, elem=8, delem=None, batch=1):
    """Algorithmic bytes per stage of the canonical two-sweep schedule, per
    cell times the cells of one launch: the y-sweep reads u (8 words) and
    writes the filtered u (8), d/dy of u (8) and of E (3) = 27 words; the
    x-sweep reads u (8), the y-derivatives (11) and R_prev (8) and writes u
    (8) and R_prev (8) = 43 words.  State words count elem bytes, derivative
    and R_prev words delem bytes."""
    n = nx * ny * batch
    s, d = elem, (elem if delem is None else delem)
    return {"ysweep": n * (16 * s + 11 * d), "xsweep": n * (16 * s + 27 * d)}


# kernels of each canonical sweep as launched here: the y-sweep is one
# kernel; the x-sweep is the derivative x-solve, the RHS + RK update and the
# filter x-solve (three kernels, since full x-lines of all fields do not fit
# on one CU), or with the fused x-derivative + RHS the carry solve, the fused
# chunk kernel and the filter x-solve
SWEEP_KERNELS = {"ysweep": ("ysweep",), "xsweep": ("xderiv", "rhs_rk", "xfilter")}
# traffic (PMC) class of each timed class where the two differ
TRAFFIC_CLASS = {("xfilter", False): "xfilter_unfused"}
SWEEP_KERNELS_FUSED = {"ysweep": ("ysweep",), "xsweep": ("xcarry", "xrhs_rk", "xfilter")}


def roofline_of(ctx, nx, ny, elem, delem, batch, profile_steps, copy_gbps, traffic=None):
    """roofline of the dominant canonical sweep (the larger device time per
    step): its algorithmic bytes over its average time per stage, measured
    with hipEvents around every launch on the kernels' stream.  The largest
    single kernel against its own reads and writes is reported beside it
    (kernel_io_frac).  traffic: HBM bytes per launch of each kernel class
    from the committed PMC summary (or None)."""
    prof = ctx.profile_run(profile_steps)
    ctx.sync()
    alg = kernel_alg_bytes(nx, ny, elem, delem, batch)
    salg = sweep_alg_bytes(nx, ny, elem, delem, batch)
    sk = SWEEP_KERNELS_FUSED if "xcarry" in prof else SWEEP_KERNELS
    # time per stage of each sweep: the sum of its kernels' average launches
    st_ms = {sw: sum(prof[k][0] for k in ks) for sw, ks in sk.items()}
    dom = "ysweep" if st_ms["ysweep"] >= st_ms["xsweep"] else "xsweep"
    achieved = salg[dom] / (st_ms[dom] * 1e-3)
    roof = {"bound": "hbm", "kernel": dom, "kernels": list(sk[dom]),
            "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "avg_ms_per_stage": st_ms[dom],
            "alg_bytes_per_stage": salg[dom],
            "frac_vs_copy_probe": achieved / (copy_gbps * 1e9), "traffic": None}
    tk = [TRAFFIC_CLASS.get((k, sk is SWEEP_KERNELS_FUSED), k) for k in sk[dom]]
    if traffic is not None and all(k in traffic for k in tk):
        roof["traffic"] = sum(traffic[k] for k in tk)
    roof["by_sweep"] = {sw: {"avg_ms_per_stage": st_ms[sw], "alg_bytes_per_stage": salg[sw],
                             "GB/s": salg[sw] / (st_ms[sw] * 1e-3) / 1e9,
                             "frac": salg[sw] / (st_ms[sw] * 1e-3) / HBM_PEAK}
                        for sw in sk}
    share = {k: prof[k][0] * prof[k][1] for k in alg if k in prof}
    big = max(share, key=share.get)
    kio = alg[big] / (prof[big][0] * 1e-3)
    roof["kernel_io_frac"] = {"kernel": big, "avg_launch_ms": prof[big][0],
                              "io_bytes_per_launch": alg[big], "GB/s": kio / 1e9,
                              "frac": kio / HBM_PEAK}
    kernels = {k: {"avg_ms": prof[k][0], "launches_per_step": prof[k][1] / profile_steps,
                   "GB/s": (alg[k] / (prof[k][0] * 1e-3) / 1e9) if k in alg and prof[k][0] > 0
                   else None}
               for k in prof}
    return roof, kernels


def c3_forms(m, value2048, profile_steps, copy_gbps, hall, di, n=4096, chunk=50, chunks=4,
             graph=True, only=None):
    """4096^2: mixed precision (FP64 state) and plain FP32 (FP32 state), each
    with the x-derivative as the tridiagonal solve and as the dense A^-1 B
    through split-BF16 MFMA GEMMs (3 and 6 products); per-field rel-L2
    against FP64 every 50 steps for 200 steps.  Each variant is timed over
    its 4 x 50 steps (hipEvents).  only: the variants to run (by name);
    errors and speed-ups against FP64 need "fp64" among them."""
    u0 = m.turbulence(n, n, seed=42, rms=0.1, kmin=1, kmax=16)
    variants = [("fp64", m.FP64, 0), ("mixed", m.MIXED, 0), ("mixed_gemm6", m.MIXED, 6),
                ("mixed_gemm3", m.MIXED, 3), ("fp32", m.FP32, 0), ("fp32_gemm6", m.FP32, 6),
                ("fp32_gemm3", m.FP32, 3)]
    if only:
        variants = [x for x in variants if x[0] in only]
    ref_states, out = [], {}
    ref_value = None
    for name, prec, xg in variants:
        ctx = m.Context(n, n, hall=hall, di=di, precision=prec, xgemm=xg)
        ctx.set_state(u0)
        if AUTOTUNE and not xg:
            ctx.autotune()
        ctx.run(1, graph=graph)  # the first step (no pending filter) and the capture
        ctx.sync()
        ctx.set_state(u0)
        ms, errs = 0.0, []
        for k in range(chunks):
            ms += ctx.time_run(chunk, graph=graph)
            ctx.sync()
            st = ctx.state()
            if name == "fp64":
                ref_states.append(st)
            elif ref_states:
                errs.append(max(field_errors(st, ref_states[k])))
        v = n * n * chunk * chunks / (ms * 1e-3)
        if name == "fp64":
            ref_value = v
        pe, pd = (8, 8) if prec == m.FP64 else ((4, 4) if prec == m.FP32 else (8, 4))
        roof, kern = roofline_of(ctx, n, n, pe, pd, 1, profile_steps, copy_gbps)
        rec = {"workload": f"turbulence {n}x{n}, {name}", "cell_updates_per_s": v,
               "ms_per_step": ms / (chunk * chunks), "speedup_vs_fp64_2048": v / value2048,
               "speedup_vs_fp64_4096": v / ref_value if ref_value else None,
               "kernels": kern, "roofline": roof,
               "step_roofline_frac": v * B_ALG["fp64" if prec == m.FP64 else
                                                ("fp32" if prec == m.FP32 else "mixed")] / HBM_PEAK}
        if errs:
            rec[f"rel_l2_vs_fp64_every_{chunk}"] = errs
            rec["rel_l2_vs_fp64_max"] = max(errs)
            rec["meets_1e-6"] = bool(max(errs) <= 1e-6)
        if xg:
            # MFMA work of one x-derivative launch (9 fields: the ideal
            # periodic RHS needs no dBx/dx) over its time
            flops = ctx.xgemm_panels() * 64 * 32 * 2 * n * 9 * xg
            t = kern["xderiv"]["avg_ms"] * 1e-3
            rec["xderiv_mfma"] = {"bf16_flops_per_launch": flops, "TFLOP/s": flops / t / 1e12,
                                  "frac_of_2500_TFLOPs": flops / t / 2.5e15,
                                  "panels": ctx.xgemm_panels(),
                                  "dense_fraction": ctx.xgemm_panels() / ((n / 64) * (n / 32))}
        out[f"{name}_{n}"] = rec
        ctx.close()
    return out


def step_traffic_run(m, prec, steps, warmup):
    """the 64 x 512^2 batch of linear waves in one precision, whole batch
    per launch, default launch shapes, eager steps only: the program that
    rocprofv3 --pmc runs to count HBM bytes per step (summarize with
    --summarize-step-pmc)"""
    enx = eny = 512
    states, dis = ensemble_states(m, enx, eny, 64, 42)
    ctx = m.Context(enx, eny, hall=1, precision=_precision(m, prec), batch=len(states))
    for b in range(len(states)):
        ctx.set_di(b, float(dis[b]))
        ctx.set_state(states[b], inst=b)
    ctx.run(warmup, graph=False)
    ctx.sync()
    ms = ctx.time_run(steps, graph=False)
    ctx.sync()
    out = {"workload": f"64 x {enx}x{eny} linear waves, {prec}, whole batch per launch, "
                       "default launch shapes, eager steps",
           "precision": prec, "steps_run": warmup + steps, "timed_steps": steps,
           "ms_per_step": ms / steps,
           "cell_updates_per_s": len(states) * enx * eny * steps / (ms * 1e-3),
           "x_fused": ctx.xfuse}
    ctx.close()
    return out


# kernels of an RK3 step (set-up and diagnostics kernels excluded)
STEP_KERNELS = ("k_ysweep", "k_xsolve", "k_xrhs_rk", "k_rhs_rk", "k_dt_finalize")


def summarize_step_pmc(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, steps):
    """HBM bytes per RK3 step: FETCH_SIZE and WRITE_SIZE (KiB per dispatch)
    summed over every dispatch of the step kernels and divided by the steps
    the program ran, scaled by what the counters read for the copy kernel of
    known size"""
    import csv

    def totals(path, counter):
        per, copy = 0.0, []
        with open(path) as f:
            for r in csv.DictReader(f):
                name = r.get("Kernel_Name") or r.get("Kernel-Name") or ""
                ctr = r.get("Counter_Name") or r.get("Counter-Name") or ""
                if ctr != counter:
                    continue
                val = float(r.get("Counter_Value") or r.get("Counter-Value") or 0) * 1024.0
                if "k_copy" in name:
                    copy.append(val)
                elif any(k in name for k in STEP_KERNELS):
                    per += val
        return per, copy

    fetch, _ = totals(fetch_csv, "FETCH_SIZE")
    write, _ = totals(write_csv, "WRITE_SIZE")
    _, cf = totals(copy_fetch_csv, "FETCH_SIZE")
    _, cw = totals(copy_write_csv, "WRITE_SIZE")
    fscale = COPY_BYTES / float(np.median(cf)) if cf else 1.0
    wscale = COPY_BYTES / float(np.median(cw)) if cw else 1.0
    rd, wr = fetch * fscale / steps, write * wscale / steps
    return {"fetch_bytes_per_step": rd, "write_bytes_per_step": wr,
            "bytes_per_step": rd + wr, "fetch_scale": fscale, "write_scale": wscale}


def dropin_form(m, u0, nx, ny, hall, di, value, steps=10, warmup=2):
    """the same workload through the reference's own entry points, called in
    mhd.c's order (mhd.c:154-206) as its main loop would: one kernel per
    entry point, host copies only on request (MHD2D_SYNC_EXPLICIT); each
    step ends in calcDt, which hands dt back to the host"""
    import time
    L = m.lib()
    G = m.RefGlobals(nx, ny, hall=hall, di=di)
    try:
        L.mhd2d_dropin_set_sync(1)
        G.array("uu")[:8] = u0
        P = G.ptr
        nvar = G.nvar

        def step(dt):
            for irk in range(3):
                L.sendrecvHaloPointsX()
                L.derivX(nvar, P("uu"), P("uuHaloX0"), P("uuHaloX1"), P("dudx"), 0, 0)
                L.derivY(nvar, P("uu"), P("dudy"), 0)
                if hall == 1:   # d2B/dx2, d2B/dy2 (mhd.c:167-172, read only with eta > 0)
                    L.derivXX(3, P("uu"), P("uuHaloX0"), P("uuHaloX1"), P("d2Bdx2"), 4, 4)
                    L.derivYY(3, P("uu"), P("d2Bdy2"), 4)
                L.calcEfield()
                L.sendrecvHaloPointsEfieldX()
                L.derivX(3, P("Efield"), P("EfieldHaloX0"), P("EfieldHaloX1"), P("dEdx"), 0, 0)
                L.derivY(3, P("Efield"), P("dEdy"), 0)
                L.calcRHS()
                L.rkExecute(irk)
                L.sendrecvHaloPointsX()
                L.filtering()
            dt = L.calcDt(dt)
            L.rkInit(dt)
            return dt

        dt = L.calcDt(0.0)
        L.rkInit(dt)
        for _ in range(warmup):
            dt = step(dt)
        t0 = time.perf_counter()
        for _ in range(steps):
            dt = step(dt)
        sec = time.perf_counter() - t0
        v = nx * ny * steps / sec
        return {"workload": f"turbulence {nx}x{ny}, fp64, the reference's entry points in "
                            "mhd.c's order (drop-in layer, host copies on request), "
                            "incl. the d2B derivatives mhd.c computes with Hall = 1",
                "cell_updates_per_s": v, "ms_per_step": sec * 1e3 / steps,
                "timed_steps": steps, "fused_engine_vs_dropin": value / v}
    finally:
        L.mhd2d_dropin_set_sync(0)
        L.mhd2d_dropin_release()
        G.close()


def batched_forms(m, value, args, copy_gbps):
    """the 64 x 512^2 batched ensemble in the three precisions (errors
    against the FP64 batch on four instances); value: the 2048^2 FP64 rate"""
    out = {}
    # batched ensemble: 64 x 512^2 linear waves, d_i per instance
    enx = eny = 512
    states, dis = ensemble_states(m, enx, eny, 64, 42)
    res = {}
    # timed steps: the warm-up plus the timed steps make the 500 of the
    # workload, after which the errors are taken
    tsteps = max(1, args.ens_steps - args.warmup)
    target = ERR_STEPS["batched"]
    for prec in ("fp64", "mixed", "fp32"):
        v, ms_p, c = run_batch(m, states, dis, enx, eny, prec, tsteps, args.warmup,
                               graph=bool(args.graph))

        def reset(c=c):
            for b in range(len(states)):
                c.set_state(states[b], inst=b)
        advance_to(c, args.warmup + tsteps, target, reset, graph=bool(args.graph))
        res[prec] = (v, ms_p, c)
    vref, _, cref = res["fp64"]
    checked = (0, 21, 42, 63)
    # errors first: profiling advances the runs
    ref_states = {b: cref.state(b) for b in checked}
    errs_of = {prec: [max(field_errors(res[prec][2].state(b), ref_states[b]))
                      for b in checked] for prec in res}
    for prec in ("fp64", "mixed", "fp32"):
        v, ms_p, c = res[prec]
        errs = errs_of[prec]
        pe, pd = _bytes(prec)
        roof, kern = roofline_of(c, enx, eny, pe, pd, 64, args.profile_steps, copy_gbps)
        out[f"batched64x512_{prec}"] = {
            "workload": f"64 x {enx}x{eny} linear waves, d_i ~ U[0, 0.2], {prec}, "
                        "one launch per kernel, stage and instance group, "
                        + ("hipGraph step" if args.graph else "eager launches"),
            "instances_per_launch": c.subbatch,
            "cell_updates_per_s": v, "ms_per_step": ms_p / tsteps, "timed_steps": tsteps,
            "speedup_vs_fp64_2048": v / value, "speedup_vs_fp64_batched": v / vref,
            "rel_l2_vs_fp64_max": max(errs), "after_steps": target,
            "instances_checked": list(checked), "meets_1e-6": bool(max(errs) <= 1e-6),
            "step_roofline_frac": v * B_ALG[prec] / HBM_PEAK, "roofline": roof}
    for prec in res:
        res[prec][2].close()
    return out


def uneven_forms(m, value, args):
    """the uneven ensemble of 96 turbulence runs on four grid sizes in the
    three precisions (errors against the FP64 ensemble on four runs);
    value: the 2048^2 FP64 rate"""
    out = {}
    graph = bool(args.graph)
    # uneven ensemble: 96 runs, sizes drawn from four shapes, per-run rms, d_i
    runs, ustates = uneven_runs(m, 96, 42)
    ures = {}
    for prec in ("fp64", "mixed", "fp32"):
        E = m.Ensemble(runs, precision=_precision(m, prec))
        for k in range(len(runs)):
            E.set_state(k, ustates[k])
        if AUTOTUNE:
            E.autotune()
        E.run(args.warmup, graph=graph)
        E.sync()
        # both launch modes, timed over the same steps from the same state;
        # the autotuned one is the reported rate
        chosen = E.launch_mode
        by_mode = {}
        for mode in (m.ENS_ONE_LAUNCH, m.ENS_PER_GROUP):
            try:
                E.set_launch(mode)
            except m.MHD2DError:
                continue
            for k in range(len(runs)):
                E.set_state(k, ustates[k])
            E.run(args.warmup, graph=graph)
            E.sync()
            t = E.time_run(args.steps, graph=graph)
            E.sync()
            by_mode["one_launch" if mode == m.ENS_ONE_LAUNCH else "per_group"] = {
                "ms_per_step": t / args.steps, "launches_per_step": E.launches_per_step}
        E.set_launch(chosen)
        for k in range(len(runs)):
            E.set_state(k, ustates[k])
        E.run(args.warmup, graph=graph)
        E.sync()
        ms_p = E.time_run(args.steps, graph=graph)
        E.sync()

        def reset(E=E):
            for k in range(len(runs)):
                E.set_state(k, ustates[k])
        advance_to(E, args.warmup + args.steps, ERR_STEPS["uneven"], reset, graph=graph)
        cells = sum(r["nx"] * r["ny"] for r in runs)
        E.bench_modes = (chosen, by_mode)
        ures[prec] = (cells * args.steps / (ms_p * 1e-3), ms_p, E)
    checked = (0, 31, 64, 95)
    refs = {k: ures["fp64"][2].state(k) for k in checked}
    for prec in ("fp64", "mixed", "fp32"):
        v, ms_p, E = ures[prec]
        errs = [max(field_errors(E.state(k), refs[k])) for k in checked]
        chosen, by_mode = E.bench_modes
        how = ("one launch per kernel class and stage over the four size groups"
               if chosen == m.ENS_ONE_LAUNCH else
               "one batched context per size, launches forked into the four group streams")
        out[f"uneven96_{prec}"] = {
            "workload": "96 turbulence runs, sizes drawn from 256x256, 512x256, "
                        "1024x1024, 2048x512, rms ~ U[0.05, 0.2], d_i ~ U[0, 0.1], "
                        f"{prec}; {how}" + ("; hipGraph step" if graph else "; eager launches"),
            "launch_mode": "one_launch" if chosen == m.ENS_ONE_LAUNCH else "per_group",
            "launches_per_step": E.launches_per_step, "by_launch_mode": by_mode,
            "cells": sum(r["nx"] * r["ny"] for r in runs), "groups": E.groups,
            "cell_updates_per_s": v, "ms_per_step": ms_p / args.steps,
            "speedup_vs_fp64_2048": v / value, "speedup_vs_fp64_uneven": v / ures["fp64"][0],
            "rel_l2_vs_fp64_max": max(errs), "after_steps": ERR_STEPS["uneven"],
            "timed_steps": args.steps,
            "instances_checked": list(checked), "meets_1e-6": bool(max(errs) <= 1e-6),
            "step_roofline_frac": v * B_ALG[prec] / HBM_PEAK}
    for prec in ures:
        ures[prec][2].close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--precision", choices=["fp64", "fp32", "mixed"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the mixed and batched measurements")
    ap.add_argument("--ens-steps", type=int, default=500,
                    help="steps of the 64 x 512^2 ensemble")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default x-solve launch shapes")
    ap.add_argument("--xfuse", choices=["auto", "0", "1"], default="auto",
                    help="fused x-derivative + RHS: the context's (autotuned) choice, or "
                         "forced off / on where it applies")
    ap.add_argument("--c3-only", action="store_true",
                    help="only the 4096^2 precision / x-derivative comparison (prints its JSON)")
    ap.add_argument("--c3-forms", default="",
                    help="with --c3-only: comma-separated variants (fp64, mixed, mixed_gemm6, "
                         "mixed_gemm3, fp32, fp32_gemm6, fp32_gemm3; default all)")
    ap.add_argument("--c3-chunk", type=int, default=50,
                    help="with --c3-only: steps between error checks (4 checks)")
    ap.add_argument("--c4-only", action="store_true",
                    help="only the 64 x 512^2 batched ensemble (prints its JSON; honours --graph)")
    ap.add_argument("--c5-only", action="store_true",
                    help="only the 96-run uneven ensemble (prints its JSON; honours --graph)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the streaming-copy calibration kernel and print its rate")
    ap.add_argument("--step-traffic", choices=["fp64", "fp32", "mixed"],
                    help="only run the 64 x 512^2 batch for --warmup + --steps eager steps "
                         "in this precision (the program for rocprofv3 --pmc byte counts "
                         "per step); prints its JSON")
    ap.add_argument("--summarize-step-pmc", nargs=4, metavar="CSV",
                    help="HBM bytes per step from the counter CSVs of a --step-traffic run "
                         "(FETCH, WRITE) and of --copy-probe (FETCH, WRITE); the program "
                         "ran --warmup + --steps steps")
    ap.add_argument("--summarize-pmc", nargs="+",
                    metavar="CSV",
                    help="turn rocprofv3 --pmc counter CSVs into profiles/traffic_*.json: "
                         "FETCH, WRITE (fused path), COPY_FETCH, COPY_WRITE, then "
                         "optionally FETCH, WRITE of the unfused path")
    args = ap.parse_args()
    global AUTOTUNE, XFUSE
    AUTOTUNE = not args.no_autotune
    XFUSE = None if args.xfuse == "auto" else int(args.xfuse)

    if args.summarize_step_pmc:
        c = args.summarize_step_pmc
        print(json.dumps(summarize_step_pmc(*c, steps=args.warmup + args.steps)))
        return

    if args.summarize_pmc:
        c = args.summarize_pmc
        if len(c) not in (4, 6):
            ap.error("--summarize-pmc takes 4 or 6 CSV files")
        summarize_pmc(*c[:4], nx=args.nx, ny=args.ny,
                      fetch_unfused=c[4] if len(c) == 6 else None,
                      write_unfused=c[5] if len(c) == 6 else None)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs the whole grid on one GPU; multi-GPU runs "
              "(--gpus > 1) are not part of it", file=sys.stderr)
        sys.exit(2)

    import mhd2dcfdfftw_open_amd as m

    if args.copy_probe:
        print(json.dumps({"copy_probe_bytes": COPY_BYTES,
                          "copy_gbps": m.copy_probe(COPY_BYTES, 5)}))
        return

    if args.step_traffic:
        print(json.dumps(step_traffic_run(m, args.step_traffic, args.steps, args.warmup)))
        return

    if args.c3_only:
        only = [f for f in args.c3_forms.split(",") if f] or None
        print(json.dumps(c3_forms(m, float("nan"), args.profile_steps, m.copy_probe(COPY_BYTES, 5),
                                  1, 0.02, chunk=args.c3_chunk, graph=bool(args.graph),
                                  only=only)))
        return

    if args.c5_only:
        print(json.dumps(uneven_forms(m, float("nan"), args)))
        return

    if args.c4_only:
        print(json.dumps(batched_forms(m, float("nan"), args, m.copy_probe(COPY_BYTES, 5))))
        return

    nx, ny = args.nx, args.ny
    hall, di = 1, 0.02
    graph = bool(args.graph)
    elem, delem = _bytes(args.precision)
    u0 = m.turbulence(nx, ny, seed=42, rms=0.1, kmin=1, kmax=16)
    copy_gbps = m.copy_probe(COPY_BYTES, 10)

    value, ms, ctx = run_single(m, u0, nx, ny, args.precision, args.steps, args.warmup, graph,
                                hall, di)
    traffic = None
    if args.precision == "fp64" and nx == 2048 and ny == 2048:
        traffic = {k: load_traffic(k)
                   for k in ("ysweep", "xderiv", "rhs_rk", "xcarry", "xrhs_rk", "xfilter",
                             "xfilter_unfused")}
        traffic = {k: v for k, v in traffic.items() if v is not None}
    roofline, kernels = roofline_of(ctx, nx, ny, elem, delem, 1, args.profile_steps, copy_gbps,
                                    traffic)
    energy, divb = ctx.diagnostics()
    # the state after warmup + steps (profile_run advanced it further: rerun)
    ref_state = None

    out = {
        "metric": "RK3 cell-updates/s: plain FP64 vs mixed/batched at rel-L2 <=1e-6 vs FP64",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "mixed": "f64 state, f32 derivatives"}[
            args.precision],
        "data": "synthetic (seeded turbulence initial state, seed 42)",
        "config": {"workload": f"hall-mhd2d turbulence {nx}x{ny}, 8 fields, "
                               f"{args.precision}, d_i={di}, {args.steps} timed RK3 steps "
                               f"after {args.warmup} warm-up steps",
                   "nx": nx, "ny": ny, "graph": graph, "precision": args.precision},
        "roofline": roofline,
        "step_roofline_frac": value * B_ALG[args.precision] / HBM_PEAK,
        "kernels": kernels,
        "x_solve_shapes": ctx.x_shapes,
        "x_fused": ctx.xfuse,
        "copy_probe_GBps": copy_gbps,
        "diagnostics": {"energy": energy, "max_divb": divb},
    }
    ctx.close()

    if not args.no_forms and args.precision == "fp64":
        forms = {}
        nsteps = ERR_STEPS["single"]
        # FP64 reference state after the workload's 200 steps, for the error
        # columns
        ref = m.Context(nx, ny, hall=hall, di=di, precision=m.FP64)
        ref.set_state(u0)
        advance_to(ref, 0, nsteps, None, graph=graph)
        ref_state = ref.state()
        ref.close()
        for prec in ("mixed", "fp32"):
            v, ms_p, c = run_single(m, u0, nx, ny, prec, args.steps, args.warmup, graph, hall,
                                    di)
            advance_to(c, args.warmup + args.steps, nsteps, lambda c=c: c.set_state(u0),
                       graph=graph)
            e = field_errors(c.state(), ref_state)
            pe, pd = _bytes(prec)
            roof, kern = roofline_of(c, nx, ny, pe, pd, 1, args.profile_steps, copy_gbps)
            forms[f"{prec}_{nx}"] = {
                "workload": f"turbulence {nx}x{ny}, {prec}", "cell_updates_per_s": v,
                "ms_per_step": ms_p / args.steps, "speedup_vs_fp64": v / value,
                "rel_l2_vs_fp64_max": max(e), "rel_l2_vs_fp64": e,
                "after_steps": nsteps, "meets_1e-6": bool(max(e) <= 1e-6),
                "step_roofline_frac": v * B_ALG[prec] / HBM_PEAK, "roofline": roof,
                "kernels": kern}
            c.close()
        # the same workload with viscosity and resistivity (addSourceTerm):
        # the cost of the Laplacians in the fused engine
        for prec in ("fp64", "mixed"):
            v, ms_p, c = run_single(m, u0, nx, ny, prec, args.steps, args.warmup, graph, hall,
                                    di, viscosity=1e-4, resistivity=1e-4)
            pe, pd = _bytes(prec)
            _, kern = roofline_of(c, nx, ny, pe, pd, 1, args.profile_steps, copy_gbps)
            forms[f"dissipative_{prec}_{nx}"] = {
                "workload": f"turbulence {nx}x{ny}, {prec}, viscosity = resistivity = 1e-4",
                "cell_updates_per_s": v, "ms_per_step": ms_p / args.steps,
                "vs_ideal_fp64": v / value, "kernels": kern}
            c.close()
        # the reference's own CPU-runnable case, the CPU baseline's workload:
        # 256 x 256 Orszag-Tang with the Hall CFL term (d_i = 0.05), 200 steps
        ot = m.orszag_tang(256, 256)
        v, ms_p, c = run_single(m, ot, 256, 256, "fp64", 200, args.warmup, graph, 1, 0.05)
        forms["orszag_tang_256_fp64"] = {
            "workload": "256x256 Orszag-Tang, d_i = 0.05, 200 RK3 steps, fp64",
            "cell_updates_per_s": v, "ms_per_step": ms_p / 200}
        c.close()
        forms["dropin_fp64_" + str(nx)] = dropin_form(m, u0, nx, ny, hall, di, value)
        forms.update(batched_forms(m, value, args, copy_gbps))
        forms.update(uneven_forms(m, value, args))
        forms.update(c3_forms(m, value, args.profile_steps, copy_gbps, hall, di))
        out["forms"] = forms
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nx, ny, u0, hall, di)
        ot_form = out.get("forms", {}).get("orszag_tang_256_fp64")
        if ot_form:
            # the GPU on the CPU baseline's own workload (a reported ratio)
            g = ot_form["cell_updates_per_s"]
            out["cpu_baseline"]["gpu_same_workload"] = {
                "value": g, "workload": ot_form["workload"],
                "vs_threads_n": g / out["cpu_baseline"]["value"],
                "vs_threads_1": g / out["cpu_baseline"]["threads_1"]["value"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
