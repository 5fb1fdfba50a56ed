"""Throughput of the RK3 substep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload zpg1536x315x1024]
                    [--precision both|mixed|fp64]

A step is one RK3 step (3 substeps: rhsp, outflow_correction, pois) of the
whole box, with every field resident in HBM; each substep runs as one
captured hipGraph.  The metric is grid-point substeps per second,
nx*ny*nz*3*K / wall time / 1e6 (Mpts*substep/s).  By default the FP64
substep and the mixed-precision substep are run from the same state for the
same steps; the line reports the mixed rate as `value`, the FP64 rate beside
it and the largest relative difference of u, v, w between the two.  One JSON
line is printed; per-kernel timings (HIP events on the library's stream) go
to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from opentbl_amd import config, ic
from opentbl_amd.fields import divergence_rel  # noqa: E402
from opentbl_amd.lib import GpuBox, OTBL_DT_EVERY_STEP, OTBL_FP64, OTBL_MIXED  # noqa: E402

METRIC = "Mpts*RK3-substep/s, mixed vs FP64 on 1 GPU at <=1e-5 rel err; % HBM peak"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
VEL_TARGET = 1e-5  # mixed vs FP64 velocity error the mixed substep is specified to


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_model(cfg, s=8):
    """Algorithmic bytes (each distinct array read once / written once) and
    flops per launch for field elements of s bytes (8: FP64, 4: the mixed
    substep's FP32 fields), for the kernels of one substep.  Returned as
    name -> (bytes per launch, flops per launch, launches per substep)."""
    X, Ru, Rv, z = cfg.nx, cfg.ny + 1, cfg.ny, cfg.nz1
    nz = cfg.nz
    f = X * z * s  # bytes of one row-plane column set (per row of a field)
    m = {}
    m["interpxx"] = ((2 * (Ru + Rv + Ru) * f) / 3, 0.0, 3)
    m["difvisxx"] = (((4 * Ru + 3 * Rv + 3 * Ru) * f) / 3, 0.0, 3)
    # complex transforms of length nz (5 nz log2 nz flops each, two real rows
    # per transform): 5 per row j <= ny and 3 for row ny+1 in plane loop 1,
    # 4 per row in plane loop 2
    m["products1"] = (X * (Rv * 8 + 5) * z * s, X * (Rv * 5 + 3) * 5 * nz * math.log2(nz), 1)
    m["products2"] = ((X - 2) * Rv * 7 * z * s, (X - 2) * Rv * 4 * 5 * nz * math.log2(nz), 1)
    m["interpyy"] = ((2 * Ru * X + (Ru + Rv) * (X - 2) + (Rv + Ru) * (X - 2)) * z * s / 3, 0.0, 3)
    m["rhsu_loop1"] = ((Rv + 2 * Ru) * f, 0.0, 1)
    # rhsu, rhsv, rhsw read and written; resu, resv, resw, vw, vv, ww read
    m["rhs_loop2"] = ((X - 2) * z * s * (2 * Ru + 2 * Rv + 2 * Ru + Ru + Rv + Ru + Ru + Rv + Rv), 0.0, 1)
    # res, rhs, rhs_prev, u read; u written (rhs becomes rhs_prev by a role swap)
    m["rk_final"] = ((5 * Ru + 5 * Rv + 5 * Ru) * f / 3, 0.0, 3)
    # the fused tail per component (FP64 / full-wave y): rhs and the field read
    # and written, res and rhs_prev read, and for v and w the two products of
    # plane loop 2 (planes 2..nx-1) read
    m["rk_tail"] = (((6 * Rv + 12 * Ru) * X + 4 * Ru * (X - 2)) * z * s / 3, 0.0, 3)
    m["save_predict"] = ((4 * Ru + 3 * Rv) * f, 0.0, 1)  # u, v, w, p read; predicted u, v, w written
    N, L = cfg.nx - 1, (cfg.ny - 1) * z
    # Bluestein length: the shortest engine length >= 2N-1 (3 x 1024 or a power of two)
    M = next(m for m in (512, 1024, 2048, 3072, 4096, 8192) if m >= 2 * N - 1)
    # Poisson x transforms: one pencil read and written; Bluestein FFTs of
    # length M, two lines per complex transform pair (forward + inverse)
    for nm in ("dct2_rows", "dct3_rows"):
        m[nm] = (2 * N * L * s, (L / 2) * 2 * 5 * M * math.log2(M), 1)
    m["div"] = ((Ru + Rv + Ru + Rv) * X * z * s, 0.0, 1)
    # the solution pencils read, the planes-layout phi written
    m["to_planes"] = (2 * N * L * s, 0.0, 1)
    m["poisson_y"] = (2 * (X - 1) * (cfg.ny - 1) * z * s, 0.0, 1)
    m["project"] = ((Rv + 2 * (Ru + Rv + Ru + Rv)) * X * z * s, 0.0, 1)
    m["boun"] = (2 * 4 * (Ru + Rv) * X * z * s / (cfg.ny), 0.0, 1)
    return m


# kernel-event names -> HIP kernel function names in rocprofv3 output
TRACE_NAMES = {"dct2_rows": "dct_rows_z", "dct3_rows": "dct_rows_z", "div": "div_pencils_kernel",
               "poisson_y": "poisson_y_chunk", "to_planes": "pencils_to_planes_kernel",
               "products1": ("products1_zp", "products1_z"), "products2": "products2_z", "interpxx": "interpxx_w",
               "difvisxx": "difvisxx_w", "interpyy": "interpyy_y", "rhsu_loop1": "rhsu_loop1_y",
               "rhs_loop2": "rhs_loop2_y", "rk_final": ("rk_final_y", "rk_final_yh"), "rk_tail": "rk_tail_y",
               "project": ("project_kernel", "project_pair_kernel"),
               "save_predict": ("save_predict_kernel", "save_predict_pair_kernel")}


def profiled_traffic(workload, precision, kname):
    """HBM bytes per launch of one kernel from the committed rocprofv3 summary
    of this workload and precision (FETCH_SIZE + WRITE_SIZE per dispatch,
    each scaled by what the counter reads for a plain 1 GiB device copy); the
    summary lists kernels by time, so the first row whose HIP function
    matches is the launch timed here.  None if absent."""
    path = os.path.join(ROOT, "profiles", f"substep_kernels_{workload}_{precision}.md")
    if not os.path.exists(path):
        return None
    fns = TRACE_NAMES.get(kname, kname + "_kernel")
    fns = (fns,) if isinstance(fns, str) else fns
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        name = cells[0] if cells else ""
        hit = any(name == fn or name.startswith(fn + "<") for fn in fns)
        if len(cells) >= 7 and hit and cells[5] and cells[6]:
            return (float(cells[5]) + float(cells[6])) * 1e9
    return None


def host_cpu():
    """CPU model name and logical CPU count of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(cfg_full, planes=256):
    """The CPU FP64 oracle (C, OpenMP over planes, every thread it is given)
    on a bounded sample of the same workload: one RK3 step (three substeps)
    of a slab of `planes` x-planes with the box's ny and nz."""
    from dataclasses import replace
    from oracle.oracle import OracleBox, build, lib as olib
    build()
    slab = replace(cfg_full, nx=min(planes, cfg_full.nx), name=cfg_full.name + f"-slab{planes}",
                   kzmax=min(cfg_full.kzmax, 8)).check()
    st = ic.make_state(slab)
    ob = OracleBox(slab, st)
    threads = olib().oc_max_threads()
    model, ncpu = host_cpu()
    log(f"[bench] CPU baseline: {slab.nx}x{slab.ny}x{slab.nz} slab, {threads} OpenMP threads on "
        f"{ncpu} logical CPUs ({model})")
    t0 = time.perf_counter()
    for m in (1, 2, 3):
        ob.substep(m)
        log(f"[bench] CPU baseline substep {m} done at {time.perf_counter() - t0:.1f} s")
    dt = time.perf_counter() - t0
    return {"value": slab.points * 3 / dt / 1e6, "unit": "Mpts*substep/s", "cores": threads, "kind": "port",
            "cpu": model, "nproc": ncpu,
            "sample": f"one RK3 step (3 substeps) of a {slab.nx}x{slab.ny}x{slab.nz} slab of the "
                      f"{cfg_full.nx}x{cfg_full.ny}x{cfg_full.nz} box, CPU FP64 oracle (C, OpenMP, "
                      f"{threads} threads), {dt:.1f} s"}


def xline_microbench(reps=10):
    """The compact x-line solves alone on 192*512 = 98,304 lines of 1024
    points, FP64 N(0,1) inputs (seed 1): an interpxx-type solve (1 input, 1
    output, 1.611 GB) and a difvisxx-type solve (2 inputs, 1 output, 2.416
    GB), through the C-ABI operators on a box whose planes hold 192 rows x
    512 reals (nx = 1024, ny = 192, nz = 768: v-type rows, nz1 = 512).
    Average launch time from HIP events on the library's stream."""
    from opentbl_amd.lib import DevArray
    cfg = config.BoxConfig("xline1024x192x768", 1024, 192, 768, 60.0, 4.0, 8.0, 600.0, 0.0, 0, "gauss", 1, 1)
    st = ic.make_state(cfg, with_perturbation=False)
    box = GpuBox(cfg, st, precision=OTBL_FP64)
    R, L, n = cfg.ny, cfg.ny * cfg.nz1, cfg.nx
    assert L == 98304
    rng = np.random.default_rng(1)
    a = DevArray.from_host(rng.standard_normal(n * L))
    b = DevArray.from_host(rng.standard_normal(n * L))
    out = DevArray(n * L)
    del st
    res = {}
    for name, nin in (("interpxx", 1), ("difvisxx", 2)):
        def call():
            if name == "interpxx":
                box.op("interpxx", a.ptr, out.ptr, R, 0)
            else:
                box.op("difvisxx", a.ptr, b.ptr, out.ptr, R, 0)
        call()
        box.sync()
        box.kprof(True)
        for _ in range(reps):
            call()
        box.sync()
        k, ms = box.kprof_report()[name]
        box.kprof(False)
        avg = ms / k
        by = (nin + 1) * n * L * 8
        gbs = by / (avg * 1e-3) / 1e9
        res[name] = {"lines": L, "length": n, "algorithmic_bytes": by, "avg_ms": avg,
                     "lower_bound_ms": by / (HBM_PEAK_GBS * 1e9) * 1e3, "achieved_GBs": gbs,
                     "frac": gbs / HBM_PEAK_GBS}
        log(f"[bench] x-line microbench {name}: {avg:.4f} ms per launch, {gbs:.0f} GB/s "
            f"({100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak)")
    for d in (a, b, out):
        d.free()
    box.close()
    return res


def run_box(cfg, st, precision, args):
    """Warm up, time args.steps RK3 steps, then one profiled step with the
    kernels bracketed by HIP events.  Returns wall seconds of the timed
    steps, the per-kernel report, the chosen paths and the final box."""
    name = "mixed" if precision == OTBL_MIXED else "fp64"
    box = GpuBox(cfg, st, precision=precision, dt_mode=OTBL_DT_EVERY_STEP, graphs=not args.no_graphs)
    paths = box.paths()
    log(f"[bench] {name}: kernel paths {paths}")
    for w in range(args.warmup):
        box.step(1)
        box.sync()
        log(f"[bench] {name}: warmup step {w + 1}/{args.warmup} done")
    box.sync()
    t0 = time.perf_counter()
    box.step(args.steps)
    box.sync()
    wall = time.perf_counter() - t0
    log(f"[bench] {name}: timed {args.steps} steps in {wall:.3f} s")
    box.kprof(True)
    box.step(1)
    box.sync()
    rep = box.kprof_report()
    box.kprof(False)
    return wall, rep, paths, box


def report(name, rep, model):
    rows = []
    for kname, (n, ms) in sorted(rep.items(), key=lambda kv: -kv[1][1]):
        avg = ms / n
        by, fl, _ = model.get(kname, (None, None, None))
        gbs = by / (avg * 1e-3) / 1e9 if by else None
        tfs = fl / (avg * 1e-3) / 1e12 if fl else None
        rows.append((kname, n, ms, avg, gbs, tfs))
        log(f"[kprof {name}] {kname:14s} launches {n:3d} total {ms:9.3f} ms avg {avg:9.4f} ms"
            + (f"  {gbs:8.1f} GB/s ({100 * gbs / HBM_PEAK_GBS:5.1f}% HBM)" if gbs else "")
            + (f"  {tfs:6.2f} TFLOP/s" if tfs else ""))
    return rows


def roofline(rows, model, workload, precision, hbm_only=False):
    """Roofline of the kernel with the largest time per substep among those
    the model covers (hbm_only: among those without a flop count, the
    streaming kernels)."""
    name, n, ms, avg, gbs, tfs = next(r for r in rows if r[0] in model and not (hbm_only and model[r[0]][1]))
    by, fl, _ = model[name]
    peak_tf = FP32_PEAK_TFLOPS if precision == "mixed" else FP64_PEAK_TFLOPS
    intensity = fl / by if fl else 0.0
    if intensity > peak_tf * 1e12 / (HBM_PEAK_GBS * 1e9):
        roof = {"bound": "mfma", "achieved": tfs, "peak": peak_tf, "unit": "TFLOP/s", "frac": tfs / peak_tf}
    else:
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
    roof["traffic"] = profiled_traffic(workload, precision, name)
    roof.update({"kernel": name, "avg_ms": avg, "algorithmic_bytes": by})
    return roof


ENSEMBLE = "zpg_ensemble8"


def run_ensemble(args):
    """The uneven eight-box ensemble (nx 256..1024, ny 96..192, nz 64..256;
    mixed precision unless --precision fp64): K RK3 steps of all boxes
    through the ensemble (one graph launch per substep for all boxes), then
    the same K steps with each box stepped alone, one box after another.
    Returns the JSON record."""
    from opentbl_amd.lib import Ensemble
    cfgs = config.ensemble_configs()
    prec = OTBL_FP64 if args.precision == "fp64" else OTBL_MIXED
    name = "fp64" if prec == OTBL_FP64 else "mixed"
    states = [ic.make_state(c) for c in cfgs]
    boxes = [GpuBox(c, s, precision=prec, dt_mode=OTBL_DT_EVERY_STEP, graphs=True) for c, s in zip(cfgs, states)]
    del states
    points = sum(c.points for c in cfgs)
    ens = Ensemble(boxes)
    for w in range(args.warmup):
        ens.step(1)
        ens.sync()
        for b in boxes:
            b.step(1)
            b.sync()
        log(f"[bench] ensemble warmup step {w + 1}/{args.warmup} done")
    ens.sync()
    t0 = time.perf_counter()
    ens.step(args.steps)
    ens.sync()
    wall = time.perf_counter() - t0
    log(f"[bench] ensemble: {args.steps} steps in {wall:.3f} s")
    t0 = time.perf_counter()
    for b in boxes:
        b.step(args.steps)
    for b in boxes:
        b.sync()
    wall_seq = time.perf_counter() - t0
    log(f"[bench] boxes one after another: {args.steps} steps in {wall_seq:.3f} s")
    for b in boxes:
        f = b.get("u")
        if not np.isfinite(f).all():
            raise SystemExit("bench.py: non-finite u in the ensemble")
    ens.close()
    for b in boxes:
        b.close()
    value = points * 3 * args.steps / wall / 1e6
    return {
        "metric": METRIC, "value": value, "unit": "Mpts*substep/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if prec == OTBL_MIXED else "f64",
        "data": "synthetic",
        "config": {"workload": "uneven 8-box ZPG ensemble (nx 256..1024, ny 96..192, nz 64..256, "
                               f"Re_theta_in 300..650), {name} substep, one hipGraph per substep for all boxes",
                   "boxes": [f"{c.nx}x{c.ny}x{c.nz}" for c in cfgs], "points": points,
                   "one_after_another_value": points * 3 * args.steps / wall_seq / 1e6,
                   "ensemble_over_one_after_another": wall_seq / wall},
        "roofline": None, "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="zpg1536x315x1024", choices=sorted(config.CONFIGS) + [ENSEMBLE])
    ap.add_argument("--precision", default="both", choices=["both", "mixed", "fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-xline", action="store_true", help="skip the x-line microbench")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one box on one GPU; --gpus {args.gpus} is not "
              "supported (the x-slab multi-GPU layout is design only)", file=sys.stderr)
        sys.exit(2)
    if args.workload == ENSEMBLE:
        print(json.dumps(run_ensemble(args)), flush=True)
        return
    cfg = config.CONFIGS[args.workload]
    log(f"[bench] workload {cfg.name} {cfg.nx}x{cfg.ny}x{cfg.nz}: generating the state")
    t0 = time.perf_counter()
    st = ic.make_state(cfg)
    log(f"[bench] state ready in {time.perf_counter() - t0:.1f} s")
    # a plain device copy of 1 GiB: the known-size reference for hardware
    # byte counters (FETCH_SIZE / WRITE_SIZE) read by rocprofv3 on this command
    from opentbl_amd.lib import DevArray, lib as _lib, check as _check
    _a, _b = DevArray(1 << 27), DevArray(1 << 27)
    _check(_lib().otbl_d2d(_b.ptr, _a.ptr, 1 << 30), "d2d")
    _a.free()
    _b.free()
    runs = {}
    ref, ref_rc = None, None
    order = {"both": [OTBL_FP64, OTBL_MIXED], "mixed": [OTBL_MIXED], "fp64": [OTBL_FP64]}[args.precision]
    for prec in order:
        name = "mixed" if prec == OTBL_MIXED else "fp64"
        wall, rep, paths, box = run_box(cfg, st, prec, args)
        vel = {n: box.get(n) for n in ("u", "v", "w", "p")}
        if not all(np.isfinite(a).all() for a in vel.values()):
            raise SystemExit(f"bench.py: non-finite velocity after the {name} steps")
        rth, cf = box.retheta_cf()
        box.close()
        model = kernel_model(cfg, 4 if prec == OTBL_MIXED else 8)
        rows = report(name, rep, model)
        runs[name] = {"wall": wall, "value": cfg.points * 3 * args.steps / wall / 1e6, "paths": paths,
                      "roofline": roofline(rows, model, cfg.name, name),
                      "hbm_roofline": roofline(rows, model, cfg.name, name, hbm_only=True),
                      # discrete divergence of pois.f90:41-54 after the run, relative to max|u|/dx
                      "divergence": divergence_rel(cfg, st["y"], vel["u"], vel["v"], vel["w"])}
        if prec == OTBL_FP64:
            ref = vel
            ref_rc = (rth, cf)
        elif ref is not None:
            # velocity error relative to the largest velocity component (the
            # free stream); each field also relative to its own maximum
            scale = max(np.abs(ref[n]).max() for n in ("u", "v", "w"))
            err = {"vel": max(float(np.abs(vel[n] - ref[n]).max()) for n in ("u", "v", "w")) / scale}
            err.update({n: float(np.abs(vel[n] - ref[n]).max() / np.abs(ref[n]).max()) for n in vel})
            # Re_theta(x) and Cf(x) (genflu.F90:76-87) against the FP64 run's
            err["Re_theta"] = float(np.abs(rth - ref_rc[0]).max() / np.abs(ref_rc[0]).max())
            err["Cf"] = float(np.abs(cf - ref_rc[1]).max() / np.abs(ref_rc[1]).max())
            runs[name]["rel_err_vs_fp64"] = err
            # the mixed substep's specification: velocity within 1e-5 of the
            # FP64 run; each velocity component is judged against its own
            # maximum as well as against the free stream
            runs[name]["within_target"] = {n: bool(err[n] <= VEL_TARGET) for n in ("vel", "u", "v", "w")}
        del vel
    del st, ref, ref_rc
    head = "mixed" if "mixed" in runs else "fp64"
    h = runs[head]
    steps = args.steps
    cfgd = {"workload": f"{cfg.nx}x{cfg.ny}x{cfg.nz} ZPG boundary layer, "
                        f"Re_theta_in={cfg.re_theta:g}, {head} substep, one hipGraph per substep",
            "nx": cfg.nx, "ny": cfg.ny, "nz": cfg.nz, "precision": head, "graphs": not args.no_graphs,
            "dt": "CFL every step", "paths": h["paths"],
            # the largest streaming kernel, when the dominant one is bound by arithmetic
            "hbm_roofline": h["hbm_roofline"]}
    for name, r in runs.items():
        cfgd[f"{name}_value"] = r["value"]
        cfgd[f"{name}_ms_per_step"] = 1e3 * r["wall"] / steps
        cfgd[f"{name}_divergence"] = r["divergence"]
    if "mixed" in runs and "fp64" in runs:
        cfgd["mixed_over_fp64"] = runs["mixed"]["value"] / runs["fp64"]["value"]
        cfgd["mixed_rel_err_vs_fp64"] = runs["mixed"]["rel_err_vs_fp64"]
        # errors after warmup + timed + one profiled RK3 step of each run
        cfgd["mixed_err_after_steps"] = args.warmup + args.steps + 1
        cfgd["mixed_within_target"] = runs["mixed"]["within_target"]
        cfgd["fp64_roofline"] = runs["fp64"]["roofline"]
    out = {
        "metric": METRIC, "value": h["value"], "unit": "Mpts*substep/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": 1e3 * h["wall"] / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64", "data": "synthetic",
        "config": cfgd,
        "roofline": h["roofline"],
    }
    if not args.no_xline:
        log("[bench] x-line microbench (98,304 lines x 1024, FP64)")
        cfgd["xline_microbench"] = xline_microbench()
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle on a slab)")
        out["cpu_baseline"] = cpu_baseline(cfg)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out, default=json_default), flush=True)


def json_default(o):
    """numpy scalars and arrays in the output line as plain JSON values."""
    if isinstance(o, np.generic):
        return o.item()
    if isinstance(o, np.ndarray):
        return o.tolist()
    raise TypeError(f"{type(o).__name__} is not JSON serializable")


if __name__ == "__main__":
    main()
