"""Throughput of the MI355X grid backend on one GPU.

One step = one grid_collocate_task_list (density -> all grid levels) plus one
grid_integrate_task_list (potential -> Hamiltonian blocks, and forces /
virial where the workload asks for them) over the same task list, with inputs
and outputs resident in HBM (the library's *_dev entry points).  value =
task-points processed per second (collocate + integrate), where a task-point
is one (task, grid point) pair inside the reference's truncation region.

Default workload (the headline): 1024 atoms at 0.1 atoms/A^3 in a triclinic
cell (angles 80-100 deg), 4 grid levels from 200^3, TZV2P-like shells l <= 3,
exponents 0.03-100, integrate with forces, FP32-mixed point evaluation with
FP64 accumulation.  The plain FP64 kernels are timed in the same run (the
`fp64` object) and the FP32-mixed outputs are compared with them per result
(rho per level, hab, forces, virial).  Synthetic, seeded (cp2k_amd/synth.py).

Also reported:
  roofline        of the dominant call (collocate or integrate), HBM-bound:
                  achieved = algorithmic bytes of one call / its average
                  duration from HIP events on the launch stream.  Algorithmic
                  bytes: collocate 16 B per task-point (FP64 read + write of
                  the grid value) + per task its Cartesian density sub-block
                  and record + 8 B per level grid point (overwrite); integrate
                  8 B per task-point + per task its hab sub-block (and density
                  sub-block with forces) and record.  traffic = measured HBM
                  bytes of one call from a committed rocprofv3 PMC summary
                  (profiles/), used only when it was measured on the current
                  kernel sources, else null.  roofline_calls holds the same for
                  both calls, FP64 throughput of each call (2 (lp + 1) flops per
                  task-point) against the FP64 peak (78.6 TFLOP/s).
  abi_host_buffers  the drop-in C-ABI (grid_collocate_task_list /
                  grid_integrate_task_list on offload_buffers): pab H2D, grids
                  and hab (forces) D2H per call, as CP2K calls it.  Never value.
  cpu_baseline    CP2K's own OpenMP cpu-backend kernels (oracle/_ref) on a
                  bounded sample of the same workload, threads bound to cores.

`python bench.py --rocprof-summary DIR` turns rocprofv3 output (kernel trace
statistics and FETCH_SIZE / WRITE_SIZE counter runs of this script) into the
text summaries kept under profiles/.
"""
import argparse
import csv
import glob
import hashlib
import json
import os
import pathlib
import re
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64
HBM_PEAK_GBS = 8000.0    # MI355X HBM3E
METRIC = "collocate+integrate Gpts/s & HBM GB/s; FP32-mixed vs FP64 at <=2e-6 rel rho err"

# kernels of one collocate / integrate call, by rocprofv3 kernel name
CALL_KERNELS = {
    "collocate": ("cxyz_lane_kernel", "cxyz_wave_kernel", "collocate_tile_kernel",
                  "collocate_wave_kernel", "general_collocate_kernel", "general_tile_kernel",
                  "general_cijk_kernel", "collocate_kernel"),
    "integrate": ("integrate_wave_kernel", "integrate_mfma_kernel", "finish_wave_kernel",
                  "finish_block_kernel", "general_integrate_kernel", "to_float_kernel",
                  "integrate_kernel"),
}
# timing ids (grid.KERNEL_NAMES) inside each call
CALL_IDS = {
    "collocate": ("cxyz", "collocate_tiles", "collocate_direct", "per_task", "general_points"),
    "integrate": ("integrate_points", "finish", "per_task", "general_points", "to_float"),
}
WORKLOADS = {
    "h2o64": "64 H2O, 12.4 A cube, 4 levels 128/72/42/24, DZVP-like l<=2",
    "tri1024": "1024 atoms, triclinic 80-100 deg, 4 levels 200/112/64/36, TZV2P-like l<=3",
    "hl2048": "2048 atoms, orthorhombic, 5 levels 256/150/84/48/28, l<=4",
    "batch256": "256 lists of 8-128 atoms, cubic 6-14 A, one level 32^3-112^3, l<=2",
}
# what each workload computes: integrate derivatives and the precision of the
# headline leg (the other precision is timed beside it)
WORKLOAD_MODE = {
    "h2o64": dict(forces=False, virial=False, headline="fp64"),
    "tri1024": dict(forces=True, virial=False, headline="fp32mixed"),
    "hl2048": dict(forces=True, virial=True, headline="fp32mixed"),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", default="tri1024")
    p.add_argument("--precision", choices=["fp64", "fp32mixed"], default=None,
                   help="headline leg (default: the workload's own)")
    p.add_argument("--forces", choices=["auto", "on", "off"], default="auto",
                   help="integrate also forces (auto: the workload's own)")
    p.add_argument("--virial", choices=["auto", "on", "off"], default="auto")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-abi", action="store_true", help="skip the host-buffer C-ABI timing")
    p.add_argument("--no-other", action="store_true",
                   help="time only the headline precision (no FP64 / FP32-mixed comparison)")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep the default tile launch shapes (no grid_mi355x_autotune)")
    p.add_argument("--secondary", default="h2o64",
                   help="comma-separated workloads also measured (no CPU / ABI legs) and "
                        "reported under 'secondary'; empty for none")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--cpu-only", action="store_true",
                   help="run only the CPU baseline of --config and print it (no GPU); "
                        "bench.py runs itself this way in a child process")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="before the steps, run 3 device-to-device copies of MIB MiB: "
                        "a plain copy kernel of known size that scales the HBM "
                        "byte counters of a rocprofv3 --pmc run")
    p.add_argument("--rocprof-summary", metavar="DIR",
                   help="summarise rocprofv3 output under DIR and exit (no GPU)")
    p.add_argument("--calls", type=int, default=1,
                   help="with --rocprof-summary: collocate/integrate calls in the run "
                        "(the copy calibration size is --copy-calibration, default 512)")
    return p.parse_args()


def mode_of(opt, name):
    m = dict(WORKLOAD_MODE.get(name, dict(forces=False, virial=False, headline="fp64")))
    if opt.forces != "auto":
        m["forces"] = opt.forces == "on"
    if opt.virial != "auto":
        m["virial"] = opt.virial == "on"
    m["forces"] = m["forces"] or m["virial"]
    if opt.precision:
        m["headline"] = opt.precision
    return m


def integrate_shift(forces, virial):
    """la_max + lb_max shift of an integrate call (grid_process_vab.h:222-251)."""
    return (1 if forces else 0) + (2 if virial else 0)


def algorithmic_bytes(args, points, which, forces=False):
    """Algorithmic HBM bytes of one collocate / integrate call (module doc)."""
    from cp2k_amd.tasklist import ncoset
    import numpy as np
    la = np.array([b.lmax[s - 1] for b, s in zip(
        [args.basis_sets[args.atom_kinds[i - 1] - 1] for i in args.iatom_list],
        args.iset_list)])
    lb = np.array([b.lmax[s - 1] for b, s in zip(
        [args.basis_sets[args.atom_kinds[j - 1] - 1] for j in args.jatom_list],
        args.jset_list)])
    nc = np.vectorize(ncoset)
    sub = 8.0 * float(np.sum(nc(la) * nc(lb)))
    record = 280.0 * args.ntasks
    if which == "collocate":
        grid_pts = float(sum(int(np.prod(n)) for n in args.npts_local))
        return 16.0 * points + sub + record + 8.0 * grid_pts
    return 8.0 * points + sub * (2.0 if forces else 1.0) + record


def source_hash():
    """Hash of the kernel sources a profile summary was measured on."""
    h = hashlib.sha256()
    for f in sorted((ROOT / "cp2k_amd" / "csrc").glob("*")):
        h.update(f.name.encode())
        h.update(f.read_bytes())
    return h.hexdigest()[:16]


def call_of(name):
    for call, keys in CALL_KERNELS.items():
        if any(k + "<" in name or k + "(" in name for k in keys):
            return call
    return None


def rocprof_summary(d, calls, copy_mib):
    """Text summary of rocprofv3 output under d: kernel statistics, and HBM
    traffic per collocate / integrate call from FETCH_SIZE and WRITE_SIZE
    counters scaled by what they read for the copy calibration (a device copy
    of known size, __amd_rocclr_copyBuffer)."""
    out = [f"source_hash: {source_hash()}"]
    for f in sorted(glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)):
        out.append(f"== kernel statistics ({os.path.relpath(f, d)})")
        out.append(f"{'kernel':70s} {'calls':>6s} {'total ms':>10s} {'avg ms':>9s} {'%':>6s}")
        for r in csv.DictReader(open(f)):
            out.append(f"{r['Name'][:70]:70s} {int(r['Calls']):6d} "
                       f"{float(r['TotalDurationNs']) / 1e6:10.3f} "
                       f"{float(r['AverageNs']) / 1e6:9.3f} {float(r['Percentage']):6.2f}")
    for f in sorted(glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)):
        # dispatches after the calibration copies (issued after autotuning,
        # right before the timed steps): the steps' kernels only
        rows = list(csv.DictReader(open(f)))
        dur = lambda r: int(r["End_Timestamp"]) - int(r["Start_Timestamp"])  # noqa: E731
        copies = [r for r in rows if "copyBuffer" in r["Kernel_Name"]]
        big = max((dur(r) for r in copies), default=0)
        start = max((int(r["End_Timestamp"]) for r in copies if dur(r) > big // 2), default=0)
        agg = {}
        for r in rows:
            if int(r["Start_Timestamp"]) < start or "copyBuffer" in r["Kernel_Name"]:
                continue
            a = agg.setdefault(r["Kernel_Name"], [0, 0.0])
            a[0] += 1
            a[1] += dur(r) / 1e6
        out.append(f"== kernel trace after the calibration copies ({os.path.relpath(f, d)}), "
                   f"{calls} call(s) of each")
        out.append(f"{'kernel':70s} {'calls':>6s} {'total ms':>10s} {'avg ms':>9s}")
        for name, (n, ms) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
            out.append(f"{name[:70]:70s} {n:6d} {ms:10.3f} {ms / n:9.4f}")
        for call in ("collocate", "integrate"):
            ms = sum(v[1] for k, v in agg.items() if call_of(k) == call)
            out.append(f"call {call}: kernels {ms / calls:.3f} ms per call")
    per = {}  # counter -> {"copy": [...], call: total KiB}
    for f in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
        for r in csv.DictReader(open(f)):
            c, v, n = r["Counter_Name"], float(r["Counter_Value"]), r["Kernel_Name"]
            e = per.setdefault(c, {"copy": [], "collocate": 0.0, "integrate": 0.0})
            if "copyBuffer" in n:
                e["copy"].append(v)
            elif call_of(n):
                e[call_of(n)] += v
    if "FETCH_SIZE" in per and "WRITE_SIZE" in per:
        out.append("== HBM traffic per call (KiB from FETCH_SIZE / WRITE_SIZE, per dispatch summed)")
        # the calibration copies (the largest copy dispatches) read and write
        # copy_mib MiB each
        copy_kib = 1024.0 * copy_mib
        fetch_copy = max(per["FETCH_SIZE"]["copy"] or [0.0])
        write_copy = max(per["WRITE_SIZE"]["copy"] or [0.0])
        ff = copy_kib / fetch_copy if fetch_copy else float("nan")
        wf = copy_kib / write_copy if write_copy else float("nan")
        out.append(f"copy calibration: a {copy_kib:.0f} KiB device copy reads FETCH_SIZE "
                   f"{fetch_copy:.0f} KiB and WRITE_SIZE {write_copy:.0f} KiB -> "
                   f"fetch factor {ff:.3f}, write factor {wf:.3f}")
        for call in ("collocate", "integrate"):
            fk = per["FETCH_SIZE"][call] / calls
            wk = per["WRITE_SIZE"][call] / calls
            tb = (fk * ff + wk * wf) * 1024.0
            out.append(f"call {call}: fetch {fk:.0f} KiB, write {wk:.0f} KiB, "
                       f"traffic_bytes {tb:.0f}")
    # any other counters: totals per kernel (all dispatches of the run)
    kern = {}
    for f in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
        for r in csv.DictReader(open(f)):
            c = r["Counter_Name"]
            if c in ("FETCH_SIZE", "WRITE_SIZE"):
                continue
            name = re.sub(r"\(.*", "", r["Kernel_Name"]).replace("void ", "")
            kern.setdefault(name, {})
            kern[name][c] = kern[name].get(c, 0.0) + float(r["Counter_Value"])
    if kern:
        out.append(f"== counters per kernel, totals over the run ({calls} call(s))")
        for name in sorted(kern):
            out.append(name)
            for c in sorted(kern[name]):
                out.append(f"    {c:26s} {kern[name][c]:14.5g}")
    return "\n".join(out)


def profile_traffic(call, tag):
    """HBM bytes per call from a committed summary measured on these sources
    (profiles/<tag>_hbm_traffic.txt)."""
    f = ROOT / "profiles" / f"{tag}_hbm_traffic.txt"
    if not f.exists():
        return None
    text = f.read_text()
    m = re.search(r"source_hash: (\w+)", text)
    if not m or m.group(1) != source_hash():
        return None
    m = re.search(rf"call {call}: .*traffic_bytes (\d+)", text)
    return float(m.group(1)) if m else None


def physical_cores():
    """Physical cores of the host (distinct (package, core) pairs)."""
    try:
        pairs, phys = set(), None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k == "physical id":
                phys = v.strip()
            elif k == "core id":
                pairs.add((phys, v.strip()))
        return len(pairs) or None
    except OSError:
        return None


def cpu_baseline(args, pab, V, seconds, forces, virial):
    """CP2K's cpu-backend kernels (grid_cpu_*_pgf_product, oracle/_ref) under
    an OpenMP block loop, on a random sample of the workload's blocks sized to
    roughly `seconds` of CPU time; integrate with the workload's forces and
    virial.  Threads: OMP_NUM_THREADS (the CPU share of this GPU) or every
    CPU this process may run on, bound one per core (OMP_PROC_BIND=spread,
    OMP_PLACES=cores).  Runs in a child process of its own (--cpu-only): the
    binding would otherwise pin this process's later threads too."""
    import dataclasses

    import numpy as np

    affinity = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or affinity
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle as O
    if O.ref_lib(fast=True) is None:
        return None
    rng = np.random.default_rng(1)
    nblk = args.nblocks

    def subset(frac):
        keep = rng.random(nblk) < frac
        sel = np.nonzero(keep[args.block_num_list - 1])[0]
        kw = {f.name: getattr(args, f.name) for f in dataclasses.fields(args)}
        for name in ("level_list", "iatom_list", "jatom_list", "iset_list", "jset_list",
                     "ipgf_list", "jpgf_list", "border_mask_list", "block_num_list",
                     "radius_list", "rab_list"):
            kw[name] = kw[name][sel]
        return type(args)(**kw)

    def run(tl):
        t0 = time.perf_counter()
        tl.collocate(100, pab, kernel="ref")
        t1 = time.perf_counter()
        tl.integrate(V, pab, forces=forces, virial=virial, kernel="ref")
        return t1 - t0, time.perf_counter() - t1

    frac = 0.002
    a, b = run(O.TaskList(subset(frac)))
    frac = min(1.0, frac * seconds / max(a + b, 1e-3))
    tl = O.TaskList(subset(frac))
    pts = tl.task_points()
    tc, ti = run(tl)
    return {"value": round(2 * pts / (tc + ti) / 1e9, 4), "unit": "Gpts/s",
            "cores": threads, "kind": "reference",
            "host_physical_cores": physical_cores(),
            "affinity_cpus": affinity,
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND",
                                                  "OMP_PLACES")},
            "sample": (f"{tl.a.ntasks} of {args.ntasks} tasks ({100 * frac:.2f}% of blocks), "
                       f"{pts:.3e} task-points; CP2K cpu-backend kernels "
                       f"grid_cpu_collocate/integrate_pgf_product (-O3 -march=x86-64-v3), "
                       f"OpenMP over blocks; integrate forces={forces} virial={virial}; "
                       f"collocate {tc:.2f} s, integrate {ti:.2f} s")}


def log(msg):
    """Progress of the run, on stderr (the JSON line goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_steps(step, steps):
    from cp2k_amd import grid
    grid.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    grid.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def outputs(gt, pab_d, rho_d, V_d, natoms, buffer_size, forces, virial):
    """One collocate and one integrate on the device, copied to the host:
    (rho levels, hab, forces, virial)."""
    from cp2k_amd import grid
    hab = grid.DeviceArray(buffer_size)
    f = grid.DeviceArray((natoms, 3)) if forces else None
    v = grid.DeviceArray((3, 3)) if virial else None
    gt.collocate_dev(100, pab_d, rho_d)
    gt.integrate_dev(V_d, hab, pab=pab_d if forces else None, forces=f, virial=v)
    grid.synchronize()
    res = ([r.numpy() for r in rho_d], hab.numpy(), f.numpy() if f else None,
           v.numpy() if v else None)
    for a in (hab, f, v):
        if a is not None:
            a.free()
    return res


def roofline_of(call, kt, by_lp, shift, abytes, traffic):
    """Roofline object of one call from the per-kernel timings."""
    from cp2k_amd import grid
    ms, n = kt[call]
    avg_ms = ms / max(n, 1)
    flops = sum(2.0 * (lp + shift + 1) * p for lp, p in enumerate(by_lp))
    tflops = flops / (avg_ms * 1e-3) / 1e12
    achieved = abytes / (avg_ms * 1e-3) / 1e9
    names = [k for k in grid.KERNEL_NAMES if k in CALL_IDS[call] and kt[k][0] > 0]
    return {
        "bound": "hbm",
        "kernel": f"{call} call ({', '.join(names)})",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "algorithmic_bytes_per_launch": abytes,
        "avg_launch_ms": round(avg_ms, 4),
        "measured_hbm_gbs": (round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None),
        "fp64_tflops": round(tflops, 3),
        "fp64_frac": round(tflops / FP64_PEAK_TFLOPS, 4),
        "algorithmic_flops_per_launch": flops,
    }


def leg(opt, gt, step, precision, name, args, points, by_lp, mode):
    """Timed steps in one precision: ms per step, kernel times, rooflines."""
    from cp2k_amd import grid
    grid.set_precision(grid.FP64 if precision == "fp64" else grid.FP32_MIXED)
    for _ in range(opt.warmup):
        step()
    grid.synchronize()
    grid.reset_timing()
    grid.set_timing(True)
    ms = timed_steps(step, opt.steps)
    grid.set_timing(False)
    kt = {k: grid.kernel_time_ms(i) for i, k in enumerate(grid.KERNEL_NAMES)}
    launches = {k: grid.point_launches(i) for i, k in enumerate(grid.KERNEL_NAMES)
                if sum(grid.point_launches(i)) > 0}
    tag = f"{name}_{'forces_' if mode['forces'] else ''}{precision}"
    calls = {}
    for call in ("collocate", "integrate"):
        shift = 0 if call == "collocate" else integrate_shift(mode["forces"], mode["virial"])
        calls[call] = roofline_of(call, kt, by_lp, shift,
                                  algorithmic_bytes(args, points, call, mode["forces"]),
                                  profile_traffic(call, tag))
    dom = max(("collocate", "integrate"), key=lambda k: kt[k][0])
    return {
        "value": round(2.0 * points / (ms * 1e-3) / 1e9, 3),
        "unit": "Gpts/s",
        "ms_per_step": round(ms, 3),
        "tasks_per_s": round(2 * args.ntasks / (ms * 1e-3), 1),
        "kernel_ms_per_step": {k: round(v[0] / max(opt.steps, 1), 4) for k, v in kt.items()},
        "point_launches_fp64_fp32": {k: list(v) for k, v in launches.items()},
        "roofline": calls[dom],
        "roofline_calls": calls,
    }


def abi_leg(opt, gt, args, pab, V, points, mode, precision):
    """The drop-in C-ABI on host buffers: every call copies pab H2D and the
    grids / hab / forces D2H (src/grid/gpu/grid_gpu_context.cu:509-529)."""
    from cp2k_amd import grid
    grid.set_precision(grid.FP64 if precision == "fp64" else grid.FP32_MIXED)
    abi = grid.AbiBuffers(gt)
    abi.set_inputs(pab, V)

    def step():
        abi.collocate(100)
        abi.integrate(forces=mode["forces"], virial=mode["virial"])

    step()
    n = max(2, opt.steps // 2)
    t0 = time.perf_counter()
    for _ in range(n):
        step()
    ms = (time.perf_counter() - t0) * 1e3 / n
    abi.free()
    grid.set_precision(grid.FP64)
    return {"value": round(2.0 * points / (ms * 1e-3) / 1e9, 3), "unit": "Gpts/s",
            "ms_per_step": round(ms, 3), "steps": n, "precision": precision,
            "note": "PCIe-inclusive (host offload_buffers), not the headline value"}


def measure(opt, name, cpu, abi):
    """One workload: the headline precision and (unless --no-other) the other
    one, per-result errors of FP32-mixed against FP64, rooflines, and with cpu
    / abi the CPU baseline and the host-buffer ABI timing."""
    from cp2k_amd import grid, precision as prec, synth
    if name.startswith("batch"):
        return measure_batch(opt, name)
    mode = mode_of(opt, name)
    cpu_res = None
    if cpu:
        cpu_res = cpu_child(opt, name)
        log(f"{name}: CPU baseline done")
    t0 = time.time()
    args, pab, V = synth.config(name, seed=0)
    gen_s = time.time() - t0
    log(f"{name}: generated {args.ntasks} tasks in {gen_s:.1f} s")
    grid.set_precision(grid.FP64)
    gt = grid.GridTaskList(args)
    points = gt.task_points()
    log(f"{name}: task list built, {points} task points")
    by_lp = gt.task_points_by_lp()
    pab_d = grid.DeviceArray.from_numpy(pab)
    rho_d = [grid.DeviceArray(s) for s in args.grid_shapes()]
    V_d = [grid.DeviceArray.from_numpy(v) for v in V]
    hab_d = grid.DeviceArray(args.buffer_size)
    f_d = grid.DeviceArray((args.natoms, 3)) if mode["forces"] else None
    v_d = grid.DeviceArray((3, 3)) if mode["virial"] else None

    def step():
        gt.collocate_dev(100, pab_d, rho_d)
        gt.integrate_dev(V_d, hab_d, pab=pab_d if mode["forces"] else None, forces=f_d,
                         virial=v_d)

    tuned = None
    if not opt.no_autotune:
        t0 = time.time()
        gt.autotune(100, pab_d, rho_d)
        tuned = {"seconds": round(time.time() - t0, 2), "levels": [
            {"ortho": gt.tuned_shape(l), "general": gt.tuned_shape(l, general=True)}
            for l in range(args.nlevels)]}
        log(f"{name}: autotuned in {tuned['seconds']} s")

    if opt.copy_calibration:
        n = opt.copy_calibration * (1 << 20) // 8
        src, dst = grid.DeviceArray((n,)), grid.DeviceArray((n,))
        for _ in range(3):
            dst.copy_from(src)
        src.free()
        dst.free()

    head = mode["headline"]
    other = "fp64" if head == "fp32mixed" else "fp32mixed"
    legs = {head: leg(opt, gt, step, head, name, args, points, by_lp, mode)}
    log(f"{name}: {head} steps of {legs[head]['ms_per_step']:.3f} ms")
    if not opt.no_other:
        legs[other] = leg(opt, gt, step, other, name, args, points, by_lp, mode)
        log(f"{name}: {other} steps of {legs[other]['ms_per_step']:.3f} ms")
    H = legs[head]
    out = {
        "metric": METRIC,
        "value": H["value"],
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": opt.steps,
        "warmup": opt.warmup,
        "ms_per_step": H["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32-mixed(f64-accumulate)",
        "data": "synthetic (seeded cp2k_amd.synth, random density/potential)",
        "config": {
            "workload": WORKLOADS.get(name, name),
            "config_name": name,
            "ntasks": args.ntasks,
            "task_points": points,
            "task_points_by_lp": by_lp,
            "levels": [int(n[0]) for n in args.npts_local],
            "forces": bool(mode["forces"]),
            "virial": bool(mode["virial"]),
            "precision": head,
            "tasks_per_s": H["tasks_per_s"],
            "generate_s": round(gen_s, 2),
            "autotune": tuned,
        },
        "kernel_ms_per_step": H["kernel_ms_per_step"],
        "point_launches_fp64_fp32": H["point_launches_fp64_fp32"],
        "roofline": H["roofline"],
        "roofline_calls": H["roofline_calls"],
        "cpu_baseline": cpu_res,
    }
    if other in legs:
        out[other] = legs[other]
        grid.set_precision(grid.FP64)
        ref = outputs(gt, pab_d, rho_d, V_d, args.natoms, args.buffer_size, mode["forces"],
                      mode["virial"])
        grid.set_precision(grid.FP32_MIXED)
        got = outputs(gt, pab_d, rho_d, V_d, args.natoms, args.buffer_size, mode["forces"],
                      mode["virial"])
        grid.set_precision(grid.FP64)
        err = prec.report(rho=(got[0], ref[0]), hab=(got[1], ref[1]),
                          forces=(got[2], ref[2]) if mode["forces"] else None,
                          virial=(got[3], ref[3]) if mode["virial"] else None)
        ms32 = legs["fp32mixed"]["ms_per_step"]
        out["fp32_mixed_vs_fp64"] = {
            "speedup": round(legs["fp64"]["ms_per_step"] / ms32, 3),
            "errors": {k: (float(f"{v:.4g}") if isinstance(v, float) else
                           [{a: float(f"{b:.4g}") for a, b in e.items()} for e in v])
                       for k, v in err.items()},
            "bars": {"rho_linf": prec.RHO_LINF_BAR, "forces_linf": prec.FORCE_BAR},
            "within_bars": bool(err["rho_linf"] <= prec.RHO_LINF_BAR and
                                err.get("forces_linf", 0.0) <= prec.FORCE_BAR),
        }
    if abi:
        out["abi_host_buffers"] = abi_leg(opt, gt, args, pab, V, points, mode, head)
        log(f"{name}: host-buffer ABI steps of {out['abi_host_buffers']['ms_per_step']:.3f} ms")
    for a in [pab_d, hab_d, f_d, v_d] + rho_d + V_d:
        if a is not None:
            a.free()
    gt.free()
    return out


def measure_batch(opt, name):
    """Batched lists: one step = one replay of the batch graph (collocate
    then integrate of every list).  Also timed: the same lists called one by
    one through the device entry points (unbatched)."""
    from cp2k_amd import grid, precision, synth
    nlists = int(name[5:] or 256)
    t0 = time.time()
    lists = synth.batch_config(nlists, seed=0)
    gen_s = time.time() - t0
    log(f"{name}: generated {nlists} lists in {gen_s:.1f} s")
    grid.set_precision(grid.FP64)
    gts = []
    for i, (a, _, _) in enumerate(lists):
        gts.append(grid.GridTaskList(a))
        if i % 32 == 31:
            log(f"{name}: {i + 1} task lists built")
    batch = grid.GridBatch(gts)
    log(f"{name}: batch built")
    for i, (a, p, V) in enumerate(lists):
        batch.set_inputs(i, pab=p, V=V)
    points = batch.task_points()
    by_lp = [0] * 11
    for g in gts:
        for lp, x in enumerate(g.task_points_by_lp()):
            by_lp[lp] += x
    ntasks = sum(a.ntasks for a, _, _ in lists)
    forces = opt.forces == "on"

    def step():
        batch.run(3, forces=forces)

    tuned = None
    if not opt.no_autotune:
        # launch shape of the batch's point kernels: one workgroup per unit
        # or persistent launches over device work queues, the faster kept
        t0 = time.time()
        grid.set_precision(grid.FP64)
        queue = batch.autotune(3, forces=forces)
        tuned = {"seconds": round(time.time() - t0, 2), "work_queue": queue}
        log(f"{name}: autotuned in {tuned['seconds']} s (work queue {queue})")

    def timed(mode):
        grid.set_precision(mode)
        for _ in range(max(opt.warmup, 1)):
            step()
        grid.synchronize()
        grid.reset_timing()
        grid.set_timing(True)
        ms = timed_steps(step, opt.steps)
        grid.set_timing(False)
        return ms

    ms_step = timed(grid.FP64)
    log(f"{name}: batch steps of {ms_step:.3f} ms")
    kt = {k: grid.kernel_time_ms(i) for i, k in enumerate(grid.KERNEL_NAMES)}
    ref = [batch.outputs(i) for i in range(nlists)]
    out = {
        "metric": METRIC,
        "value": round(2.0 * points / (ms_step * 1e-3) / 1e9, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": opt.steps,
        "warmup": opt.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded cp2k_amd.synth.batch_config, random density/potential)",
        "config": {"workload": WORKLOADS.get(name, name), "config_name": name, "lists": nlists,
                   "ntasks": ntasks, "task_points": points, "task_points_by_lp": by_lp,
                   "tasks_per_s": round(2 * ntasks / (ms_step * 1e-3), 1),
                   "forces": forces, "generate_s": round(gen_s, 2), "autotune": tuned},
        "kernel_ms_per_step": {k: round(v[0] / max(opt.steps, 1), 4) for k, v in kt.items()},
        "cpu_baseline": None,
    }
    # the same work list by list, unbatched (device entry points, no graph)
    bufs = []
    for g, (a, p, V) in zip(gts, lists):
        bufs.append((grid.DeviceArray.from_numpy(p), [grid.DeviceArray(s) for s in a.grid_shapes()],
                     [grid.DeviceArray.from_numpy(v) for v in V], grid.DeviceArray(a.buffer_size)))

    def unbatched():
        for g, (p_d, r_d, v_d, h_d) in zip(gts, bufs):
            g.collocate_dev(100, p_d, r_d)
            g.integrate_dev(v_d, h_d)

    unbatched()
    ms_unb = timed_steps(unbatched, max(1, opt.steps // 2))
    log(f"{name}: unbatched steps of {ms_unb:.3f} ms")
    out["unbatched_ms_per_step"] = round(ms_unb, 3)
    out["batch_speedup"] = round(ms_unb / ms_step, 3)
    for p_d, r_d, v_d, h_d in bufs:
        for x in [p_d, h_d] + r_d + v_d:
            x.free()
    if not opt.no_other:
        ms32 = timed(grid.FP32_MIXED)
        got = [batch.outputs(i) for i in range(nlists)]
        grid.set_precision(grid.FP64)
        worst = {}
        for g_, r_ in zip(got, ref):
            e = precision.report(rho=(g_[0], r_[0]), hab=(g_[1], r_[1]))
            for k in ("rho_linf", "rho_l2", "hab_linf"):
                worst[k] = max(worst.get(k, 0.0), e[k])
        out["fp32_mixed"] = {
            "value": round(2.0 * points / (ms32 * 1e-3) / 1e9, 3),
            "unit": "Gpts/s",
            "ms_per_step": round(ms32, 3),
            "speedup_vs_fp64": round(ms_step / ms32, 3),
            "errors_vs_fp64_worst_list": {k: float(f"{v:.4g}") for k, v in worst.items()},
            "within_bars": bool(worst["rho_linf"] <= precision.RHO_LINF_BAR),
        }
    batch.free()
    return out


def cpu_child(opt, name):
    """The CPU baseline of workload `name`, run by a child process of this
    script (--cpu-only) with its own OpenMP environment."""
    import subprocess
    cmd = [sys.executable, str(pathlib.Path(__file__).resolve()), "--cpu-only", "--config", name,
           "--cpu-seconds", str(opt.cpu_seconds), "--forces", opt.forces, "--virial", opt.virial]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        log(f"CPU baseline failed (rc {r.returncode}): {r.stderr[-500:]}")
        return None
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    return json.loads(lines[-1]) if lines else None


def main():
    opt = parse()
    if opt.cpu_only:
        from cp2k_amd import synth
        mode = mode_of(opt, opt.config)
        args, pab, V = synth.config(opt.config, seed=0)
        print(json.dumps(cpu_baseline(args, pab, V, opt.cpu_seconds, mode["forces"],
                                      mode["virial"])))
        return
    if opt.rocprof_summary:
        print(rocprof_summary(opt.rocprof_summary, opt.calls, opt.copy_calibration or 512))
        return
    if opt.gpus != 1:
        print(f"bench.py measures one GPU only (this backend has no multi-GPU path); "
              f"got --gpus {opt.gpus}", file=sys.stderr)
        sys.exit(2)
    res = measure(opt, opt.config, cpu=not opt.no_cpu, abi=not opt.no_abi)
    if opt.secondary and not opt.config.startswith("batch"):
        res["secondary"] = {}
        for name in [x for x in opt.secondary.split(",") if x and x != opt.config]:
            r = measure(opt, name, cpu=False, abi=False)
            res["secondary"][name] = {k: r[k] for k in (
                "value", "unit", "ms_per_step", "dtype", "config", "kernel_ms_per_step",
                "roofline", "roofline_calls", "fp64", "fp32mixed", "fp32_mixed_vs_fp64")
                if k in r}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
