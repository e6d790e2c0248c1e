"""Throughput of the MI355X grid backend on one GPU.

One step = one grid_collocate_task_list (density -> all grid levels) plus one
grid_integrate_task_list (potential -> Hamiltonian blocks) over the same task
list, with inputs and outputs resident in HBM (the library's *_dev entry
points).  value = task-points processed per second (collocate + integrate),
where a task-point is one (task, grid point) pair inside the reference's
truncation region, counted on the device with the kernels' own membership.

Default workload: 64 rigid H2O in a 12.4 A cube, 4 grid levels (128/72/42/24),
DZVP-like contracted s/p/d shells, FP64 (the plain FP64 GPU collocate+
integrate configuration).  Synthetic, seeded (cp2k_amd/synth.py).

Also reported: the roofline of the dominant kernel (HBM-bound, algorithmic
bytes per launch / average launch time from HIP events recorded on the
library's launch stream inside the timed region) and a CPU baseline: CP2K's
own OpenMP cpu-backend kernels (oracle/_ref) on a bounded sample of the same
workload, on this host's cores.
"""
import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--config", default="h2o64")
    p.add_argument("--precision", choices=["fp64", "fp32mixed"], default="fp64")
    p.add_argument("--forces", action="store_true", help="integrate also forces+virial")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="before the steps, run 3 device-to-device copies of MIB MiB: "
                        "a plain copy kernel of known size that scales the HBM "
                        "byte counters of a rocprofv3 --pmc run")
    return p.parse_args()


def algorithmic_bytes(args, points, which):
    """HBM bytes per launch of the point kernel:
    collocate: 16 B per task-point (FP64 grid read + write) + per task its
               Cartesian density sub-block and task record;
    integrate: 8 B per task-point (grid read) + per task hab sub-block and
               task record (+ density sub-block for forces)."""
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
        return 16.0 * points + sub + record
    return 8.0 * points + sub + record


def cpu_baseline(args, pab, V, seconds):
    """CP2K's cpu-backend kernels (grid_cpu_*_pgf_product, oracle/_ref) under
    an OpenMP block loop, on a random sample of the workload's blocks sized to
    roughly `seconds` of CPU time."""
    import dataclasses

    import numpy as np

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

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    frac = 0.01
    probe = O.TaskList(subset(frac))
    t0 = time.perf_counter()
    probe.collocate(100, pab, kernel="ref")
    probe.integrate(V, kernel="ref")
    dt = max(time.perf_counter() - t0, 1e-3)
    frac = min(1.0, frac * seconds / dt)
    tl = O.TaskList(subset(frac))
    pts = tl.task_points()
    t0 = time.perf_counter()
    tl.collocate(100, pab, kernel="ref")
    t1 = time.perf_counter()
    tl.integrate(V, kernel="ref")
    t2 = time.perf_counter()
    return {"value": round(2 * pts / (t2 - t0) / 1e9, 4), "unit": "Gpts/s",
            "cores": threads, "kind": "reference",
            "sample": (f"{tl.a.ntasks} of {args.ntasks} tasks ({100 * frac:.1f}% of blocks), "
                       f"{pts:.3e} task-points; CP2K cpu-backend kernels "
                       f"grid_cpu_collocate/integrate_pgf_product, OpenMP over blocks; "
                       f"collocate {t1 - t0:.2f} s, integrate {t2 - t1:.2f} s")}


def main():
    opt = parse()
    if opt.gpus != 1:
        print(f"bench.py measures one GPU only (this backend has no multi-GPU path); "
              f"got --gpus {opt.gpus}", file=sys.stderr)
        sys.exit(2)
    from cp2k_amd import grid, synth

    grid.set_precision(grid.FP64 if opt.precision == "fp64" else grid.FP32_MIXED)
    t0 = time.time()
    args, pab, V = synth.config(opt.config, seed=0)
    gen_s = time.time() - t0
    gt = grid.GridTaskList(args)
    points = gt.task_points()
    pab_d = grid.DeviceArray.from_numpy(pab)
    rho_d = [grid.DeviceArray(s) for s in args.grid_shapes()]
    V_d = [grid.DeviceArray.from_numpy(v) for v in V]
    hab_d = grid.DeviceArray(args.buffer_size)
    f_d = grid.DeviceArray((args.natoms, 3)) if opt.forces else None
    v_d = grid.DeviceArray((3, 3)) if opt.forces else None

    def step():
        gt.collocate_dev(100, pab_d, rho_d)
        gt.integrate_dev(V_d, hab_d, pab=pab_d if opt.forces else None, forces=f_d,
                         virial=v_d)

    if opt.copy_calibration:
        n = opt.copy_calibration * (1 << 20) // 8
        src, dst = grid.DeviceArray((n,)), grid.DeviceArray((n,))
        for _ in range(3):
            dst.copy_from(src)
        src.free()
        dst.free()

    for _ in range(opt.warmup):
        step()
    grid.synchronize()
    grid.reset_timing()
    grid.set_timing(True)
    grid.synchronize()
    t0 = time.perf_counter()
    for _ in range(opt.steps):
        step()
    grid.synchronize()
    t1 = time.perf_counter()
    grid.set_timing(False)
    ms_step = (t1 - t0) * 1e3 / opt.steps
    kt = {k: grid.kernel_time_ms(i) for i, k in enumerate(
        ["collocate", "integrate", "decontract", "contract", "fill", "reduce"])}
    value = 2.0 * points / (ms_step * 1e-3) / 1e9
    # dominant kernel
    dom = max(("collocate", "integrate"), key=lambda k: kt[k][0])
    ms, n = kt[dom]
    avg_ms = ms / max(n, 1)
    abytes = algorithmic_bytes(args, points, dom)
    achieved = abytes / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": "collocate+integrate Gpts/s & HBM GB/s; FP32-mixed vs FP64 at <=2e-6 rel rho err",
        "value": round(value, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": opt.steps,
        "warmup": opt.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if opt.precision == "fp64" else "f32-mixed(f64-accumulate)",
        "data": "synthetic (seeded cp2k_amd.synth, random density/potential)",
        "config": {
            "workload": {"h2o64": "64 H2O, 12.4 A cube, 4 levels 128/72/42/24, DZVP-like l<=2",
                         }.get(opt.config, opt.config),
            "config_name": opt.config,
            "ntasks": args.ntasks,
            "task_points": points,
            "levels": [int(n[0]) for n in args.npts_local],
            "forces": bool(opt.forces),
            "tasks_per_s": round(2 * args.ntasks / (ms_step * 1e-3), 1),
        },
        "kernel_ms_per_step": {k: round(v[0] / max(opt.steps, 1), 4) for k, v in kt.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes_per_launch": abytes,
            "avg_launch_ms": round(avg_ms, 4),
        },
        "cpu_baseline": None,
    }
    if not opt.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, pab, V, opt.cpu_seconds)
    out["config"]["generate_s"] = round(gen_s, 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
