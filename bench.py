#!/usr/bin/env python3
"""Throughput of batched GFN2-xTB energy+forces on one MI355X.

One step = one ``xtbloom_compute`` (energy + forces + charges, FRESH start,
default SCC policy: charge 1e-6, energy 1e-8, at most 250 iterations, Broyden
8/0.4, 300 K) over the 512-system ragged alkane batch (CnH2n+2, n ~ U{2..38},
half of them with a terminal OH or NH2), with every input and output buffer
resident in HBM (device pointers through the C ABI).  XTBLOOM_AMD_LANES > 1
runs the batch as that many concurrent lanes (contiguous slices on their own
streams); the default is one.  The warmup calls record HIP events around every launch (the per-family
breakdown); the timed calls record them around the launches of the dominant
family only, on the stream each launch runs on, for the roofline entry.  The CPU baseline is
the FP64 C restatement run one system per thread on a bounded, size-stratified
sample of the same batch; the same sample gives the parity figures.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("systems/sec GFN2 energy+forces, ragged batch; speedup vs plain FP64 HIP at "
          "1e-6 Eh/bohr")
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def to_device(batch, torch, dev):
    t = dict(
        atom_offsets=torch.from_numpy(batch.atom_offsets).to(dev),
        numbers=torch.from_numpy(batch.atomic_numbers).to(dev),
        positions=torch.from_numpy(batch.positions).to(dev),
        charges=torch.from_numpy(batch.molecular_charges).to(dev),
        uhf=torch.from_numpy(batch.unpaired_electrons).to(dev))
    if batch.point_offsets is not None:
        t.update(point_offsets=torch.from_numpy(batch.point_offsets).to(dev),
                 point_positions=torch.from_numpy(batch.point_positions).to(dev),
                 point_values=torch.from_numpy(batch.point_charges).to(dev),
                 point_gammas=torch.from_numpy(batch.point_gammas).to(dev))
    return t


WORKLOAD_TEXT = {
    "ragged512": "ragged alkanes CnH2n+2, n~U{2..38}, p=0.5 terminal OH/NH2",
    "c20h42_128": "uniform C20H42 conformers (62 atoms, 122 AOs), torsions N(180, 15 deg)",
}


def cpu_baseline(batch, gpu_outs, sample_stride, opts=None):
    """FP64 C restatement on a strided sample; returns (baseline dict, {path: parity dict})."""
    import numpy as np

    from oracle import oracle

    idx = list(range(0, batch.batch_size, sample_stride))
    sub = batch.subset(idx)
    threads = host_threads()
    oracle.lib()
    t0 = time.perf_counter()
    pcs = {}
    if sub.point_offsets is not None:
        pcs = dict(point_offsets=sub.point_offsets, pc_xyz=sub.point_positions,
                   pc_q=sub.point_charges, pc_gamma=sub.point_gammas)
    ref = oracle.batch(sub.atom_offsets, sub.atomic_numbers, sub.positions,
                       sub.molecular_charges, sub.unpaired_electrons, opts=opts, threads=threads,
                       **pcs)
    wall = time.perf_counter() - t0
    parity = {}
    for name, gpu_out in gpu_outs.items():
        e = gpu_out["energies"].cpu().numpy()
        f = gpu_out["forces"].cpu().numpy()
        q = gpu_out["charges"].cpu().numpy()
        it = gpu_out["scc_iterations"].cpu().numpy()
        de = df = dq = 0.0
        dit = 0
        for k, i in enumerate(idx):
            a0, a1 = batch.atom_offsets[i], batch.atom_offsets[i + 1]
            b0, b1 = sub.atom_offsets[k], sub.atom_offsets[k + 1]
            de = max(de, abs(e[i] - ref["energies"][k]))
            df = max(df, float(np.abs(f[a0:a1] + ref["gradient"][b0:b1]).max()))
            dq = max(dq, float(np.abs(q[a0:a1] - ref["charges"][b0:b1]).max()))
            dit = max(dit, abs(int(it[i]) - int(ref["iterations"][k])))
        parity[name] = {"systems": len(idx), "max_abs_dE": de, "max_abs_dF": df,
                        "max_abs_dq": dq, "max_abs_diter": dit}
    atoms = int(sub.atom_offsets[-1])
    base = {"value": len(idx) / wall, "unit": "systems/s", "cores": threads, "kind": "port",
            "sample": f"{len(idx)} of {batch.batch_size} systems (stride {sample_stride}, "
                      f"{atoms} atoms), FP64 C restatement, one system per thread, "
                      f"{wall:.2f} s wall"}
    return base, parity


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ragged512")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-stride", type=int, default=2,
                    help="CPU baseline / parity sample: every k-th system")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=None, help="override the workload's batch size")
    ap.add_argument("--charge-tol", type=float, default=1e-6,
                    help="SCC charge tolerance (the ABI default; the reference's published "
                         "cross-engine numbers use 1e-4)")
    ap.add_argument("--energy-tol", type=float, default=1e-8,
                    help="SCC energy tolerance (the ABI default; published numbers use 1e-6)")
    ap.add_argument("--scc-start", choices=("fresh", "warm"), default="fresh",
                    help="warm: the warmup call converges FRESH, every timed call starts WARM "
                         "from the previous call's converged state")
    ap.add_argument("--warm-step", type=float, default=0.0,
                    help="with --scc-start warm: Gaussian displacement (bohr) of every atom "
                         "before each timed call (0: the same geometry)")
    ap.add_argument("--path", choices=("both", "plain", "batched"), default="both",
                    help="execution paths timed: the headline value is the batched path, "
                         "vs_baseline its ratio to the plain FP64 path")
    ap.add_argument("--split32", action="store_true",
                    help="also time the batched path with split-precision FP32 density products")
    ap.add_argument("--plan", action="store_true",
                    help="timed calls through a fixed-topology plan (no per-launch events: "
                         "the roofline entry then comes from the warmup calls)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the batched GFN2 path runs every system of a batch "
              "on a single device (multi-GPU runs shard systems across independent processes "
              "and are not part of this benchmark)", file=sys.stderr)
        sys.exit(2)

    import torch

    from xtbloom_amd import interface, library, workloads

    library.load()
    over = {} if args.batch is None else {"batch": args.batch}
    batch = workloads.make(args.workload, seed=args.seed, **over)
    tol = dict(charge_tolerance=args.charge_tol, energy_tolerance=args.energy_tol)
    dev = torch.device("cuda", 0)
    inp = to_device(batch, torch, dev)
    ctx = interface.Context("hip", 0)
    flags = library.COMPUTE_ENERGY | library.COMPUTE_FORCES | library.COMPUTE_ATOMIC_CHARGES
    out = None

    def new_agg():
        return {name: dict(launches=0, ms=0.0, flops=0.0, bytes=0.0)
                for name in library.KERNEL_FAMILIES}

    def add(agg, stats):
        for name, st in stats.items():
            for k in ("launches", "ms", "flops", "bytes"):
                agg[name][k] += st[k]

    start = library.SCC_START_WARM if args.scc_start == "warm" else library.SCC_START_FRESH
    geoms = [inp["positions"]] * args.steps
    if args.scc_start == "warm" and args.warm_step > 0.0:
        gen = torch.Generator(device=dev).manual_seed(args.seed)
        geoms, cur = [], inp["positions"]
        for _ in range(args.steps):
            cur = cur + args.warm_step * torch.randn(cur.shape, generator=gen, device=dev,
                                                     dtype=cur.dtype)
            geoms.append(cur)

    def run_path(path, stats, timed_stats, split32=False):
        """Warmup and timed calls of one execution path (XTBLOOM_AMD_PATH)."""
        os.environ["XTBLOOM_AMD_PATH"] = path
        os.environ["XTBLOOM_AMD_SPLIT32"] = "1" if split32 else "0"
        out = None
        # warmup: with stats, events around every launch (the per-family breakdown)
        wagg = new_agg()
        for _ in range(max(args.warmup, 1)):
            out, timing = interface.compute_tensors(ctx, **inp, flags=flags, outputs=out,
                                                    kernel_stats=True if stats else None, **tol)
            if stats:
                add(wagg, timing["kernel_stats"])
        torch.cuda.synchronize()
        dom = max(wagg, key=lambda k: wagg[k]["ms"]) if stats else None
        agg = new_agg()
        rounds = 0
        device_ms = 0.0
        plan = None
        if args.plan:
            plan = interface.Plan(ctx, batch.atom_offsets, batch.atomic_numbers, batch.positions,
                                  batch.molecular_charges, batch.unpaired_electrons, flags=flags,
                                  **tol)
            # the plan's own checkpoint for WARM calls
            out, _ = interface.compute_tensors(ctx, **inp, flags=flags, outputs=out, plan=plan,
                                               **tol)
        # the timed loop sums iteration counts on the device: load torch's
        # reduction kernel outside the timed region
        iter_sum = 0 * int(out["scc_iterations"].sum())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            step_inp = dict(inp, positions=geoms[k])
            out, timing = interface.compute_tensors(
                ctx, **step_inp, flags=flags, outputs=out,
                kernel_stats=[dom] if (timed_stats and not plan) else None,
                scc_start_mode=start, plan=plan, **tol)
            if timed_stats and not plan:
                add(agg, timing["kernel_stats"])
            rounds = max(rounds, timing["scc_rounds"])
            device_ms += timing["total_ms"]
            if args.scc_start == "warm":
                iter_sum += int(out["scc_iterations"].sum())
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if plan:
            plan.close()
        res = dict(out={k: v.clone() for k, v in out.items()}, ms_per_step=1000.0 * wall / args.steps,
                   wagg=wagg, agg=agg, dom=dom, rounds=rounds, device_ms=device_ms,
                   iter_sum=iter_sum, plan=bool(plan), timed_stats=timed_stats)
        res["value"] = batch.batch_size / (res["ms_per_step"] / 1000.0)
        return res

    # the plain FP64 path (per-launch events on its dominant family give the
    # roofline entry), then the batched path timed without events
    runs = {}
    if args.path in ("both", "plain"):
        runs["plain"] = run_path("plain", True, True)
    if args.path in ("both", "batched"):
        # timed without events (events would split the captured rounds)
        runs["batched"] = run_path("batched", args.path == "batched", False)
    if args.split32:
        # the split-precision FP32 density products on the batched path (reported
        # beside the headline, which stays the FP64 batched path)
        runs["batched_split32"] = run_path("batched", False, False, split32=True)
    os.environ.pop("XTBLOOM_AMD_SPLIT32", None)
    head = runs["batched"] if "batched" in runs else runs["plain"]
    ref_run = runs["plain"] if "plain" in runs else runs["batched"]
    out = head["out"]
    ms_per_step = head["ms_per_step"]
    value = head["value"]
    status = out["per_system_status"].cpu().numpy()
    wagg, agg, dom, plan = ref_run["wagg"], ref_run["agg"], ref_run["dom"], ref_run["plan"]
    from_warmup = plan or not ref_run["timed_stats"]
    rounds, device_ms, iter_sum = head["rounds"], head["device_ms"], head["iter_sum"]

    # roofline of the kernel family with the largest device time
    d = wagg[dom] if from_warmup else agg[dom]
    avg_ms = d["ms"] / max(d["launches"], 1)
    if d["flops"] > 0.0:
        per_launch = d["flops"] / d["launches"]
        achieved = per_launch / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None}
    else:
        per_launch = d["bytes"] / max(d["launches"], 1)
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    # HBM bytes per launch of the dominant kernel from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated on a copy)
    tpath = os.path.join(ROOT, "profiles", f"hbm_traffic_{args.workload}_fp64.json")
    if os.path.exists(tpath) and args.scc_start == "fresh" and args.batch is None:
        with open(tpath) as fh:
            tr = json.load(fh)
        if tr.get("kernel") == dom:
            roof["traffic"] = tr["traffic_bytes_per_launch"]
            roof["traffic_source"] = f"profiles/hbm_traffic_{args.workload}_fp64.json (rocprofv3 --pmc)"
    roof.update(kernel=dom, launches_per_step=d["launches"] / (max(args.warmup, 1) if from_warmup else args.steps),
                avg_launch_ms=avg_ms, algorithmic_per_launch=per_launch)
    roof["source"] = (("warmup calls (" + ("plan calls" if plan else "graph-captured rounds")
                       + " record no events)" if from_warmup else
                       "timed calls, events on this family only") +
                      (" of the plain path (the batched path runs the same kernels inside "
                       "captured graphs, without events)" if "batched" in runs and "plain" in runs
                       else ""))
    nw = max(args.warmup, 1)
    total_kernel_ms = sum(v["ms"] for v in wagg.values())
    kernels = {k: {"ms_per_step": v["ms"] / nw, "launches_per_step": v["launches"] / nw,
                   "share": v["ms"] / total_kernel_ms if total_kernel_ms else 0.0}
               for k, v in sorted(wagg.items(), key=lambda kv: -kv[1]["ms"]) if v["launches"]}

    cpu = parity = None
    path_parity = {}
    if not args.no_cpu and args.scc_start == "fresh":
        from oracle import oracle as _o
        cpu, path_parity = cpu_baseline(batch, {name: r["out"] for name, r in runs.items()},
                                        args.cpu_stride, _o.options(**tol))
        parity = path_parity.get("batched", path_parity.get("plain"))
    n_atoms = int(batch.atom_offsets[-1])
    nao = workloads_nao(batch)
    n_ao = int(sum(nao))
    # whole-path FP64-equivalent work, F(n, k) = (25/3) k n^3 + (7/3) n^3 per system
    # with k its SCC iterations (Cholesky, per-iteration TRSM, tridiagonalization,
    # back-transformation, C = XY and P, and W once), over the wall time of a step
    iters = out["scc_iterations"].cpu().numpy()
    fpath = sum((25.0 / 3.0) * float(k) * n ** 3 + (7.0 / 3.0) * n ** 3 for n, k in zip(nao, iters))
    path_tflops = fpath / (ms_per_step * 1e-3) / 1e12
    line = {
        "metric": METRIC, "value": value, "unit": "systems/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak",
        # the metric's own baseline: the plain all-FP64 HIP path on the same batch
        "vs_baseline": (runs["batched"]["value"] / runs["plain"]["value"]
                        if "batched" in runs and "plain" in runs else None),
        "dtype": "f64",
        "data": "synthetic (seeded alkane generator)",
        "config": {"workload": f"{args.workload}: {batch.batch_size} "
                               f"{WORKLOAD_TEXT.get(args.workload, 'systems')}, seed {args.seed}",
                   "systems": batch.batch_size, "atoms": n_atoms, "aos": n_ao,
                   "scc": f"{args.scc_start.upper()}"
                          + (f" (step {args.warm_step:g} bohr)" if args.scc_start == "warm" else "")
                          + f", charge {args.charge_tol:g}, energy {args.energy_tol:g}, max 250, "
                          "Broyden 8/0.4, 300 K",
                   "path": ("batched (graph-captured SCC rounds)" if "batched" in runs
                            else "plain FP64 HIP") + (", fixed-topology plan" if plan else ""),
                   "inputs": "device-resident (ROCm buffers)",
                   "lanes": lanes_for_batch(batch.batch_size)},
        "roofline": roof,
        "paths": {name: {"value": r["value"], "ms_per_step": r["ms_per_step"],
                         "parity_vs_oracle": path_parity.get(name)} for name, r in runs.items()},
        "path_fp64_equivalent": {"flops_per_step": fpath, "achieved": path_tflops,
                                 "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                 "frac": path_tflops / FP64_PEAK_TFLOPS},
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
        "scc_rounds": rounds,
        "mean_scc_iterations": (iter_sum / (args.steps * batch.batch_size)
                                if args.scc_start == "warm" else float(iters.mean())),
        "failed_systems": int((status != 0).sum()),
        "device_ms_per_step": device_ms / args.steps,
        "kernels": kernels,
        "kernels_source": "warmup calls, events around every launch, summed over lanes",
    }
    ctx.close()
    print(json.dumps(line), flush=True)


def lanes_for_batch(batch_size):
    # concurrent slices the library runs a batch in (XTBLOOM_AMD_LANES, default 1,
    # at least 32 systems per slice)
    want = min(max(int(os.environ.get("XTBLOOM_AMD_LANES", "1") or 1), 1), 4)
    return max(1, min(want, batch_size // 32))


def workloads_nao(batch):
    # 1 AO for H, 4 for C/N/O (the only elements of the generated workloads)
    import numpy as np
    per_atom = np.where(batch.atomic_numbers == 1, 1, 4)
    return [int(per_atom[batch.atom_offsets[i]:batch.atom_offsets[i + 1]].sum())
            for i in range(batch.batch_size)]


if __name__ == "__main__":
    main()
