"""Throughput benchmark of the FDTD time step on one MI355X.

Workload (default): plain 2nd-order Yee, float32, 1024^3 cells, 200 seeded
non-overlapping dielectric spheres (radius U[8,40], eps_r U[1,12], background 1),
12-cell CPML on all faces (m=4, R=1e-6), Ez point source at the centre driven by the
Gaussian-modulated pulse, Ez probe; synthetic data.  One step = E update + CPML-E +
source + H update + CPML-H (+ probe gather).

    python bench.py [--steps K] [--warmup W] [--engine fused|kemp|slab] [--n 1024]
        (default engine fused: single-pass interior sweep + CPML shell passes, CE as
         uint8 material indices; kemp: the in-place drop-in kernels)
    python bench.py --workload s4 [--n 512]   (4th-order symplectic, 100 spheres, CPML 10,
                                              Courant 0.9 x stability limit)
    python bench.py --workload batch [--steps 300] [--tune]   (64 uneven Yee+CPML 8
                       instances in one batched step loop vs the same instances run
                       one after another)
    python bench.py --workload adjoint   (160^3 design problem: forward 800 steps with
                       checkpoints every 50, adjoint gradient on the 64^3 box; reports
                       backward / forward time)
    python bench.py --workload pbc   (the 1024^3 sphere workload with periodic x and y and
                       CPML on the z faces; steps timed without and with the periodic
                       boundaries)
    python bench.py --workload tfsf  (1024^3, CPML 12, a TFSF plane wave along x on the box
                       [64, n-65]^3 around the sphere dielectric; steps timed without
                       and with the plane wave)
    python bench.py --workload drude (1024^3, CPML 12, a Drude box [256, 767]^3 (wp 0.5,
                       gamma 0.05) inside the sphere dielectric; steps timed without and
                       with the Drude medium)
        (pbc, tfsf, drude: fused engine by default, --engine kemp for the in-place one)
    python bench.py --workload snapshot   (the 1024^3 step with an Ez plane saved every
                       10 steps to gzip HDF5 files, against the same steps without; and
                       one whole-grid Ez snapshot: its cost to the step loop and the
                       time until its file is complete)
    python bench.py --copy-calibration MiB   (plain device copy of known size, for
                                              scaling the HBM byte counters)

Prints one JSON line: Gcell-updates/s over K timed steps (inputs resident in HBM),
the dominant kernel's roofline (algorithmic bytes / its average launch time, from
HIP events around each launch on the kernel's stream) and a CPU baseline (the
build's CPU oracle, timed on a bounded sample on this host).  The default run also
carries the metric's other two legs as keys of the same line: "batched" (the
64-instance uneven batch against the same instances run one after another) and
"adjoint" (the 160^3 design problem: backward / forward time and the gradient's error
against the float64 references); --plain-only skips them.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s + % HBM peak, plain vs batched vs adjoint-grad, 1 MI355X"
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the dominant kernel (GB), from two rocprofv3 --pmc runs
# (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0: the factors a plain 1 GiB copy gives), keyed by
# (engine, n, npml, CE storage, --z-phase); profiles/update_e_yee_cpml12_1024cube_hbm_counters.txt,
# profiles/update_e_lut_yee_cpml12_1024cube_hbm_counters.txt and
# profiles/k_fused_sweep_yee_cpml12_1024cube_hbm_counters.txt (the interior sweep; the
# automatic line phase is 16 there, the sweep box z [16, 1008), CE read as uint16 triple
# indices; at phase 0 the box is [32, 992) and the measurement predates the triple indices).
MEASURED_TRAFFIC_GB = {("kemp", 1024, 12, "float32", None): 64.69,
                       ("kemp", 1024, 12, "uint8 index + table", None): 53.24,
                       ("fused", 1024, 12, "uint8 index + table", None): 52.36,
                       ("fused", 1024, 12, "uint8 index + table", 16): 52.36,
                       ("fused", 1024, 12, "uint8 index + table", 0): 51.83}
# The 4th-order E stage (k_s4_e_zm, CE as uint8 material indices), same method, keyed by
# (n, npml); profiles/k_s4_zmarch_symplectic_cpml10_512cube_hbm_counters.txt.
MEASURED_S4_TRAFFIC_GB = {(512, 10): 5.81}
# The batched step (both launches of one step, all instances), same method, keyed by
# (instances, cube edges drawn from); profiles/k_batch_uneven_64cubes_kernel_trace.txt.
MEASURED_BATCH_TRAFFIC_GB = {(64, "64,96,128,160,192"): 16.86}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--n", type=int, default=1024)
    p.add_argument("--npml", type=int, default=12)
    p.add_argument("--spheres", type=int, default=200)
    p.add_argument("--engine", default="fused", choices=("fused", "kemp", "slab"))
    p.add_argument("--pairs", action="store_true",
                   help="fused engine: advance steps in pairs (two-step interior sweep; slower on MI355X)")
    p.add_argument("--slab-planes", type=int, default=None)
    p.add_argument("--z-phase", type=int, default=None,
                   help="fused: 128-byte line phase (floats) of the field arrays (default: chosen from npml)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MiB")
    p.add_argument("--workload", default="yee", choices=("yee", "s4", "batch", "adjoint", "pbc", "tfsf", "drude", "snapshot"))
    p.add_argument("--instances", type=int, default=64)
    p.add_argument("--sizes", default="64,96,128,160,192", help="batch: cube edges to draw from")
    p.add_argument("--tune", action="store_true", help="batch: autotune launch options first")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-ce-lut", action="store_true",
                   help="kemp and fused engines: read CE as float arrays instead of uint8 indices + table")
    p.add_argument("--cpu-sample-n", type=int, default=256)
    p.add_argument("--plain-only", action="store_true",
                   help="default workload: time the 1024^3 step only (skip the batched and adjoint legs)")
    return p.parse_args()


def cpml_face_cells(n, npml):
    """Slab cells of the six faces, each face counted separately (edges included)."""
    return 6 * npml * n * n


def build_sim(args):
    import torch
    from fdtd_accelerate_amd import geometry
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.sources import gaussian_pulse

    n, npml = args.n, args.npml
    sim = YeeSim(n, n, n, npml=npml, coeff_use="e", engine=args.engine)
    lo, hi = (npml,) * 3, (n - npml,) * 3
    rmax = min(40.0, (n - 2 * npml) / 8)
    spheres = geometry.random_spheres(2024, args.spheres, (min(8.0, rmax / 2), rmax), (1.0, 12.0), lo, hi)
    geometry.paint_ce(sim.fields.ces, spheres, (n, n, n))
    c = n // 2
    sim.add_source("ez", (c, c, c), (c, c, c), gaussian_pulse())
    sim.add_probe("ez", (c + n // 4, c, c))
    if args.engine in ("kemp", "fused") and not args.no_ce_lut:
        sim.fields.compress_ce()
    torch.cuda.synchronize()
    return sim


def time_kemp(sim, steps, warmup):
    """In-place engine; events bracket the update_e kernel (the dominant launch) on the
    fields' stream inside the timed loop."""
    import torch
    f = sim.fields
    core = sim.core
    others_e = [i for i in f.instance_list if i is not core]
    for _ in range(warmup):
        f.update_e()
        f.update_h()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    t0.record(f.stream)
    for s in range(steps):
        evs[s][0].record(f.stream)
        core.update_e()
        evs[s][1].record(f.stream)
        for inst in others_e:
            inst.update_e()
        f.update_h()
    t1.record(f.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1)
    kern_ms = sum(a.elapsed_time(b) for a, b in evs) / steps
    if f.ce_mat is not None:
        return ms / steps, wall, kern_ms, "update_e (CE as uint8 index + table)", 39.0
    return ms / steps, wall, kern_ms, "update_e", 48.0


def time_fused(sim, steps, warmup, pairs=False, z_phase=None):
    """Fused engine: events bracket the whole run (all launches of the K steps); a second
    run of K steps brackets, inside the C loop, every interior-sweep launch (the dominant
    kernel) with timing-only HIP events on the fields' stream.
    Returns the sweep's algorithmic bytes per grid cell (bytes per swept cell x swept
    cells / grid cells), so that bytes x cells / sweep time is the sweep's rate."""
    import torch
    from fdtd_accelerate_amd.fused import FusedStepper
    st = FusedStepper(sim, step_pairs=pairs, z_phase=z_phase)
    st.run(warmup, None)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record(sim.fields.stream)
    st.run(steps, None)
    t1.record(sim.fields.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1)
    # second pass of the same K steps with timing-only events around each sweep launch
    kern_ms = st.run(steps, None, time_kernel=True)
    cells = sim.fields.nx * sim.fields.ny * sim.fields.nz
    sweep = st.sweep_cells()
    # algorithmic bytes on the per-cell basis of the plain Yee step: read 6 fields and 3 CE,
    # write 6 fields = 60 B (the sweep box holds no CPML cell); the kernel itself reads CE
    # as 3 B of material indices or a 2 B triple index, so it moves 51 or 50 B of those 60
    if st.pairs:
        # one launch advances its box two steps: state and CE read once, state written once
        name = (f"k_fused2 (two-step interior sweep {sweep} cells, 60 B/cell per launch = 30 B/cell per step "
                f"algorithmic, {st.bytes_per_cell:g} B/cell per launch read/written)")
        return ms / steps, wall, kern_ms, name, 30.0 * sweep / cells
    name = (f"k_fused (interior sweep {sweep} cells, field arrays at z line phase {st.z_phase}, 60 B/cell "
            f"algorithmic, {st.bytes_per_cell:g} B/cell read/written)")
    return ms / steps, wall, kern_ms, name, 60.0 * sweep / cells


def time_slab(sim, steps, warmup, slab_planes=None):
    """Slab engine (E pass + H pass per slab, CPML and sources fused); events bracket each
    step's launches on the fields' stream, the step is the timed unit."""
    import torch
    from fdtd_accelerate_amd.fused import SlabStepper
    st = SlabStepper(sim, slab_planes=slab_planes)
    st.run(warmup, None)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record(sim.fields.stream)
    kern_ms = st.run(steps, None, time_kernel=True)
    t1.record(sim.fields.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1)
    return ms / steps, wall, kern_ms, f"k_slab_e+k_slab_h (S={st.S})", None


def run_s4(args):
    """4th-order symplectic step: 512^3, 100 spheres (eps_r U[1,12], r U[8,40] scaled to
    the grid), 10-cell CPML, S = 0.9 S_max, Ez point source at the centre."""
    import torch
    from fdtd_accelerate_amd import geometry
    from fdtd_accelerate_amd.s4 import S4Sim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    n = args.n
    npml = 10 if args.npml == 12 else args.npml
    nsph = 100 if args.spheres == 200 else args.spheres
    rmax = min(40.0, (n - 2 * npml) / 8)
    spheres = geometry.random_spheres(2024, nsph, (min(8.0, rmax / 2), rmax), (1.0, 12.0), (npml,) * 3, (n - npml,) * 3)
    sim = S4Sim(n, n, n, npml=npml)
    sim.set_eps(geometry.eps_map_numpy((n, n, n), spheres))
    c = n // 2
    sim.add_source("ez", (c, c, c), (c, c, c), gaussian_pulse())
    for _ in range(args.warmup):
        sim.step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record(sim.stream)
    for k in range(args.steps):
        sim.step(events=evs[k])
    t1.record(sim.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1) / args.steps
    kern_ms = sum(a.elapsed_time(b) for a, b in evs) / args.steps
    cells = n ** 3
    step_bytes = 336.0 * cells + 32.0 * 4 * 6 * npml * n * n
    achieved = 48.0 * cells / (kern_ms * 1e-3) / 1e9
    return {
        "metric": METRIC, "value": round(cells / (ms * 1e-3) / 1e9, 3), "unit": "Gcell/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"4th-order symplectic FP32 {n}^3, {nsph} spheres, CPML {npml}, S=0.9*S_max={sim.S:.4f}, Ez point source",
                   "cells": cells},
        "hbm_frac_step": round(step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": "k_s4_e_zm (E stage)", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_S4_TRAFFIC_GB.get((n, npml)), "traffic_unit": "GB per launch",
                     "bytes_per_cell": 48.0, "bytes_basis": "read E, H, CE as float32, write E (the kernel reads CE as "
                     "3 B of material indices: 39 B moved)", "avg_launch_ms": round(kern_ms, 4)},
        "wall_s": round(wall, 3), "device": torch.cuda.get_device_name(0),
    }


def run_batch(args):
    """Batched plain Yee: 64 cubes with edges from {64,96,128,160,192}, own sphere maps,
    CPML 8, seeded source; batched step loop vs sequential per-instance runs."""
    import torch
    from fdtd_accelerate_amd import batch as B
    from fdtd_accelerate_amd.fused import SlabStepper
    sims = B.make_uneven_batch(seed=3, count=args.instances, sizes=tuple(int(v) for v in args.sizes.split(",")))
    cells = sum(s.fields.nx * s.fields.ny * s.fields.nz for s in sims)
    bt = B.BatchedYee(sims)
    tuned = None
    if args.tune:
        tuned = B.tune(bt)
    bt.time_steps(max(args.warmup, 1))
    wall0 = time.perf_counter()
    ms = bt.time_steps(args.steps)
    wall = time.perf_counter() - wall0
    # sequential: each instance alone, (a) slab engine with its C step loop (2 launches
    # per step), (b) the in-place engine driven step by step from Python
    st = torch.cuda.current_stream()
    steppers = [SlabStepper(s) for s in sims]
    for sp in steppers:
        sp.run(1, None)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(st)
    for sp in steppers:
        sp.run(args.steps, None)
    b.record(st)
    torch.cuda.synchronize()
    seq_slab_ms = a.elapsed_time(b) / args.steps
    ksteps = min(args.steps, 50)
    for s in sims:
        s.fields.update_e()
        s.fields.update_h()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    a.record(st)
    for s in sims:
        for _ in range(ksteps):
            s.fields.update_e()
            s.fields.update_h()
    b.record(st)
    torch.cuda.synchronize()
    seq_kemp_ms = max(a.elapsed_time(b), (time.perf_counter() - t0) * 1e3) / ksteps
    seq_best = min(seq_slab_ms, seq_kemp_ms)
    sizes = sorted(s.fields.nx for s in sims)
    step_bytes = sum(60.0 * s.fields.nx * s.fields.ny * s.fields.nz +
                     32.0 * (2 * s.cpml.npml * (s.fields.ny * s.fields.nz + s.fields.nx * s.fields.nz + s.fields.nx * s.fields.ny)
                             if s.cpml is not None else 0) for s in sims)
    return {
        "metric": METRIC, "value": round(cells / (ms * 1e-3) / 1e9, 3), "unit": "Gcell/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batched plain Yee FP32: {len(sims)} cubes, edges {dict((n, sizes.count(n)) for n in sorted(set(sizes)))}, "
                               f"10-40 spheres each, CPML 8, seeded Ez source", "cells": cells,
                   "use_graph": bt.use_graph, "remap_rows": bt.remap_rows},
        "batched_vs_sequential": round(seq_best / ms, 3),
        "sequential": {"slab_engine_ms_per_step": round(seq_slab_ms, 4),
                       "inplace_engine_ms_per_step": round(seq_kemp_ms, 4),
                       "inplace_steps_timed": ksteps},
        "tuning": None if tuned is None else {"best": list(tuned[0]), "ms": {f"{k[0]}/{k[1]}": round(v, 4) for k, v in tuned[1].items()}},
        "roofline": {"bound": "hbm", "kernel": bt.kernel_name, "basis": "60 B per cell + 32 B per CPML face-slab cell per step",
                     "achieved": round(step_bytes / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_BATCH_TRAFFIC_GB.get((args.instances, args.sizes)), "traffic_unit": "GB per step",
                     "algorithmic_gb": round(step_bytes / 1e9, 3), "ms_per_step": round(ms, 4)},
        "wall_s": round(wall, 3), "device": torch.cuda.get_device_name(0),
    }


def run_pbc(args):
    """Periodic boundaries: the sphere workload of the default run with periodic x and y
    (Pbc(fields, "xy")) and CPML only on the z faces, on the fused engine (default) or
    the in-place engine (--engine kemp).  The same grid is timed first without and then
    with the periodic boundaries."""
    import torch
    from fdtd_accelerate_amd import Pbc, geometry
    from fdtd_accelerate_amd.sim import YeeSim
    from fdtd_accelerate_amd.sources import gaussian_pulse
    n, npml = args.n, args.npml
    fused = args.engine == "fused"
    sim = YeeSim(n, n, n, npml=npml, coeff_use="e", cpml_faces=("", "", "fb"), engine="fused" if fused else "kemp")
    lo, hi = (npml,) * 3, (n - npml,) * 3
    rmax = min(40.0, (n - 2 * npml) / 8)
    spheres = geometry.random_spheres(2024, args.spheres, (min(8.0, rmax / 2), rmax), (1.0, 12.0), lo, hi)
    geometry.paint_ce(sim.fields.ces, spheres, (n, n, n))
    c = n // 2
    sim.add_source("ez", (c, c, c), (c, c, c), gaussian_pulse())
    sim.add_probe("ez", (c + n // 4, c, c))
    sim.fields.compress_ce()
    torch.cuda.synchronize()
    runner = (lambda s, k, w: time_fused(s, k, w, z_phase=args.z_phase)) if fused else time_kemp
    ms0, _, _, _, _ = runner(sim, args.steps, args.warmup)
    Pbc(sim.fields, "xy")
    ms, wall, kern_ms, kname, kbytes = runner(sim, args.steps, args.warmup)
    cells = n ** 3
    achieved = kbytes * cells / (kern_ms * 1e-3) / 1e9
    return {
        "metric": METRIC, "value": round(cells / (ms * 1e-3) / 1e9, 3), "unit": "Gcell/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain Yee FP32 {n}^3, {args.spheres} spheres, periodic x and y, CPML {npml} on the "
                               f"z faces, Ez point source", "engine": "fused" if fused else "kemp", "cells": cells},
        "ms_per_step_without_periodic": round(ms0, 4),
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_cell": kbytes, "avg_launch_ms": round(kern_ms, 4)},
        "wall_s": round(wall, 3), "device": torch.cuda.get_device_name(0),
    }


def run_tfsf(args):
    """TFSF plane wave: 1024^3 with CPML on all faces, the sphere dielectric inside the
    box [64, n-65]^3 (its faces in vacuum), a plane wave along x (wavelength 20 cells,
    polarization angle 0.3) entering through all six faces; fused engine (default) or
    in-place (--engine kemp).  The same grid is timed without and then with the plane wave."""
    import torch
    from fdtd_accelerate_amd import Tfsf, geometry
    from fdtd_accelerate_amd.sim import YeeSim
    n, npml = args.n, args.npml
    lo_box, hi_box = 64, n - 65
    fused = args.engine == "fused"
    sim = YeeSim(n, n, n, npml=npml, coeff_use="e", engine="fused" if fused else "kemp")
    lo, hi = (lo_box + 8,) * 3, (hi_box - 8,) * 3
    rmax = min(40.0, (hi_box - lo_box) / 8)
    spheres = geometry.random_spheres(2024, args.spheres, (min(8.0, rmax / 2), rmax), (1.0, 12.0), lo, hi)
    geometry.paint_ce(sim.fields.ces, spheres, (n, n, n))
    sim.fields.compress_ce()
    torch.cuda.synchronize()
    runner = (lambda s, k, w: time_fused(s, k, w, z_phase=args.z_phase)) if fused else time_kemp
    ms0, _, _, _, _ = runner(sim, args.steps, args.warmup)
    Tfsf(sim.fields, (lo_box,) * 3, (hi_box,) * 3, ("fb", "fb", "fb"), 20.0, ("normal", "x"), 0.3)
    ms, wall, kern_ms, kname, kbytes = runner(sim, args.steps, args.warmup)
    cells = n ** 3
    achieved = kbytes * cells / (kern_ms * 1e-3) / 1e9
    return {
        "metric": METRIC, "value": round(cells / (ms * 1e-3) / 1e9, 3), "unit": "Gcell/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain Yee FP32 {n}^3, {args.spheres} spheres inside the TFSF box [{lo_box},{hi_box}]^3, "
                               f"CPML {npml} all faces, x-directed TFSF plane wave (wavelength 20, polarization 0.3)",
                   "engine": "fused" if fused else "kemp", "cells": cells},
        "ms_per_step_without_tfsf": round(ms0, 4),
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_cell": kbytes, "avg_launch_ms": round(kern_ms, 4)},
        "wall_s": round(wall, 3), "device": torch.cuda.get_device_name(0),
    }


def run_drude(args):
    """Drude dispersion: the default sphere workload (CPML on all faces) with a Drude box
    [n/4, 3n/4)^3 (plasma frequency 0.5, damping 0.05 per unit time) on the in-place
    engine (--engine kemp) or the fused engine (default).  The same grid is timed without
    and then with the Drude medium; the Drude update reads J, CJA, CJB (and, in place, E and
    CE) and writes J (and, in place, E): 28 B per component, 84 B per box cell."""
    import torch
    from fdtd_accelerate_amd import Drude
    a = argparse.Namespace(**vars(args))
    fused = args.engine == "fused"
    a.engine = "fused" if fused else "kemp"
    a.no_ce_lut = not fused  # the in-place Drude update reads the float CE arrays
    sim = build_sim(a)
    n = args.n
    p0, p1 = (n // 4,) * 3, (3 * n // 4 - 1,) * 3
    nd = (p1[0] - p0[0] + 1) ** 3
    cells = n ** 3
    if fused:
        # the sweep and shell E pass carry the Drude update: the sweep's 60 B per cell plus
        # 84 B per Drude cell (the box lies inside the sweep box)
        ms0, _, _, _, _ = time_fused(sim, args.steps, args.warmup, z_phase=args.z_phase)
        Drude(sim.fields, p0, p1, 0.5, 0.05)
        ms, wall, kern_ms, kname, kbytes = time_fused(sim, args.steps, args.warmup, z_phase=args.z_phase)
        kbytes += 84.0 * nd / cells
        kname = kname.replace("k_fused (", "k_fused with a Drude box (", 1)
    else:
        ms0, _, _, _, _ = time_kemp(sim, args.steps, args.warmup)
        Drude(sim.fields, p0, p1, 0.5, 0.05)
        ms, wall, kern_ms, kname, kbytes = time_kemp(sim, args.steps, args.warmup)
    achieved = kbytes * cells / (kern_ms * 1e-3) / 1e9
    return {
        "metric": METRIC, "value": round(cells / (ms * 1e-3) / 1e9, 3), "unit": "Gcell/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain Yee FP32 {n}^3, {args.spheres} spheres, CPML {args.npml} all faces, Ez point "
                               f"source, Drude box [{p0[0]},{p1[0]}]^3 (wp 0.5, gamma 0.05)",
                   "engine": "fused" if fused else "kemp", "cells": cells, "drude_cells": nd},
        "ms_per_step_without_drude": round(ms0, 4),
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_cell": kbytes, "avg_launch_ms": round(kern_ms, 4)},
        "wall_s": round(wall, 3), "device": torch.cuda.get_device_name(0),
    }


def run_snapshot(args):
    """Snapshots during the step loop: the default sphere workload on the fused engine,
    K steps without snapshots, then K steps with an Ez z-plane through the source saved
    every 10 steps (gzip HDF5 files, SaveFields), then one whole-grid Ez snapshot: the
    time it takes from the step loop (staging copy on the fields' stream) and until its
    file is complete (device-to-host copy and 16-thread deflate off the step loop)."""
    import shutil
    import tempfile
    import torch
    from fdtd_accelerate_amd import snapshot
    a = argparse.Namespace(**vars(args))
    a.engine = "fused"
    sim = build_sim(a)
    n, c = args.n, args.n // 2
    sim.max_tstep = 10 ** 6
    st = sim.fields.stream
    d = tempfile.mkdtemp(prefix="mifdtd_snap_") + "/"
    try:
        def timed(k):
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            wall0 = time.perf_counter()
            t0.record(st)
            sim.run(k)
            t1.record(st)
            torch.cuda.synchronize()
            sim.wait_savefields()
            return t0.elapsed_time(t1) / k, time.perf_counter() - wall0

        sim.run(args.warmup)
        ms0, _ = timed(args.steps)
        print(f"snapshot bench: {ms0:.3f} ms/step without snapshots", file=sys.stderr, flush=True)
        plane = sim.set_savefields("ez", (0, 0, c), (-1, -1, c), (0, -1, 10), d, nthreads=16)
        ms1, wall1 = timed(args.steps)
        nplane = len(plane.saved)
        print(f"snapshot bench: {ms1:.3f} ms/step with {nplane} plane snapshots", file=sys.stderr, flush=True)
        sim.savef_list.clear()
        whole = snapshot.SaveFields(sim.fields, "ez", (0, 0, 0), (-1, -1, -1), d, 99, 10 ** 6, nslots=1, nthreads=16)
        sim.run(10)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        wall0 = time.perf_counter()
        e0.record(st)
        whole.save_fields(sim.tstep)
        e1.record(st)
        sim.run(10)
        torch.cuda.synchronize()
        steps_wall = time.perf_counter() - wall0
        whole.wait()
        file_s = time.perf_counter() - wall0
        stage_ms = e0.elapsed_time(e1)
        path = whole.saved[-1]
        size = os.path.getsize(path)
        whole.close()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    cells = n ** 3
    return {
        "metric": METRIC, "value": round(cells / (ms1 * 1e-3) / 1e9, 3), "unit": "Gcell/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms1, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain Yee FP32 {n}^3, {args.spheres} spheres, CPML {args.npml} all faces, Ez point "
                               f"source, fused engine; Ez plane z={c} saved every 10 steps (gzip level 4 HDF5)",
                   "engine": "fused", "cells": cells},
        "ms_per_step_without_snapshots": round(ms0, 4),
        "plane_snapshots": nplane, "plane_loop_wall_s": round(wall1, 3),
        "whole_grid_snapshot": {"bytes": 4 * cells, "stage_copy_ms": round(stage_ms, 3),
                                "ten_steps_after_wall_s": round(steps_wall, 3), "file_complete_s": round(file_s, 2),
                                "file_bytes": size},
        "device": torch.cuda.get_device_name(0),
    }


ADJ_GOLD = os.path.join(ROOT, "tests", "golden", "adjoint_design160.npz")


def run_adjoint(args):
    """The 160^3 design problem: J(eps) forward (800 steps, checkpoints every 50) and dJ/deps on the
    64^3 design box by the adjoint (recompute + adjoint kernels); GPU times of both.  With the
    committed float64 reference data, the gradient's error against the float64 central
    differences at its 32 voxels and against the float64 gradient over the box."""
    import numpy as np
    import torch
    from fdtd_accelerate_amd.adjoint import DesignProblem
    p = DesignProblem()
    gold = np.load(ADJ_GOLD) if os.path.exists(ADJ_GOLD) else None
    seed = int(gold["eps_seed"]) if gold is not None else 4
    eps = np.random.default_rng(seed).uniform(1.0, 4.0, p.box[3:]).astype(np.float32)
    p.forward(eps)  # warm-up (allocations, first launches)
    p.backward()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    c = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    a.record(p.stream)
    J = p.forward(eps)
    b.record(p.stream)
    g = p.backward()
    c.record(p.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    fwd, bwd = a.elapsed_time(b), b.elapsed_time(c)
    n = p.n[0] * p.n[1] * p.n[2]
    check = None
    if gold is not None:
        gh = g.cpu().numpy().astype(np.float64)
        fd = gold["fd"]
        ref = gold["grad_ref"].astype(np.float64)
        check = {"fd_voxels": int(len(fd)), "fd_max_rel": float(np.max(np.abs(gh[tuple(gold["voxels"].T)] - fd) / np.abs(fd))),
                 "grad_rel_l2_vs_fp64": float(np.linalg.norm(gh - ref) / np.linalg.norm(ref)),
                 "J_rel_vs_fp64": float(abs(J - float(gold["J64"])) / float(gold["J64"]))}
    return {
        "metric": METRIC, "value": round(n * p.nsteps / (fwd * 1e-3) / 1e9, 3), "unit": "Gcell/s", "n_gpus": 1,
        "steps": p.nsteps, "warmup": 1, "ms_per_step": round(fwd / p.nsteps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "adjoint gradient: 160^3 Yee, CPML 10, 64^3 design box eps_r U[1,4], Ez line source on x=20 (y=80, all interior z), "
                               "J = sum of |E|^2 on x=140 over steps 601-800, checkpoints every 50", "cells": n},
        "forward_ms": round(fwd, 3), "backward_ms": round(bwd, 3), "backward_over_forward": round(bwd / fwd, 3),
        "J": J, "grad_norm": float(g.norm()), "check": check,
        "roofline": None,
        "wall_s": round(wall, 3), "device": torch.cuda.get_device_name(0),
    }


def copy_calibration(mib):
    """A plain device-to-device copy of ``mib`` MiB, 5 times: the reference kernel whose
    FETCH_SIZE / WRITE_SIZE counters (known bytes) scale the counters of the step kernels."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4, "launches": 5}))


def cpu_baseline(args):
    """The oracle (float32, OpenMP) on a bounded sample: an n_s^3 Yee+CPML grid with
    the same npml and a random-sphere CE map, timed over a few steps on this host."""
    import numpy as np
    from oracle import oracle as O
    from fdtd_accelerate_amd import geometry
    ns = args.cpu_sample_n
    npml = args.npml
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    spheres = geometry.random_spheres(7, 12, (4.0, 12.0), (1.0, 12.0), (npml,) * 3, (ns - npml,) * 3)
    ces = geometry.paint_ce_numpy((ns, ns, ns), ns, spheres)
    s = O.OracleSim(ns, ns, ns, 0, npml, np.float32, ces=ces)
    s.add_source("ez", (ns // 2,) * 3, (ns // 2,) * 3, O.pulse)
    s.step()
    steps = 0
    t0 = time.perf_counter()
    while True:
        s.step()
        steps += 1
        el = time.perf_counter() - t0
        if el > 10.0 or steps >= 1000:
            break
    gcells = ns ** 3 * steps / el / 1e9
    return {"value": round(gcells, 4), "unit": "Gcell/s", "cores": cores, "kind": "port", "label": "build CPU oracle",
            "sample": f"oracle float32 Yee+CPML({npml}) {ns}^3, 12 spheres, {steps} steps, {el:.1f} s, OpenMP {cores} threads"}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the package has no multi-GPU path "
              "(sharding would split the grid into x slabs with one-plane halo exchange)", file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.workload == "adjoint":
        print(json.dumps(run_adjoint(args)), flush=True)
        return
    if args.workload == "batch":
        if args.steps == 200:
            args.steps = 300
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.workload == "snapshot":
        print(json.dumps(run_snapshot(args)), flush=True)
        return
    if args.workload == "drude":
        print(json.dumps(run_drude(args)), flush=True)
        return
    if args.workload == "tfsf":
        print(json.dumps(run_tfsf(args)), flush=True)
        return
    if args.workload == "pbc":
        print(json.dumps(run_pbc(args)), flush=True)
        return
    if args.workload == "s4":
        if args.n == 1024:
            args.n = 512
        print(json.dumps(run_s4(args)), flush=True)
        return
    sim = build_sim(args)
    if args.engine == "slab":
        ms, wall, kern_ms, kname, kbytes = time_slab(sim, args.steps, args.warmup, args.slab_planes)
    else:
        if args.engine == "fused":
            ms, wall, kern_ms, kname, kbytes = time_fused(sim, args.steps, args.warmup, args.pairs, args.z_phase)
        else:
            ms, wall, kern_ms, kname, kbytes = time_kemp(sim, args.steps, args.warmup)
    n = args.n
    cells = n ** 3
    ce_storage = "uint8 index + table" if getattr(sim.fields, "ce_mat", None) is not None else "float32"
    gcells = cells / (ms * 1e-3) / 1e9
    step_bytes = 60.0 * cells + 32.0 * cpml_face_cells(n, args.npml)
    if kbytes is None:  # whole-step unit: algorithmic bytes of the step
        kbytes = step_bytes / cells
    achieved = kbytes * cells / (kern_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(gcells, 3),
        "unit": "Gcell/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain Yee FP32 {n}^3, {args.spheres} spheres, CPML {args.npml} all faces, Ez point source",
                   "engine": args.engine + (" (step pairs)" if args.pairs else ""), "cells": cells,
                   "ce_storage": ce_storage},
        "hbm_frac_step": round(step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC_GB.get((args.engine + ("-pairs" if args.pairs else ""), n, args.npml, ce_storage,
                                                         args.z_phase)), "traffic_unit": "GB per launch",
                     "bytes_per_cell": round(kbytes, 3), "avg_launch_ms": round(kern_ms, 4)},
        "wall_s": round(wall, 3),
        "device": torch.cuda.get_device_name(0),
    }
    del sim
    torch.cuda.empty_cache()
    if not args.plain_only and args.n == 1024:
        out["batched"] = batched_leg(args)
        torch.cuda.empty_cache()
        out["adjoint"] = adjoint_leg(args)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


def batched_leg(args):
    """The metric's batched leg: the 64-instance uneven batch (300 steps) against the
    same instances run one after another."""
    import copy
    b = copy.copy(args)
    b.steps, b.warmup = 300, 5
    r = run_batch(b)
    print(f"batched leg: {r['ms_per_step']} ms/step, {r['batched_vs_sequential']}x sequential", file=sys.stderr, flush=True)
    return {"value": r["value"], "unit": r["unit"], "ms_per_step": r["ms_per_step"], "steps": r["steps"],
            "config": r["config"], "batched_vs_sequential": r["batched_vs_sequential"], "sequential": r["sequential"],
            "roofline": r["roofline"]}


def adjoint_leg(args):
    """The metric's adjoint leg: the 160^3 design problem, backward / forward time and the
    gradient's error against the float64 references."""
    r = run_adjoint(args)
    print(f"adjoint leg: backward/forward {r['backward_over_forward']}", file=sys.stderr, flush=True)
    return {"forward_ms": r["forward_ms"], "backward_ms": r["backward_ms"],
            "backward_over_forward": r["backward_over_forward"], "forward_gcells": r["value"],
            "config": r["config"], "check": r["check"]}


if __name__ == "__main__":
    main()
