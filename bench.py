"""MPCD particle-steps/s on one MI355X, with the HBM roofline and a CPU baseline.

Workload (default): periodic 256^3 cells, 10 solvent particles per cell
(167,772,160 particles), uniform positions, Maxwell-Boltzmann velocities at
kT = 1, SRD rotation by 130 degrees with the reference's constants, cell MBS
thermostat, h = 0.1, a new random grid shift every step; FP32-bulk form, and
the plain FP64 form beside it.  One step = collide + stream + rebin of every
particle (one launch of the tile-fused step kernel plus the overflow fix-up).

Prints ONE JSON line.  The roofline figure uses the step kernel's own average
duration measured with HIP events on the stream it runs on, against the
algorithmic bytes per particle-step: 56 B (FP32-bulk: 7 x 4 B read + written)
or 96 B (FP64: 6 x 8 B read + written).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multi-ciliated-microswimmers_amd"))

PEAK_GBS = 8000.0
BYTES = {"fp32bulk": 56, "fp64": 96}


def run_form(mcm, form, L, steps, warmup, h=0.1, seed=2, variant=-1, download=False):
    """Steps a periodic L^3 box of the given form; with download, reads the
    fluid back once after the timed steps (its download kernel reads every
    record once, the read pattern of the step kernel's loads)."""
    prec = mcm.FP32BULK if form == "fp32bulk" else mcm.FP64
    n = 10 * L ** 3
    with mcm.MPCD((L, L, L), precision=prec, seed=seed, tile_variant=variant) as m:
        m.create_fluid(n)
        for _ in range(warmup):
            m.step(h)
        m.sync()
        m.set_timing(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            m.step(h)
        m.sync()
        t1 = time.perf_counter()
        kms, launches = m.timing()
        d = m.diagnostics()
        if download:
            m.download_fluid()
    if d["flags"]:
        raise RuntimeError(f"device flags {d['flags']} after the timed steps")
    wall = t1 - t0
    kavg = kms / max(launches, 1) / 1e3
    achieved = BYTES[form] * n / kavg / 1e9
    return {
        "n": n,
        "value": n * steps / wall,
        "ms_per_step": wall / steps * 1e3,
        "kernel_ms": kavg * 1e3,
        "achieved_gbs": achieved,
        "roofline_frac": achieved / PEAK_GBS,
        "kT": d["twice_kinetic"] / (3 * n),
        "momentum": [float(x) for x in d["momentum"]],
        "tile_variant": variant,
    }


def run_swimmer(mcm, steps, warmup, form="fp64", L=32, radius=5.0, n_cilia=16, rodlength=10):
    """A swimmer configuration: periodic L^3 box at 10 solvent particles per
    cell, one Clami-type swimmer (643-bead shell of the given radius,
    n_cilia cilia of 3 x rodlength beads at Fibonacci sites, random beat
    phases) at the centre, h = 0.1 with 50 MD sub-steps.  Times whole steps
    (collision + stream + MD) and the MD half alone.  Defaults: the small
    FP64 configuration (1,123 beads)."""
    from mcm_amd import swimmer as SW

    h, nsub = 0.1, 50
    n = 10 * L ** 3
    c = L / 2.0
    sw = SW.clami(levels=3, radius=radius, center=(c, c, c), n_cilia=n_cilia, rodlength=rodlength, seed=1)
    prec = mcm.FP32BULK if form == "fp32bulk" else mcm.FP64
    with mcm.MPCD((L, L, L), precision=prec, seed=1) as m:
        m.create_fluid(n)
        m.md_attach(sw, globalt=0.0, ststep=h / nsub)
        for _ in range(warmup):
            m.step(h)
        m.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            m.step(h)
        m.sync()
        t1 = time.perf_counter()
        for _ in range(steps):
            m.md_run(h)
        m.sync()
        t2 = time.perf_counter()
        d = m.diagnostics()
    if d["flags"]:
        raise RuntimeError(f"device flags {d['flags']} in the swimmer run")
    return {
        "workload": f"periodic {L}^3, {n} solvent particles, swimmer of {len(sw['pos3'])} beads "
                    f"(643-bead shell of radius {radius:g}, {n_cilia} cilia x {3 * rodlength} beads), h={h}, "
                    f"{nsub} MD sub-steps, {form} form",
        "ms_per_step": (t1 - t0) / steps * 1e3,
        "md_ms_per_step": (t2 - t1) / steps * 1e3,
        "particle_steps_per_s": n * steps / (t1 - t0),
    }


def run_dense_cilia(mcm, steps, warmup, form="fp64"):
    """The dense-cilia configuration: periodic 128^3 cells at 10 particles
    per cell (20,971,520 particles), MPC-AT+a collision, a 643-bead shell of
    radius 20 with 1000 cilia of 3 x 20 beads (60,643 beads), h = 0.1 with
    100 MD sub-steps, in the FP64 or the FP32-bulk form.  Times whole steps
    and the MD half alone; reports kT."""
    from mcm_amd import swimmer as SW

    L, h, nsub = 128, 0.1, 100
    n = 10 * L ** 3
    sw = SW.clami(levels=3, radius=20.0, center=(L / 2.0,) * 3, n_cilia=1000, rodlength=20, seed=4)
    prec = mcm.FP32BULK if form == "fp32bulk" else mcm.FP64
    with mcm.MPCD((L, L, L), precision=prec, seed=4, collision=mcm.AT_A) as m:
        m.create_fluid(n)
        m.md_attach(sw, globalt=0.0, ststep=h / nsub)
        d0 = m.diagnostics()
        for _ in range(warmup):
            m.step(h)
        m.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            m.step(h)
        m.sync()
        t1 = time.perf_counter()
        for _ in range(steps):
            m.md_run(h)
        m.sync()
        t2 = time.perf_counter()
        d = m.diagnostics()
    if d["flags"]:
        raise RuntimeError(f"device flags {d['flags']} in the dense-cilia run")
    return {
        "workload": f"periodic {L}^3, {n} solvent particles, MPC-AT+a, swimmer of {len(sw['pos3'])} beads "
                    f"(643-bead shell of radius 20, 1000 cilia x 60 beads), h={h}, {nsub} MD sub-steps, {form} form",
        "ms_per_step": (t1 - t0) / steps * 1e3,
        "md_ms_per_step": (t2 - t1) / steps * 1e3,
        "particle_steps_per_s": n * steps / (t1 - t0),
        "kT": d["twice_kinetic"] / (3.0 * n),
        "deferred_tiles_last_step": d["deferred_tiles"],
    }


# the paper's swimmer as create.py writes its configuration file (100^3
# cells, radius 8, rings at -45/0/45 degrees with 7, 10, 7 cilia of 26
# beads, tstep 0.05, start delays in create.py's metachronal wave), as
# generated by tests/golden/make_volvox_cfg.py
PAPER_CFG = os.path.join(ROOT, "tests", "golden", "volvox_sim.cfg")


def run_driver(steps):
    """The paper's swimmer as its generator writes the configuration file
    (create.py: 100^3 cells, radius 8, rings at -45/0/45 degrees with 7, 10,
    7 cilia of 26 beads, tstep 0.05, start delays in a metachronal wave)
    run by the configuration driver as `cilia sim.cfg` runs it: steps,
    50 MD sub-steps each, and the .xyz/.san outputs every 50 steps.  Times
    the run loop's blocks (steps, then outputs) for `steps` steps in each
    form; the end-of-run checkpoint is not timed."""
    import tempfile

    from mcm_amd import FP32BULK, FP64
    from mcm_amd import sim as S

    with open(PAPER_CFG) as f:
        text = f.read()
    out = {}
    for form, prec in (("fp64", FP64), ("fp32bulk", FP32BULK)):
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "sim.cfg")
            with open(path, "w") as f:
                f.write(text)
            with S.Simulation(path) as sim:
                info = sim.info()
                sim.start(precision=prec, seed=5, measureinterval=50, restart=False, quiet=True)
                sim.next_tstep(50)  # warm-up block
                sim.writedata()
                t0 = time.perf_counter()
                for _ in range(steps // 50):
                    sim.next_tstep(50)
                    sim.writedata()
                t1 = time.perf_counter()
                lines = sum(1 for _ in open(os.path.join(d, "volvox.san")))
        n = 10 * 100 ** 3
        out[form] = {"ms_per_step": (t1 - t0) / steps * 1e3, "particle_steps_per_s": n * steps / (t1 - t0),
                     "san_lines": lines}
    out["workload"] = (f"the paper's swimmer from its configuration file through the driver: 100^3 cells, "
                       f"{10 * 100 ** 3} solvent particles, {info['nbeads']} beads ({info['ncilia']} cilia of 26), "
                       f"tstep 0.05, 50 MD sub-steps, outputs every 50 steps, {steps} steps")
    return out



def run_ensemble(mcm, steps, warmup, nbox=64, seed=7, form="fp32bulk", modes=("graph", "launches", "md_block"),
                 variant=-1):
    """The batched ensemble: nbox independent swimmers, each in a periodic
    cubic box of edge uniform in [32, 64] cells at 10 particles per cell,
    with a 643-bead shell of radius uniform in [5, 10], cilia count uniform
    in 8..128 of 3 x rodlength beads (rodlength uniform in 10..20) and its
    own seed; h = 0.1 with 50 MD sub-steps; the step kernel's tile variant
    chosen by the autotuner.  Times the ensemble step (one
    launch per phase for all boxes, replayed from a captured HIP graph), the
    same launches without the graph, the graph with the MD half run by a
    workgroup per swimmer, and sequential per-box mpcd_step calls (each box
    on its own stream), as aggregate particle-steps/s."""
    import numpy as np

    from mcm_amd import swimmer as SW

    rng = np.random.default_rng(seed)
    h, nsub = 0.1, 50
    prec = mcm.FP32BULK if form == "fp32bulk" else mcm.FP64
    # the tile variant every box shares, timed by the library's autotuner on
    # a box of the middle edge (or the one given)
    tune_ms = None
    if variant < 0:
        variant, tune_ms = mcm.autotune((48, 48, 48), 10 * 48 ** 3, steps=4, precision=prec, seed=seed)
    boxes, n_tot, nb_tot, edges = [], 0, 0, []
    try:
        for i in range(nbox):
            L = int(rng.integers(32, 65))
            ncil, rod, R = int(rng.integers(8, 129)), int(rng.integers(10, 21)), float(rng.uniform(5.0, 10.0))
            m = mcm.MPCD((L, L, L), precision=prec, seed=1000 + i, tile_variant=variant)
            boxes.append(m)
            m.create_fluid(10 * L ** 3)
            sw = SW.clami(levels=3, radius=R, center=(L / 2.0,) * 3, n_cilia=ncil, rodlength=rod, seed=1000 + i)
            m.md_attach(sw, globalt=0.0, ststep=h / nsub)
            n_tot += 10 * L ** 3
            nb_tot += len(sw["pos3"])
            edges.append(L)

        def seq(k):
            for _ in range(k):
                for m in boxes:
                    m.step(h)
            for m in boxes:
                m.sync()

        t_seq = None
        if len(modes) > 1:
            seq(warmup)
            k_seq = max(2, steps // 4)
            t0 = time.perf_counter()
            seq(k_seq)
            t_seq = (time.perf_counter() - t0) / k_seq
        out = {}
        mode_flags = {"graph": (True, True), "launches": (False, True), "md_block": (True, False)}
        with mcm.Ensemble(boxes) as e:
            # the MD half's launch shape, by the ensemble's autotuner on its
            # own first steps (which are steps of the run)
            tuned = e.autotune(h, steps=2)
            for name in modes:
                graph, wide = mode_flags[name]
                e.set_graph(graph)
                e.set_md_wide(wide and tuned["md_wide"])
                e.set_md_block(tuned["md_block"])
                e.step(h, warmup)
                e.sync()
                t0 = time.perf_counter()
                e.step(h, steps)
                e.sync()
                out[name] = (time.perf_counter() - t0) / steps
        flags = [m.diagnostics()["flags"] for m in boxes]
    finally:
        for m in boxes:
            m.close()
    if any(flags):
        raise RuntimeError(f"device flags {flags} in the ensemble run")
    res = {
        "workload": f"{nbox} swimmers, cubic boxes of edge {min(edges)}..{max(edges)} cells at 10 particles/cell "
                    f"({n_tot} particles, {nb_tot} beads in all), h={h}, {nsub} MD sub-steps, {form} form",
        "particles": n_tot,
        "beads": nb_tot,
        "ms_per_step": out["graph"] * 1e3,
        "particle_steps_per_s": n_tot / out["graph"],
        "tile_variant": {"chosen": variant, "shape": list(mcm.tile_variants(prec)[variant]),
                         "autotuner_ms_on_48cube": tune_ms},
        "md_launch_shape": tuned,
    }
    if "launches" in out:
        res["launches_ms_per_step"] = out["launches"] * 1e3
    if "md_block" in out:
        res["md_workgroup_per_swimmer_ms_per_step"] = out["md_block"] * 1e3
    if t_seq is not None:
        res["sequential_ms_per_step"] = t_seq * 1e3
        res["sequential_particle_steps_per_s"] = n_tot / t_seq
        res["speedup_over_sequential"] = t_seq / out["graph"]
    return res


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_threads():
    """The host's share of cores for this process: OMP_NUM_THREADS when set
    (the GPU box sets it to its per-GPU share), else every core."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return os.cpu_count() or 1


def cpu_baseline(configs=("C1", "C0", "C2"), repeats=3, mem_cap_gb=64.0):
    """The reference's threaded CPU step (per-thread per-cell pointer
    buckets, static OpenMP schedules, one MT19937 per thread: the oracle's
    OmpSim, built with the reference's -O3 -ffast-math -fopenmp) on the
    host's cores, timed around the step loop only (state creation and the
    bucket set-up excluded), median of `repeats` runs per configuration:
    C1 (256^3, 167,772,160 particles, 3 steps per run), C0 (32^3, 1000
    steps per run), C2 (100^3, 10M particles, 5 steps per run), SRD with the
    thermostat, h = 0.1, fluid only.  The buckets take cells x threads x
    152 B; the thread count is capped so they stay under mem_cap_gb.  The
    first configuration is the headline sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O

    sizes = {"C0": (32, 1000), "C1": (256, 3), "C2": (100, 5)}
    want = _cpu_threads()
    out = {}
    for name in configs:
        L, k = sizes[name]
        cells = L ** 3
        cap = max(1, int(mem_cap_gb * 1e9 / (cells * 152.0)))
        threads = min(want, cap)
        n = 10 * cells
        rng = np.random.default_rng(1)
        pos = rng.random((n, 3), dtype=np.float32).astype(np.float64) * L
        vel = np.ones((n, 4))
        vel[:, :3] = rng.standard_normal((n, 3), dtype=np.float32)
        sim = O.OmpSim((L, L, L), pos, vel, threads=threads, seed=5, baseline=True)
        sim.steps(1, 0.1)  # first touch of the buckets' growth
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            sim.steps(k, 0.1)
            times.append(time.perf_counter() - t0)
        th = sim.threads
        sim.close()
        del pos, vel
        dt = sorted(times)[len(times) // 2]
        out[name] = {"value": n * k / dt, "particles": n, "steps": k, "threads": th,
                     "thread_cap": cap if cap < want else None, "runs_s": times}
    head = out[configs[0]]
    return {
        "value": head["value"],
        "unit": "particle-steps/s",
        "cores": head["threads"],
        "kind": "port",
        "sample": f"the reference's OpenMP bucket step restated (oracle OmpSim, -O3 -ffast-math -fopenmp, "
                  f"FP64), {configs[0]}: {head['particles']} particles, {head['steps']} steps per run, median of "
                  f"{repeats} runs, {head['threads']} threads of {os.cpu_count()} cores ({_cpu_model()})",
        "configs": out,
        "cpu_model": _cpu_model(),
        "host_cores": os.cpu_count(),
    }


def copy_calibration(gib):
    """A plain copy of known size, so that rocprofv3's FETCH_SIZE/WRITE_SIZE
    can be scaled to bytes: reads and writes gib GiB once per launch."""
    import torch

    n = int(gib * 2 ** 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}), flush=True)
    return n * 4


COPY_KERNEL = "__amd_rocclr_copyBuffer"


def device_copy_bandwidth(gib=4.0, launches=10):
    """The achievable HBM rate of a plain device copy (torch's copy kernel,
    gib GiB read and written per launch), timed with events on torch's
    stream, which the copy runs on: 2 x bytes per launch over its time."""
    import torch

    n = int(gib * 2 ** 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    for _ in range(launches):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / launches
    del src, dst
    torch.cuda.empty_cache()
    gbs = 2.0 * n * 4 / (ms * 1e-3) / 1e9
    return {"gbs": gbs, "bytes_each_way": n * 4, "ms_per_copy": ms, "frac_of_peak": gbs / PEAK_GBS}


def _pmc_values(csv_path, name_part):
    import csv

    vals = []
    with open(csv_path) as f:
        for row in csv.DictReader(f):
            if name_part in row["Kernel_Name"]:
                vals.append(float(row["Counter_Value"]))
    return vals


def measure_traffic(form, size, variant, gib=1.0):
    """HBM bytes per launch of the step kernel from two rocprofv3 --pmc passes
    (FETCH_SIZE, then WRITE_SIZE) over a child run of this script, each scaled
    by what the same counter reads for a device copy of gib GiB in the same
    child.  Returns None when rocprofv3 is unavailable or a pass fails."""
    import glob
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3") else None)
    if exe is None:
        return None
    # the main pass only (not the bead or large pass, k_step_*_later)
    kname = "k_step_f32<" if form == "fp32bulk" else "k_step_f64<"
    n = 10 * size ** 3
    # the fluid's download after the steps reads each record once with the
    # step kernel's per-lane record loads (W words from the tile segments)
    # and writes 7 doubles per particle: a cross-check of the copy's factors
    rec_bytes = n * (28 if form == "fp32bulk" else 48)
    dl_write_bytes = n * 56
    env = dict(os.environ, TMPDIR="/tmp")
    res = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="mcm_pmc_", dir="/tmp")
        cmd = [exe, "--pmc", ctr, "--output-format", "csv", "-d", d, "-o", "pmc", "--",
               sys.executable, os.path.abspath(__file__), "--pmc-child", str(gib), "--forms", form,
               "--size", str(size), "--variant", str(variant)]
        try:
            subprocess.run(cmd, cwd="/tmp", env=env, timeout=150, check=True,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            paths = glob.glob(os.path.join(d, "**", "*counter_collection*.csv"), recursive=True)
            if not paths:
                return None
            step = _pmc_values(paths[0], kname)
            copy_bytes = int(gib * 2 ** 30) // 4 * 4
            copy = [v for v in _pmc_values(paths[0], COPY_KERNEL) if v * 1024 > 0.1 * copy_bytes]
            dl = _pmc_values(paths[0], "k_download<")
            if not step or not copy:
                return None
            copy_kib = sorted(copy)[len(copy) // 2]
            factor = copy_bytes / (copy_kib * 1024.0)
            res[ctr] = {"kib_per_launch": sum(step) / len(step), "factor": factor}
            if dl:
                known = rec_bytes if ctr == "FETCH_SIZE" else dl_write_bytes
                res[ctr]["download_factor"] = known / (dl[-1] * 1024.0)
        except (subprocess.SubprocessError, OSError, KeyError, ValueError):
            return None
        finally:
            shutil.rmtree(d, ignore_errors=True)
    f, w = res["FETCH_SIZE"], res["WRITE_SIZE"]
    return {
        "bytes": f["kib_per_launch"] * 1024 * f["factor"] + w["kib_per_launch"] * 1024 * w["factor"],
        "fetch_kib": f["kib_per_launch"], "fetch_factor": f["factor"],
        "write_kib": w["kib_per_launch"], "write_factor": w["factor"],
        # the same counters' factors on the download kernel (known bytes,
        # record-granular reads from the tile segments)
        "fetch_factor_record_reads": f.get("download_factor"),
        "write_factor_download": w.get("download_factor"),
    }


# the workloads beside the headline box, each alone (--workload)
WORKLOADS = {
    "swimmer": lambda mcm, k, w: run_swimmer(mcm, k, w),
    "swimmer_large": lambda mcm, k, w: run_swimmer(mcm, k, w, form="fp32bulk", L=100, radius=16.0, n_cilia=400,
                                                   rodlength=20),
    "dense_cilia_fp64": lambda mcm, k, w: run_dense_cilia(mcm, k, w, "fp64"),
    "dense_cilia_fp32bulk": lambda mcm, k, w: run_dense_cilia(mcm, k, w, "fp32bulk"),
    "ensemble": lambda mcm, k, w, v=-1: run_ensemble(mcm, k, w, modes=("graph",), variant=v),
    "driver": lambda mcm, k, w: run_driver(max(50, 50 * (k // 50))),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--forms", default="fp32bulk,fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep", action="store_true",
                    help="time every tile variant of each form (printed to stderr) before the measured run")
    ap.add_argument("--variant", type=int, default=-1, help="tile variant (-1: the library default)")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="only run a plain device-to-device copy of this many GiB (HBM counter calibration)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 --pmc passes")
    ap.add_argument("--no-swimmer", action="store_true", help="skip the swimmer configurations")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched ensemble of 64 swimmers")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the batched ensemble (graph replays, --steps steps), one JSON line")
    ap.add_argument("--pmc-child", type=float, default=0.0,
                    help="(used under rocprofv3) copy this many GiB, then 1 + 2 steps of the first form "
                         "and one download of its fluid")
    ap.add_argument("--workload", choices=sorted(WORKLOADS),
                    help="only this workload (--steps timed steps after --warmup), one JSON line; "
                         "for per-kernel traces and counters")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.pmc_child > 0:
        copy_calibration(args.pmc_child)
        import mcm_amd as mcm

        run_form(mcm, args.forms.split(",")[0], args.size, 2, 1, variant=args.variant, download=True)
        return
    if args.workload:
        import mcm_amd as mcm

        run = WORKLOADS[args.workload]
        res = run(mcm, args.steps, args.warmup, args.variant) if args.workload == "ensemble" else \
            run(mcm, args.steps, args.warmup)
        print(json.dumps({args.workload: res}), flush=True)
        return
    if args.ensemble_only:
        import mcm_amd as mcm

        print(json.dumps(run_ensemble(mcm, args.steps, args.warmup, modes=("graph",))), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU box has one GPU).", file=sys.stderr)
        sys.exit(2)
    forms = args.forms.split(",")
    head = "fp32bulk" if "fp32bulk" in forms else forms[0]
    # counter passes first, in child processes, before this process uses the GPU
    traffic = None if args.no_traffic else measure_traffic(head, args.size, args.variant)
    # the achievable peak, a plain device copy by torch, before this process
    # loads the library (torch's HIP runtime starts first, as in the counter
    # children)
    copy = device_copy_bandwidth()
    import mcm_amd as mcm

    if args.sweep:
        for f in forms:
            prec = mcm.FP32BULK if f == "fp32bulk" else mcm.FP64
            for v, shape in enumerate(mcm.tile_variants(prec)):
                r = run_form(mcm, f, args.size, max(3, args.steps // 2), 1, variant=v)
                print(f"[sweep] {f} variant {v} tile {shape}: kernel {r['kernel_ms']:.3f} ms "
                      f"({r['achieved_gbs']:.0f} GB/s), step {r['ms_per_step']:.3f} ms", file=sys.stderr, flush=True)
    res = {f: run_form(mcm, f, args.size, args.steps, args.warmup, variant=args.variant) for f in forms}
    r = res[head]
    out = {
        "metric": "MPCD particle-steps/s at 1 GPU + % HBM roofline; FP32 vs FP64 at fixed error",
        "value": r["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "fp32bulk" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"pure MPCD fluid, periodic {args.size}^3 cells, 10 particles/cell ({r['n']} particles), "
                        f"SRD 130 deg + MBS thermostat, h=0.1, random shift, {head} form",
            "cells": args.size ** 3,
            "particles": r["n"],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": r["achieved_gbs"],
            "peak": PEAK_GBS,
            "unit": "GB/s",
            "frac": r["achieved_gbs"] / PEAK_GBS,
            "traffic": None if traffic is None else traffic["bytes"],
        },
        "forms": {f: {k: v for k, v in x.items() if k != "n"} for f, x in res.items()},
    }
    if traffic is not None:
        out["traffic_detail"] = dict(traffic, algorithmic_bytes=BYTES[head] * r["n"])
    # the step kernels' rates against the achievable peak
    copy["step_frac_of_copy"] = {f: x["achieved_gbs"] / copy["gbs"] for f, x in res.items()}
    out["device_copy"] = copy
    if not args.no_swimmer:
        out["swimmer"] = run_swimmer(mcm, max(5, args.steps // 2), 2)
        # the paper-like swimmer: 100^3 cells, 10M particles, 400 cilia of
        # 3 x 20 beads on a shell of radius 16 (24,643 beads), FP32-bulk
        out["swimmer_large"] = run_swimmer(mcm, 5, 1, form="fp32bulk", L=100, radius=16.0, n_cilia=400,
                                           rodlength=20)
    if not args.no_ensemble:
        out["ensemble"] = run_ensemble(mcm, max(8, args.steps // 2), 2)
    if not args.no_swimmer:
        out["dense_cilia"] = {f: run_dense_cilia(mcm, 3, 1, f) for f in ("fp64", "fp32bulk")}
        out["driver"] = run_driver(200)
    if "fp64" in res and "fp32bulk" in res:
        out["fp32_over_fp64"] = res["fp32bulk"]["value"] / res["fp64"]["value"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
