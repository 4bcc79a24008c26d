"""Benchmark of the Yee timestep on one MI355X.

Workload: 3D Yee + CPML, FP32, 256^3 nodes, CPML 12 cells on all faces,
TF/SF plane wave along +x (FP64 aux-1D line, Ez-polarised Gaussian-modulated
pulse), centred eps_r = 6 cube of side 64, 8 Ez probes recorded every step.
A step is one full timestep (H half step + E half step) of the whole grid.
--workload phc_slab measures the 1280x1280x640 photonic-crystal slab
waveguide (soft Ez dipole source) instead, main line only.

Prints ONE JSON line: Gcell-updates/s (whole job), the roofline of the
dominant kernel (algorithmic bytes per launch / average launch duration,
timed with HIP events on the simulation's stream), that kernel's HBM bytes
per launch ("traffic": FETCH_SIZE and WRITE_SIZE from two rocprofv3 --pmc
runs of a short child bench with the same tuned configuration, each scaled
by what the counter reads for a plain device copy of known size) and a CPU
baseline (the FP64 reference-structured oracle on a bounded sample of the
same workload).

More objects ride on the same line:
  "batched": 256 seeded 2D photonic-crystal simulations of sizes
     {256..2048}^2 (FP32, CPML 16, TF/SF plane wave), 3000 steps, advanced
     together by one batch (temporally tiled kernel over a flattened tile
     list, hipGraph-captured chunk pairs), against the same simulations run
     one at a time with the single-simulation kernels: solves/s of both, and
     the relative L2 between their final fields on a sample.
  "batched_3d": 32 seeded 3D scatterers of 64^3-128^3 advanced together by
     the batched fused step, against the same simulations one at a time.
  "snapshots": the main line's simulation with Ez HDF5 snapshots every 50
     steps and with the reference's |E|^2-every-3-steps output, against the
     same steps without output (the PCIe- and disk-inclusive cost).
  "serving": a seeded Poisson request stream (80% 2D 256^2-2048^2, 20% 3D
     64^3-320^3 dielectric scatterers, 1000-5000 steps, FP32/FP16 storage,
     200 GB admission cap) through the continuous-batching server and through
     the plain per-request server: solves/s, p50/p99 latency, Gcell-updates/s;
     and the same for a stream of 2D requests alone arriving faster than
     either server clears them ("stream_2d": the batched 2D path's
     throughput against one request at a time).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _usable_cores():
    """Host cores this process may run on: the affinity mask, capped by the
    cgroup's CPU quota (a container's CPU share), and the machine's count."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, -(-int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return n, os.cpu_count() or n


def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_run(w, steps, threads, monitor=False, snap_every=0, snap_comp=None, e2=False, tmpdir=None):
    """Time `steps` oracle timesteps after one untimed step.  monitor: the
    reference's per-step max|E|^2 (serial, as blitz::max); snap_every: an
    HDF5 snapshot every that many steps in the reference's layout (|E|^2 of
    the 3D solver or the field of the 2D one), written like fileio."""
    import numpy as np
    from luxum_amd.params import LX_EX, LX_EY, LX_EZ
    from luxum_amd.sim import h5_write
    from oracle.lxoracle import OracleSim
    o = OracleSim(w.p, threads=threads)
    o.set_materials(w.ids, w.eps_r)
    o.set_probes(w.probes)
    o.step(1)
    path = os.path.join(tmpdir, "snap.h5") if snap_every else None
    nsnap = 0
    t0 = time.perf_counter()
    for t in range(steps):
        o.step(1)
        if monitor:
            o.max_e2()
        if snap_every and t % snap_every == 0:
            if e2:
                ex, ey, ez = o.field(LX_EX), o.field(LX_EY), o.field(LX_EZ)
                nx, ny, nz = w.p.dims
                out = np.zeros((nx, ny, nz - 1))
                c = (slice(0, nx - 1), slice(0, ny - 1), slice(0, nz - 1))
                out[c] = ez[c] ** 2 + ey[c] ** 2 + ex[c] ** 2
            else:
                out = o.field(snap_comp)
            h5_write(path, t, out, append=nsnap > 0)
            nsnap += 1
    dt = time.perf_counter() - t0
    o.close()
    if path and os.path.exists(path):
        os.remove(path)
    return {"steps": steps, "seconds": round(dt, 3), "gcell_per_s": round(w.cells * steps / dt / 1e9, 5),
            "snapshots": nsnap}


def cpu_child(args):
    """Runs in a child process whose OpenMP settings the parent chose: the
    FP64 reference-structured oracle on configuration 1 (the main line's
    workload) and configuration 0 (512^2 2D CPML, the designated CPU case),
    compute-only and as the reference times its loop (per-step monitor and
    snapshots at the reference's cadences).  Prints one JSON object."""
    import tempfile
    from luxum_amd import configs
    from luxum_amd.params import LX_EY
    th = args.cpu_child
    one = th == 1
    w1 = configs.config1(n=args.n, steps=4000)
    w0 = configs.config0(n=512, steps=4000)
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        res["config1"] = _oracle_run(w1, args.cpu_steps_1t if one else args.cpu_steps, th)
        res["config0"] = _oracle_run(w0, 1000, th)
        if not one:
            # maxwell_three_dim_pml.cc:266-281: |E|^2 + max every step, |E|^2 snapshot every 3 steps
            res["config1_reference_timed"] = _oracle_run(w1, args.cpu_ref_steps, th, monitor=True, snap_every=3,
                                                         e2=True, tmpdir=tmp)
            # maxwell_two_dim_pml.cc:179-185: max Ey^2 every step, Ey snapshot every 10 steps
            res["config0_reference_timed"] = _oracle_run(w0, 1000, th, monitor=True, snap_every=10,
                                                         snap_comp=LX_EY, tmpdir=tmp)
    print(json.dumps(res), flush=True)


def cpu_baseline(args):
    """The FP64 reference-structured oracle on the host cores, in child
    processes with OMP_PROC_BIND=close: all usable cores, then one thread."""
    import subprocess
    cores, nproc = _usable_cores()
    out = {}
    for th in (cores, 1):
        env = dict(os.environ, OMP_NUM_THREADS=str(th), OMP_PROC_BIND="close", OMP_PLACES="cores")
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", str(th), "--n", str(args.n),
               "--cpu-steps", str(args.cpu_steps), "--cpu-steps-1t", str(args.cpu_steps_1t),
               "--cpu-ref-steps", str(args.cpu_ref_steps)]
        print(f"bench: CPU baseline, {th} thread(s)", file=sys.stderr, flush=True)
        r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError(f"CPU baseline child failed ({r.returncode}): {r.stderr[-500:]}")
        out[th] = json.loads(r.stdout.strip().splitlines()[-1])
    a, b = out[cores], out[1]
    c1 = a["config1"]
    return {"value": c1["gcell_per_s"], "unit": "Gcell-updates/s", "cores": cores, "kind": "port",
            "sample": f"FP64 reference-structured oracle (separate curl and update passes, full-domain psi, per-cell "
                      f"Cb, serial TF/SF, OpenMP schedule(dynamic)), {args.n}^3 grid of the main line's workload, "
                      f"{c1['steps']} timesteps ({c1['seconds']} s), compute only, {cores} threads, "
                      f"OMP_PROC_BIND=close",
            "cpu_model": _cpu_model(), "nproc": nproc, "threads": cores,
            "one_thread": {"config1": b["config1"], "config0": b["config0"]},
            "config0": {"workload": "2D TM 512^2, CPML 16, TF/SF aux-1D plane wave, eps_r 10.89 square, 4 probes, "
                                    "1000 steps", "compute_only": a["config0"],
                        "reference_timed": a["config0_reference_timed"]},
            "config1_reference_timed": a["config1_reference_timed"]}


def batch_section(args):
    """Batched 2D solves/s against one-at-a-time single-simulation solves/s."""
    import numpy as np
    from luxum_amd import configs
    from luxum_amd.sim import Batch, Context, Simulation
    ctx = Context(0)
    sizes = configs.batch_sizes(args.batch_sims)
    ws = [configs.phc2d(n, i, steps=args.batch_steps) for i, n in enumerate(sizes)]
    cells = sum(w.cells for w in ws)
    print(f"bench: batched {len(ws)} sims, {cells / 1e6:.1f} Mcells, {args.batch_steps} steps", file=sys.stderr,
          flush=True)

    def make(w):
        s = Simulation(w.p, ctx)
        s.set_materials(w.ids, w.eps_r)
        s.set_probes(w.probes)
        return s

    sample = sorted(range(len(ws)), key=lambda q: sizes[q])[:: max(1, len(ws) // 6)][:6]
    sims = [make(w) for w in ws]
    b = Batch(ctx)
    for s in sims:
        b.add(s)
    b.sync()
    t0 = time.perf_counter()
    b.step(args.batch_steps)
    b.sync()
    t_batch = time.perf_counter() - t0
    ref = {q: {c: sims[q].field(c) for c in ws[q].p.components()} for q in sample}
    b.close()
    for s in sims:
        s.close()
    del sims
    n = len(ws)
    out = {"workload": f"{n} seeded 2D PhC sims, sizes {{256..2048}}^2, FP32, CPML 16, TF/SF aux-1D plane wave, "
                       f"{args.batch_steps} steps",
           "sims": n, "cells": cells, "steps": args.batch_steps,
           "solves_per_s": round(n / t_batch, 3), "gcell_per_s": round(cells * args.batch_steps / t_batch / 1e9, 2)}
    if args.no_single:
        return out
    print(f"bench: batched {t_batch:.3f} s; one at a time ...", file=sys.stderr, flush=True)
    t_single = 0.0
    num = den = 0.0
    for q, w in enumerate(ws):
        s = make(w)
        s.sync()
        t0 = time.perf_counter()
        s.step(args.batch_steps)
        s.sync()
        t_single += time.perf_counter() - t0
        if q in ref:
            for c, a in ref[q].items():
                o = s.field(c)
                num += float(np.sum((a - o) ** 2))
                den += float(np.sum(o ** 2))
        s.close()
    print(f"bench: one at a time {t_single:.3f} s", file=sys.stderr, flush=True)
    out.update({
            "one_at_a_time_solves_per_s": round(n / t_single, 3),
            "one_at_a_time_gcell_per_s": round(cells * args.batch_steps / t_single / 1e9, 2),
            "speedup": round(t_single / t_batch, 2),
            "rel_l2_vs_one_at_a_time": (num / den) ** 0.5 if den > 0 else None,
            "l2_sample_sims": len(sample)})
    if not args.no_traffic:
        try:
            out["traffic"] = batch_traffic(args, cells, out["gcell_per_s"])
            print(f"bench: batched 2D HBM traffic {out['traffic']}", file=sys.stderr, flush=True)
        except Exception as exc:   # the traffic stays unmeasured; the reason goes to stderr
            print(f"bench: batched 2D traffic not measured: {exc}", file=sys.stderr, flush=True)
    return out


def batch3d_section(args):
    """Batched 3D solves/s (the batched fused step over a flattened block
    list) against the same simulations one at a time."""
    import numpy as np
    from luxum_amd import configs
    from luxum_amd.sim import Batch, Context, Simulation
    ctx = Context(0)
    rng = np.random.default_rng(21)
    sizes = [int(n) for n in rng.choice([64, 80, 96, 112, 128], size=args.batch3d_sims)]
    ws = [configs.scatterer3d(n, rng, steps=args.batch3d_steps) for n in sizes]
    cells = sum(w.cells for w in ws)
    print(f"bench: batched 3D {len(ws)} sims, {cells / 1e6:.1f} Mcells, {args.batch3d_steps} steps", file=sys.stderr,
          flush=True)

    def make(w):
        s = Simulation(w.p, ctx)
        s.set_materials(w.ids, w.eps_r)
        s.set_probes(w.probes)
        return s

    sims = [make(w) for w in ws]
    b = Batch(ctx)
    for s in sims:
        b.add(s)
    b.step(8)   # descriptors, graphs
    b.sync()
    t0 = time.perf_counter()
    b.step(args.batch3d_steps)
    b.sync()
    t_batch = time.perf_counter() - t0
    ref = {c: sims[0].field(c) for c in ws[0].p.components()}
    b.close()
    for s in sims:
        s.close()
    t_single = 0.0
    num = den = 0.0
    for q, w in enumerate(ws):
        s = make(w)
        s.step(8)
        s.sync()
        t0 = time.perf_counter()
        s.step(args.batch3d_steps)
        s.sync()
        t_single += time.perf_counter() - t0
        if q == 0:
            for c, a in ref.items():
                o = s.field(c)
                num += float(np.sum((a - o) ** 2))
                den += float(np.sum(o ** 2))
        s.close()
    n = len(ws)
    return {"workload": f"{n} seeded 3D scatterers of config 4 (CPML 12, aux TF/SF, 8 probes), sizes "
                        f"{sorted(set(sizes))}^3, FP32, {args.batch3d_steps} steps",
            "sims": n, "cells": cells, "steps": args.batch3d_steps,
            "solves_per_s": round(n / t_batch, 3), "gcell_per_s": round(cells * args.batch3d_steps / t_batch / 1e9, 2),
            "one_at_a_time_solves_per_s": round(n / t_single, 3),
            "one_at_a_time_gcell_per_s": round(cells * args.batch3d_steps / t_single / 1e9, 2),
            "speedup": round(t_single / t_batch, 2),
            "rel_l2_vs_one_at_a_time": (num / den) ** 0.5 if den > 0 else None}


def _serve_stream(ctx, stream):
    """Run one request stream through the continuous-batching server and the
    plain per-request server."""
    import numpy as np
    from luxum_amd.server import Server
    out = {}
    series = {}
    for mode in ("batched", "plain"):
        srv = Server(ctx, hbm_cap=200 * 10 ** 9, plain=(mode == "plain"))
        ids = [srv.submit(w.p, w.ids, w.eps_r, steps, arr, w.probes) for w, steps, arr in stream]
        srv.run()
        st = srv.stats()
        lat = np.array([srv.result(i).latency for i in ids])
        series[mode] = [srv.probes(i) for i in ids]
        res = {"solves_per_s": round(st.completed / st.makespan, 3),
               "p50_s": round(float(np.percentile(lat, 50)), 4), "p99_s": round(float(np.percentile(lat, 99)), 4),
               "makespan_s": round(st.makespan, 3), "gcell_per_s": round(st.cell_updates / st.makespan / 1e9, 2),
               "completed": st.completed, "failed": st.failed, "peak_active": st.peak_active,
               "peak_reserved_GB": round(st.peak_bytes / 1e9, 2), "ticks": st.ticks}
        # the half-step method's algorithmic bytes over the makespan: a work
        # rate, not a bandwidth (the tiled 2D kernel keeps fields in registers
        # for 8 steps and the fused 3D step reads each value once)
        res["alg_GBps_per_makespan"] = round(st.alg_bytes / st.makespan / 1e9, 1)
        out[mode] = res
        print(f"bench: serving {mode}: {res}", file=sys.stderr, flush=True)
        srv.close()
    out["speedup_solves_per_s"] = round(out["batched"]["solves_per_s"] / out["plain"]["solves_per_s"], 2)
    # each request's probe series served in the mix against the same request
    # run alone (the plain server runs every request by itself), by storage
    # precision: the mix runs 2D requests in the tiled kernel, whose rounding
    # differs from the single-simulation kernels'
    from luxum_amd.params import LX_FP16
    errs = {"fp32": [], "fp16": []}
    for (w, _, _), a, b in zip(stream, series["batched"], series["plain"]):
        nb = float(np.linalg.norm(b))
        e = float(np.linalg.norm(a - b)) / nb if nb > 0 else float(np.linalg.norm(a))
        errs["fp16" if w.p.precision == LX_FP16 else "fp32"].append(e)
    out["rel_l2_mix_vs_alone"] = {k: {"max": max(v), "median": float(np.median(v)), "requests": len(v)}
                                  for k, v in errs.items() if v}
    return out


def serve_section(args):
    """Served request streams: continuous batching against the plain server,
    for the mixed 2D/3D stream and for its 2D requests alone."""
    from luxum_amd import configs
    from luxum_amd.sim import Context
    ctx = Context(0)
    stream = configs.serve_stream(args.serve_requests, args.serve_rate, seed=11)
    cells_steps = sum(w.cells * steps for w, steps, _ in stream)
    n3 = sum(1 for w, _, _ in stream if not w.p.is_planar)
    print(f"bench: serving {len(stream)} requests ({n3} 3D), {cells_steps / 1e9:.1f} Gcell-updates, "
          f"rate {args.serve_rate}/s", file=sys.stderr, flush=True)
    out = {"workload": f"{len(stream)} seeded requests, Poisson {args.serve_rate}/s, 80% 2D 256^2-2048^2 / "
                       f"20% 3D 64^3-320^3 dielectric scatterers, 1000-5000 steps, FP32/FP16 storage, "
                       f"200 GB admission cap",
           "requests": len(stream), "requests_3d": n3, "rate_per_s": args.serve_rate,
           "cell_updates": cells_steps}
    out.update(_serve_stream(ctx, stream))
    # the 2D requests alone (same generator, no 3D), arriving faster than
    # either server clears them: the throughput of the batched 2D path
    s2 = configs.serve_stream(args.serve2d_requests, args.serve2d_rate, seed=13, frac3d=0.0)
    print(f"bench: serving 2D stream, {len(s2)} requests, rate {args.serve2d_rate}/s", file=sys.stderr, flush=True)
    d2 = {"workload": f"{len(s2)} seeded 2D requests, Poisson {args.serve2d_rate}/s, 256^2-2048^2 dielectric "
                      f"scatterers, 1000-5000 steps, FP32/FP16 storage, 200 GB admission cap",
          "requests": len(s2), "rate_per_s": args.serve2d_rate,
          "cell_updates": sum(w.cells * steps for w, steps, _ in s2)}
    d2.update(_serve_stream(ctx, s2))
    out["stream_2d"] = d2
    return out


def snapshot_section(args, w, ctx, sim):
    """Snapshot output beside the time loop, on the main line's simulation:
    Ez HDF5 snapshots every --snap-every steps (gather kernel on the compute
    stream, D2H on the side stream into pinned slots, writer thread), and the
    reference's solve3d_pml cadence, |E|^2 every 3 steps.  Reported against the
    same steps without output; the files go to a temporary directory."""
    import shutil
    import tempfile

    import numpy as np
    from luxum_amd.params import LX_EZ
    from luxum_amd.sim import LX_OUT_E2, LX_OUT_FIELD, Writer
    n, every = args.snap_steps, args.snap_every
    d = tempfile.mkdtemp(prefix="lx_snap_")
    out = {"workload": f"main line, {n} steps, Ez snapshots every {every} steps (HDF5, FP64, the reference's "
                       f"layout); |E|^2 every 3 steps over 30 steps (maxwell_three_dim_pml.cc:279 cadence)"}
    try:
        sim.sync()
        t0 = time.perf_counter()
        sim.step(n)
        sim.sync()
        t_plain = time.perf_counter() - t0
        # the device-to-host part alone: gather + D2H into pinned host memory
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        m = int(np.prod(sim.p.field_shape(LX_EZ)))
        ptrs = [ctypes.c_void_p() for _ in range(2)]
        for q in ptrs:
            if hip.hipHostMalloc(ctypes.byref(q), ctypes.c_size_t(8 * m), 0) != 0:
                raise RuntimeError("hipHostMalloc failed")
        pinned = [np.ctypeslib.as_array(ctypes.cast(q, ctypes.POINTER(ctypes.c_double)), shape=(m,)) for q in ptrs]
        t0 = time.perf_counter()
        for q in range(n // every):
            sim.step(every)
            sim.snapshot_async(LX_EZ, pinned[q & 1])
        sim.sync()
        t_d2h = time.perf_counter() - t0
        del pinned
        for q in ptrs:
            hip.hipHostFree(q)
        wr = Writer(sim, os.path.join(d, "Ez_data.h5"), LX_OUT_FIELD, LX_EZ)
        t0 = time.perf_counter()
        for _ in range(n // every):
            sim.step(every)
            wr.write(sim.info().step)
        wr.close()
        sim.sync()
        t_snap = time.perf_counter() - t0
        nsnap = n // every
        per = os.path.getsize(os.path.join(d, "Ez_data.h5")) / max(nsnap, 1)
        out.update({"snapshots": nsnap, "bytes_per_snapshot": int(per),
                    "ms_per_step_plain": round(t_plain * 1e3 / n, 4),
                    "ms_per_step_with_d2h_only": round(t_d2h * 1e3 / (nsnap * every), 4),
                    "ms_per_step_with_snapshots": round(t_snap * 1e3 / (nsnap * every), 4),
                    "overhead_frac": round(t_snap / (nsnap * every) / (t_plain / n) - 1, 4)})
        print(f"bench: snapshots every {every}: {out['ms_per_step_with_snapshots']} ms/step against "
              f"{out['ms_per_step_plain']}", file=sys.stderr, flush=True)
        wr = Writer(sim, os.path.join(d, "Ez_data_e2.h5"), LX_OUT_E2, 0)
        t0 = time.perf_counter()
        for _ in range(10):
            sim.step(3)
            wr.write(sim.info().step)
        wr.close()
        sim.sync()
        t_e2 = time.perf_counter() - t0
        per2 = os.path.getsize(os.path.join(d, "Ez_data_e2.h5")) / 10
        out["e2_every_3"] = {"snapshots": 10, "bytes_per_snapshot": int(per2),
                             "ms_per_step": round(t_e2 * 1e3 / 30, 4),
                             "written_GBps": round(10 * per2 / t_e2 / 1e9, 3)}
        print(f"bench: |E|^2 every 3 steps: {out['e2_every_3']['ms_per_step']} ms/step "
              f"({out['e2_every_3']['written_GBps']} GB/s written)", file=sys.stderr, flush=True)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def _pmc_pass(counter, args, tune, outdir, child_args=None, fp16=None):
    """One `rocprofv3 --pmc <counter>` run of a short bench.py child (same
    workload, the tuned configuration pinned through LX_TUNE, plus a device
    copy of known size; or the child arguments given); returns
    {dispatch id: (kernel name, value)}."""
    import csv
    import glob
    import signal
    import subprocess
    prof = None
    for cand in (os.environ.get("ROCPROFV3"), "/opt/rocm/bin/rocprofv3"):
        if cand and os.path.exists(cand):
            prof = cand
            break
    if prof is None:
        raise RuntimeError("rocprofv3 not found")
    cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", outdir, "-o", counter.lower(), "--",
           sys.executable, os.path.join(ROOT, "bench.py")]
    if child_args is None:
        # the main line only, in the storage precision asked for (the FP16
        # line of the child would run under the same pin)
        if fp16 is None:
            fp16 = args.fp16
        cmd += ["--workload", args.workload, "--n", str(args.n), "--steps", "4", "--warmup", "1", "--kernel-steps",
                "2", "--copy-calib", str(args.traffic_copy_mib), "--no-batch", "--no-serve", "--no-cpu", "--no-snap",
                "--no-traffic", "--no-fp16"]
        if fp16:
            cmd.append("--fp16")
    else:
        cmd += child_args
    env = dict(os.environ, TMPDIR="/tmp")
    if tune:
        env["LX_TUNE"] = tune
    print(f"bench: traffic pass {counter} ({'LX_TUNE=' + tune if tune else ' '.join(child_args)})", file=sys.stderr,
          flush=True)
    p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                         start_new_session=True)
    try:
        _, err = p.communicate(timeout=args.traffic_timeout)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.communicate()
        raise RuntimeError(f"rocprofv3 --pmc {counter} timed out")
    rows = {}
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                d = r["Dispatch_Id"]
                name, v = rows.get(d, (r["Kernel_Name"], 0.0))
                rows[d] = (name, v + float(r["Counter_Value"]))
    if not rows:
        # the profiler can exit non-zero after complete results; only missing rows are a failure
        raise RuntimeError(f"rocprofv3 --pmc {counter}: no counter rows (exit {p.returncode}): "
                           f"{err.decode(errors='replace')[-400:]}")
    return rows


def fp16_section(args, ctx):
    """The main line's workload with E and H stored in FP16 (FP32
    arithmetic): the same timed run (autotune, warmup, --steps steps), the
    whole step's and the interior launch's fraction of the HBM roofline (the
    algorithmic bytes count each stored value at 2 B)."""
    from luxum_amd import configs
    from luxum_amd.params import LX_FP16
    from luxum_amd.sim import Simulation
    total = max(args.steps + args.warmup + args.kernel_steps, 2000)
    w = configs.config1(n=args.n, steps=total, precision=LX_FP16)
    sim = Simulation(w.p, ctx)
    sim.set_materials(w.ids, w.eps_r)
    sim.set_probes(w.probes)
    sim.autotune()
    sim.step(args.warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    elapsed = time.perf_counter() - t0
    ms_h, ms_e, _ = sim.time_kernels(args.kernel_steps)
    info = sim.info()
    sim.close()
    ms_step = elapsed * 1e3 / args.steps
    step_gbs = info.bytes_per_step / (ms_step * 1e-3) / 1e9
    core_gbs = info.core_bytes / (ms_h * 1e-3) / 1e9
    print(f"bench: FP16 storage {ms_step:.4f} ms/step", file=sys.stderr, flush=True)
    kern = (f"k_step3r interior box (fused H+E step, {info.interior_rows} rows per block sharing neighbour rows "
            f"through LDS), FP16 storage" if info.interior_rows else "k_step3 interior box (fused H+E step), FP16 storage")
    return info, {"workload": "the main line's workload, E and H stored in FP16 (FP32 arithmetic, automatic "
                        "power-of-two storage scale)",
            "value": round(w.cells * args.steps / elapsed / 1e9, 3), "unit": "Gcell-updates/s",
            "steps": args.steps, "ms_per_step": round(ms_step, 5), "bytes_per_step": info.bytes_per_step,
            "step_GBps": round(step_gbs, 1), "step_roofline_frac": round(step_gbs / HBM_PEAK_GBS, 4),
            "roofline": {"bound": "hbm", "kernel": kern,
                         "achieved": round(core_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(core_gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": info.core_bytes,
                         "ms_per_launch": round(ms_h, 5), "traffic": None},
            "ms_shell": round(ms_e, 5), "kz": info.kz, "cells_per_lane": info.cells_per_lane,
            "shell_concurrent": bool(info.shell_concurrent), "shell_pipelined": bool(info.shell_pipelined),
            "interior_rows": info.interior_rows}


def batch_child(args):
    """A short run of the batched 2D workload for the PMC passes of
    batch_traffic: the same 256 members, --batch-child steps, and a device
    copy of --traffic-copy-mib MiB (the counters' calibration)."""
    import torch
    from luxum_amd import configs
    from luxum_amd.sim import Batch, Context, Simulation
    a = torch.ones(args.traffic_copy_mib * (1 << 18), device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    ctx = Context(0)
    sizes = configs.batch_sizes(args.batch_sims)
    ws = [configs.phc2d(n, i, steps=args.batch_steps) for i, n in enumerate(sizes)]
    sims = []
    bt = Batch(ctx)
    for w in ws:
        s = Simulation(w.p, ctx)
        s.set_materials(w.ids, w.eps_r)
        s.set_probes(w.probes)
        sims.append(s)
        bt.add(s)
    bt.step(args.batch_child)
    bt.sync()
    bt.close()


def batch_traffic(args, cells, gcell_per_s):
    """HBM bytes per cell-update of the batched 2D step: FETCH_SIZE and
    WRITE_SIZE summed over every dispatch of the batch's kernels (the tiled
    kernel of each slab class and the aux-line kernel) in a short child run
    of the same members, each counter scaled by what it reads for a device
    copy of known size; and the HBM bandwidth those bytes take at the
    measured cell rate, as a fraction of the peak."""
    import re
    import shutil
    import tempfile
    steps = args.batch_traffic_steps
    child = ["--batch-child", str(steps), "--batch-sims", str(args.batch_sims), "--batch-steps", str(args.batch_steps),
             "--traffic-copy-mib", str(args.traffic_copy_mib)]
    pat = re.compile(r"k_tile2d|k_aux_chunk")
    copy_bytes = args.traffic_copy_mib * (1 << 20)
    tot = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        outdir = tempfile.mkdtemp(prefix="lx_pmc_", dir="/tmp")
        try:
            rows = _pmc_pass(counter, args, None, outdir, child_args=child)
        finally:
            shutil.rmtree(outdir, ignore_errors=True)
        ks = [v for name, v in rows.values() if pat.search(name)]
        cp = [v for name, v in rows.values() if "copy" in name.lower()]
        if not ks or not cp:
            raise RuntimeError(f"{counter}: no dispatch of the batch's kernels or of the copy kernel")
        factor = copy_bytes / (max(cp) * 1024.0)
        tot[counter] = (sum(ks) * 1024.0 * factor, factor, len(ks))
    cu = cells * steps
    rd, wr = tot["FETCH_SIZE"][0] / cu, tot["WRITE_SIZE"][0] / cu
    return {"bytes_per_cell_update": round(rd + wr, 3), "read_per_cell_update": round(rd, 3),
            "write_per_cell_update": round(wr, 3), "steps": steps, "dispatches": tot["FETCH_SIZE"][2],
            "fetch_factor": round(tot["FETCH_SIZE"][1], 3), "write_factor": round(tot["WRITE_SIZE"][1], 3),
            "hbm_GBps": round((rd + wr) * gcell_per_s, 1),
            "hbm_frac": round((rd + wr) * gcell_per_s / HBM_PEAK_GBS, 4)}


def traffic_section(args, info, dom_pattern, fp16=None):
    """HBM bytes per launch of the dominant kernel: FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch) from two rocprofv3 --pmc runs, each scaled by what the
    same counter reads for a plain device copy of known size in the same run."""
    import re
    import shutil
    import tempfile
    import statistics
    nw = info.waves if info.waves in (4, 8) else 4
    tune = (f"{int(info.fused_step)},{nw},{info.cells_per_lane},{info.kz},{int(info.shell_concurrent)},"
            f"{info.interior_rows}")
    copy_bytes = args.traffic_copy_mib * (1 << 20)
    pat = re.compile(dom_pattern)
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        outdir = tempfile.mkdtemp(prefix="lx_pmc_", dir="/tmp")
        try:
            rows = _pmc_pass(counter, args, tune, outdir, fp16=fp16)
        finally:
            shutil.rmtree(outdir, ignore_errors=True)
        dom = [v for name, v in rows.values() if pat.search(name)]
        cp = [v for name, v in rows.values() if "copy" in name.lower()]
        if not dom or not cp:
            raise RuntimeError(f"{counter}: no dispatch of the dominant kernel or of the copy kernel")
        factor = copy_bytes / (max(cp) * 1024.0)
        res[counter] = (statistics.median(dom) * 1024.0 * factor, factor, len(dom))
    rd, wr = res["FETCH_SIZE"], res["WRITE_SIZE"]
    return {"read": round(rd[0]), "write": round(wr[0]), "fetch_factor": round(rd[1], 3),
            "write_factor": round(wr[1], 3), "dispatches": rd[2], "copy_MiB": args.traffic_copy_mib}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--workload", choices=["cube", "phc_slab"], default="cube",
                    help="cube: 3D CPML dielectric cube with a TF/SF plane wave (--n^3); phc_slab: 1280x1280x640 "
                         "photonic-crystal slab waveguide driven by a soft Ez dipole (no batch, serving or CPU "
                         "sections: its FP64 CPU state would not fit in host memory)")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--fp16", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--cpu-steps-1t", type=int, default=5, help="config-1 steps of the one-thread CPU run")
    ap.add_argument("--cpu-ref-steps", type=int, default=30,
                    help="config-1 steps of the CPU run with the reference's monitor and |E|^2 output")
    ap.add_argument("--cpu-child", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--batch-child", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-steps", type=int, default=50)
    ap.add_argument("--batch-sims", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=3000)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch3d-sims", type=int, default=32)
    ap.add_argument("--batch3d-steps", type=int, default=500)
    ap.add_argument("--no-single", action="store_true", help="skip the one-at-a-time baseline of the batch section")
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve2d-requests", type=int, default=64)
    ap.add_argument("--serve2d-rate", type=float, default=200.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-snap", action="store_true", help="skip the snapshot-output section")
    ap.add_argument("--no-fp16", action="store_true", help="skip the FP16-storage run of the main line's workload")
    ap.add_argument("--snap-steps", type=int, default=400)
    ap.add_argument("--snap-every", type=int, default=50)
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (PMC traffic calibration)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc passes that measure the dominant kernel's HBM bytes")
    ap.add_argument("--traffic-copy-mib", type=int, default=512)
    ap.add_argument("--traffic-timeout", type=float, default=240.0)
    ap.add_argument("--batch-traffic-steps", type=int, default=48, dest="batch_traffic_steps",
                    help="steps of the batched 2D child run under rocprofv3 --pmc")
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args)
        return
    if args.batch_child:
        batch_child(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the whole domain is resident on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    from luxum_amd import configs
    from luxum_amd.params import LX_FP16, LX_FP32
    from luxum_amd.sim import Context, Simulation

    if args.copy_calib:
        import torch
        a = torch.ones(args.copy_calib * (1 << 18), device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(f"bench: copy calibration {args.copy_calib} MiB read + {args.copy_calib} MiB write",
              file=sys.stderr, flush=True)
        del a, b

    prec = LX_FP16 if args.fp16 else LX_FP32
    # every step this simulation takes (the timed run, the kernel timing, the
    # reference-shaped loop, the snapshot section's three runs and its |E|^2
    # run), so that its probe buffer never grows mid-run (a growth drops the
    # captured step graphs, and the next calls would time their capture)
    total = max(args.steps + args.warmup + args.kernel_steps + 80 + 3 * args.snap_steps + 30 + 64, 2000)
    tag = "FP16 storage" if args.fp16 else "FP32"
    if args.workload == "phc_slab":
        w = configs.config2(steps=total, precision=prec)
        args.no_batch = args.no_serve = args.no_cpu = True
        desc = (f"3D PhC slab waveguide {'x'.join(map(str, w.p.dims))}, rods eps_r 12 pitch 32 side 13 in a "
                f"64-cell slab, one row removed, CPML 16, soft Ez dipole (a/lambda 0.3), 4 probes, {tag}")
    else:
        w = configs.config1(n=args.n, steps=total, precision=prec)
        desc = f"3D Yee+CPML {args.n}^3, CPML 12, TF/SF aux-1D plane wave +x, eps_r=6 cube, 8 probes, {tag}"
    ctx = Context(0)
    sim = Simulation(w.p, ctx)
    sim.set_materials(w.ids, w.eps_r)
    sim.set_probes(w.probes)
    kz = sim.autotune()
    print(f"bench: autotuned z-block depth {kz}", file=sys.stderr, flush=True)
    print(f"bench: {w.name} {'fp16' if args.fp16 else 'fp32'} warmup {args.warmup}", file=sys.stderr, flush=True)
    sim.step(args.warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed * 1e3 / args.steps
    gcells = w.cells * args.steps / elapsed / 1e9
    print(f"bench: timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    # per-kernel durations with HIP events on the simulation's stream
    ms_h, ms_e, ms_ev_step = sim.time_kernels(args.kernel_steps)
    info = sim.info()
    s = 2 if args.fp16 else 4
    slab = sum(info.slab_cells)
    fused_frac = None
    if info.fused_step:
        # fused H+E step: the interior-box launch (ms_h) and, with CPML, the
        # shell launch around it (ms_e); the box launch is the dominant kernel
        dom, dbytes, dms = "k_step3 interior box (fused H+E step)", info.core_bytes, ms_h
        dom_pat = rf"k_step3r?<{'__half' if args.fp16 else 'float'}, [01], \d+, \d+>"
        if info.interior_rows:
            dom = (f"k_step3r interior box (fused H+E step, {info.interior_rows} rows per block sharing neighbour "
                   f"rows through LDS)")
        fused_frac = info.bytes_per_step / ((ms_h + ms_e) * 1e-3) / 1e9 / HBM_PEAK_GBS
    else:
        bytes_h = info.cells * 9 * s + slab * 16
        bytes_e = info.cells * (9 * s + 1) + slab * 16
        if ms_e >= ms_h:
            dom, dbytes, dms, dom_pat = "e_update_3d", bytes_e, ms_e, r"k_e3<"
        else:
            dom, dbytes, dms, dom_pat = "h_update_3d", bytes_h, ms_h, r"k_h3<"
    achieved = dbytes / (dms * 1e-3) / 1e9
    step_gbs = info.bytes_per_step / (ms_step * 1e-3) / 1e9
    mx, _ = sim.monitor()
    # the reference's time loop prints max|E|^2 every step
    # (maxwell_three_dim_pml.cc:167-173, 266-276): at the top of iteration t
    # the max that the end of iteration t-1 computed.  The C++ facade runs it
    # so by default: step t, then the max of step t-1 (its monitor, started
    # after step t-1, ran beside step t), then start the monitor of step t.
    # Timed beside it: one step, then the synchronous lx_sim_max_e2.
    loop_n = 40
    t1 = time.perf_counter()
    for _ in range(loop_n):
        sim.monitor()
    ms_mon = (time.perf_counter() - t1) * 1e3 / loop_n
    t1 = time.perf_counter()
    for _ in range(loop_n):
        sim.max_e2()
    ms_max = (time.perf_counter() - t1) * 1e3 / loop_n
    t1 = time.perf_counter()
    for _ in range(loop_n):
        sim.step(1)
        sim.max_e2()
    ms_loop_sync = (time.perf_counter() - t1) * 1e3 / loop_n
    sim.sync()
    t1 = time.perf_counter()
    for q in range(loop_n):
        sim.step(1)
        if q:
            sim.max_e2_finish()
        sim.max_e2_start()
    sim.max_e2_finish()
    ms_loop = (time.perf_counter() - t1) * 1e3 / loop_n

    out = {
        "metric": "Gcell-updates/s, %HBM roofline; batched solves/s, p50/p99 @ fixed L2 err",
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16-storage/f32" if args.fp16 else "f32",
        "data": "synthetic",
        "config": {"workload": desc,
                   "cells": info.cells, "bytes_per_step": info.bytes_per_step,
                   "step_GBps": round(step_gbs, 1), "step_roofline_frac": round(step_gbs / HBM_PEAK_GBS, 4),
                   "ms_h": round(ms_h, 5), "ms_e": round(ms_e, 5), "max_E2": mx, "kz": info.kz,
                   "waves_per_block": info.waves, "cells_per_lane": info.cells_per_lane, "shell_kz": info.shell_kz,
                   "shell_concurrent": bool(info.shell_concurrent), "interior_rows": info.interior_rows,
                   "fused_step": bool(info.fused_step),
                   "fused_step_roofline_frac": None if fused_frac is None else round(fused_frac, 4)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": dbytes, "ms_per_launch": round(dms, 5)},
    }
    out["config"]["shell_pipelined"] = bool(info.shell_pipelined)
    out["reference_loop"] = {"iterations": loop_n, "ms_per_step": round(ms_loop, 5),
                             "ms_per_step_synchronous": round(ms_loop_sync, 5), "ms_max_e2": round(ms_max, 5),
                             "ms_monitor": round(ms_mon, 5),
                             "note": "the facade's default loop: lx_sim_step(1), then lx_sim_max_e2_finish (FP64 "
                                     "max|E|^2 over the E components of the previous step, non-finite guard; its "
                                     "monitor ran beside this step), then lx_sim_max_e2_start, per iteration; "
                                     "synchronous: lx_sim_step(1) then lx_sim_max_e2; ms_monitor: lx_sim_monitor "
                                     "(max, energy over all six components); the main line above steps in graphs "
                                     "without a per-step monitor"}
    if not args.no_snap and args.workload == "cube":
        out["snapshots"] = snapshot_section(args, w, ctx, sim)
    sim.close()
    info16 = None
    if not args.no_fp16 and args.workload == "cube" and not args.fp16:
        info16, out["fp16"] = fp16_section(args, ctx)
    if not args.no_traffic:
        # HBM traffic of the dominant kernel (and of the FP16 line's interior
        # launch), each from children running only that line, pinned to the
        # configuration the timed run used
        lines = [(out["roofline"], info, dom_pat, args.fp16)]
        if info16 is not None and info16.fused_step:
            lines.append((out["fp16"]["roofline"], info16, r"k_step3r?<__half, [01], \d+, \d+>", True))
        for roof, inf, pat, f16 in lines:
            try:
                t = traffic_section(args, inf, pat, fp16=f16)
                roof["traffic"] = t["read"] + t["write"]
                roof["traffic_detail"] = t
                print(f"bench: {'FP16' if f16 else 'main line'} interior HBM traffic {t}", file=sys.stderr, flush=True)
            except Exception as exc:   # traffic stays null; the reason goes to stderr
                print(f"bench: traffic not measured: {exc}", file=sys.stderr, flush=True)
    if not args.no_batch:
        out["batched"] = batch_section(args)
        if args.batch3d_sims > 0:
            out["batched_3d"] = batch3d_section(args)
    if not args.no_serve:
        out["serving"] = serve_section(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
