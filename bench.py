#!/usr/bin/env python3
"""FP64 particle-steps/s of NairnMPM's explicit step on one MI355X.

Workload: the 3D elastic-plastic block -- 128^3 grid of 1 mm cells, 2x2x2 particles per cell in
the central 96^3 cells (7,077,888 particles), IsoPlasticity J2 with linear hardening (E=70 GPa,
nu=0.33, yield 300 MPa, Ep 1 GPa, rho 2.7 g/cm^3), uGIMP, USAVG+ with FLIP, the two halves
colliding at +-50 m/s. A step is one NairnMPM::MPMStep over all particles; inputs are resident in
HBM before the timed region. The line's "served" object is the metric's other half: a short
served 3D stream (8 J2 / neo-Hookean blocks of 50k-2M particles, Poisson arrivals) through the
solve server, with p50/p99 latency and aggregate particle-steps/s against the same solves run one
at a time. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nairn-mpm-fea_amd"))

import numpy as np  # noqa: E402

import nairnmpm as nm  # noqa: E402
from nairnmpm import configs  # noqa: E402

METRIC = "FP64 particle-steps/s (3D J2 GIMP) + served p50/p99 latency vs serial"
HBM_PEAK = 8.0e12

# Algorithmic HBM bytes of each kernel group for the 3D J2 path: (bytes per particle, bytes per
# active node). Particle terms count every particle word the kernel must read or write once; node
# terms count the grid values it must read or write once.
ALG_BYTES = {
    "k_p2g_mass": (56 + 4, 32),        # pos, vel, mass, elem -> m, p
    "k_node_mass": (0, 2 * 32 + 24),   # m, p (partials) -> m, p, v
    "k_usf": (24 + 4 + 4 + 2 * 184, 24),              # pos, elem, mat, state r/w; v in
    "k_p2g_force": (24 + 4 + 8 + 56, 24),             # pos, elem, mass, stress+P; f out
    "k_node_force": (0, 24 + 24 + 24 + 24),           # f in; p r/w; v out
    "k_g2p_p2g": (24 + 24 + 8 + 4 + 48, 56 + 24),     # pos, vel, mass, elem; vel, pos out; v,f,m in; p out
    "k_node_mom2": (0, 24 + 24),
    "k_usl_reset": (24 + 24 + 4 + 2 * 184 + 8, 24),   # pos, new pos, elem, state r/w, elem+key out; v in
    "binning": (16, 0),
    # multimaterial: two contact extrapolations (pos, orig, F, mass, elem in; volume, displacement,
    # gradient out per node) and three node-local contact passes
    "contact": (2 * (24 + 24 + 72 + 8 + 4), 2 * 56 + 3 * 160),
}
CPDI_EXTRA = 72      # lCPDI: semi-side vectors read by every tile kernel (written by the first)

# The same groups for 2D plane strain: positions, velocities and node vectors have two components,
# the constitutive state is ep(4) + wrot(1) + sp(4) + pressure + alt(4) + history = 15 words.
ALG_BYTES_2D = {
    "k_p2g_mass": (16 + 16 + 8 + 4, 24),
    "k_node_mass": (0, 24 + 24 + 16),
    "k_usf": (16 + 4 + 4 + 2 * 120, 16),
    "k_p2g_force": (16 + 4 + 8 + 40, 16),
    "k_node_force": (0, 16 + 32 + 16),
    "k_g2p_p2g": (16 + 16 + 8 + 4 + 32, 40 + 16),
    "k_node_mom2": (0, 16 + 16),
    "k_usl_reset": (16 + 16 + 4 + 2 * 120 + 8, 16),
    "binning": (16, 0),
}


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic_j2_block_128.md")


def measured_traffic(kernel, cells):
    """(HBM bytes per launch of `kernel` on this workload, build the figure was measured on) from the
    committed FETCH_SIZE/WRITE_SIZE counter summary (a 'Build: <commit>' line, then rows
    '| kernel | cells | fetch KiB | write KiB | traffic bytes |'); (None, None) when the kernel or size
    was not profiled. The figure is a table lookup, not a measurement of this run."""
    build = None
    try:
        with open(TRAFFIC_TABLE) as f:
            for line in f:
                if line.startswith("Build:"):
                    build = line.split(":", 1)[1].strip()
                cols = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cols) >= 5 and cols[0] == kernel and cols[1] == str(cells):
                    return float(cols[4]), build
    except OSError:
        pass
    return None, None


def copy_calibration(nbytes, launches=5):
    """A plain device copy of known size (torch.neg over float64, nbytes read + nbytes written per
    launch): the reference point that scales the FETCH_SIZE/WRITE_SIZE counters to bytes."""
    import torch
    n = nbytes // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(launches):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / launches
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8,
                      "launches": launches, "ms_per_launch": dt * 1e3,
                      "GBps": 2 * n * 8 / dt / 1e9}), flush=True)


def cpu_threads(asked=0):
    """Host threads of the CPU baseline: the asked number, else OMP_NUM_THREADS, else the host's
    cores, at most 16 (one GPU's share of the host)."""
    if asked > 0:
        return asked
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_sample(pairs, seconds, steps=3, threads=1):
    """Oracle particle-steps/s with `threads` host threads over [(problem, (d, i))] run in order
    until about `seconds` of wall time is spent stepping."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    done = ps = 0
    t = 0.0
    for prob, (d, i) in pairs:
        g, mats, run, n, _ = prob.describe()
        o = orc.Oracle(g, mats, run, d, i)
        o.set_threads(threads)
        tc = time.perf_counter()
        o.step(steps)
        t += time.perf_counter() - tc
        del o
        done += 1
        ps += n * steps
        if t >= seconds:
            break
    return ps / t, done, ps, t


def bench_batch(a):
    """The batched 2D sweep: a.instances independent plane-strain blocks (elastic or J2, 2k-200k
    particles log-uniform, their own materials and time steps) stepped together in one context."""
    t0 = time.time()
    pairs = []
    for k in range(a.instances):
        xml, fix, _ = configs.sweep_instance_2d(k)
        prob = nm.Problem.from_xml(xml)
        d, i = prob.particles()
        fix(d)
        pairs.append((prob, (d, i)))
    batch = nm.Batch([p for p, _ in pairs], [q for _, q in pairs])
    npart = int(sum(batch.counts))
    print(f"# setup {time.time() - t0:.1f} s: {a.instances} instances, {npart} particles "
          f"({min(batch.counts)}..{max(batch.counts)} per instance)", file=sys.stderr, flush=True)
    batch.step(a.warmup)
    batch.sync()
    gsteps, tune_ms = batch.autotune(trial_steps=32)
    tshape = batch.tile_shape
    print(f"# autotuner: tile shape {tshape}, graphs of {gsteps} steps ({tune_ms:.3f} ms/step in trials)",
          file=sys.stderr, flush=True)

    # timed region: K batch steps replayed from the tuned graph
    batch.sync()
    t1 = time.perf_counter()
    batch.step(a.steps)
    batch.sync()
    elapsed = time.perf_counter() - t1
    value = npart * a.steps / elapsed

    # kernel-group times from HIP events on the context's stream (eager launches)
    pst = min(a.steps, 20)
    batch.set_profiling(True)
    batch.step(pst)
    batch.sync()
    times = batch.kernel_times()
    batch.set_profiling(False)
    mass, _, _ = batch.grid_fields()
    active_nodes = int(np.count_nonzero(mass))
    for g, (ms, n) in times.items():
        print(f"#   {g:14s} {ms / max(n, 1):9.4f} ms/launch ({n} launches)", file=sys.stderr)
    dom = max(times, key=lambda g: times[g][0])
    dom_ms = times[dom][0] / max(times[dom][1], 1)
    bp, bn = ALG_BYTES_2D[dom]
    alg_bytes = bp * npart + bn * active_nodes
    achieved = alg_bytes / (dom_ms * 1e-3)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": None, "kernel": dom,
                "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": dom_ms}
    batch.close()

    # the same instances solved one at a time (own context, default two-step graphs)
    ser_ps = 0
    ser_t = 0.0
    for prob, (d, i) in pairs:
        g, mats, run, n, _ = prob.describe()
        sim = nm.Simulation(g, mats, run)
        sim.upload(d, i)
        sim.step(2)
        sim.sync()
        tc = time.perf_counter()
        sim.step(a.serial_steps)
        sim.sync()
        ser_t += time.perf_counter() - tc
        ser_ps += n * a.serial_steps
        sim.close()
    serial = ser_ps / ser_t
    print(f"# batch {value:.4e} particle-steps/s, one at a time {serial:.4e} ({value / serial:.2f}x)",
          file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        th = cpu_threads(a.cpu_threads)
        v, done, ps, t = cpu_sample(pairs, a.cpu_seconds, threads=th)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": th, "kind": "port",
               "sample": f"3 steps of each of the first {done} sweep instances ({ps} particle-steps), "
                         f"oracle/mpm_oracle.cpp in its threaded mode on {th} threads of {cpu_model()}, {t:.1f} s"}
    return {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 * elapsed / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched 2D sweep: {a.instances} plane-strain elastic/J2 blocks, 2k-200k particles "
                               "log-uniform, E 0.5-5 GPa, nu 0.25-0.4, yield 5-50 MPa, uGIMP, one step of "
                               "each instance per batch step",
                   "particles": npart, "instances": a.instances, "active_nodes": active_nodes,
                   "graph_steps": gsteps, "tile_shape": tshape, "one_at_a_time_value": serial,
                   "one_at_a_time_sample": f"{a.serial_steps} graph-replayed steps of every instance alone",
                   "speedup_vs_one_at_a_time": value / serial},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


def served_stream(requests, rate, graph_steps=8, max_resident=0):
    """The served 3D stream: `requests` J2 / neo-Hookean blocks (50k-2M particles log-uniform,
    200-2000 steps) arriving as a Poisson process of `rate` per second, served by the solve
    server's request loop; then the same solves one at a time as plain solves (create, upload,
    step, download), first come first served. Returns (summary dict, requests)."""
    t0 = time.time()
    reqs = []
    for k in range(requests):
        xml, steps, j2 = configs.serve_request_3d(k)
        prob = nm.Problem.from_xml(xml)
        g, mats, run, n, _ = prob.describe()
        d, i = prob.particles()
        reqs.append({"k": k, "prob": prob, "grid": g, "mats": mats, "run": run, "n": n, "d": d, "i": i,
                     "steps": steps, "j2": j2})
    rng = configs.SplitMix64(777)
    gaps = -np.log(1.0 - rng.uniform_array(requests)) / rate
    arrivals = np.cumsum(gaps) - gaps[0]               # first request at t = 0
    total_ps = sum(r["n"] * r["steps"] for r in reqs)
    print(f"# served stream setup {time.time() - t0:.1f} s: {requests} requests, {sum(r['n'] for r in reqs)} "
          f"particles, {total_ps:.3e} particle-steps, arrivals over {arrivals[-1]:.2f} s", file=sys.stderr, flush=True)

    # warm-up: one small solve through a server (module loads, first graph capture)
    srv = nm.Server(analysis=nm.THREED, graph_steps=graph_steps)
    w = reqs[min(range(len(reqs)), key=lambda j: reqs[j]["n"])]
    wid = srv.submit(w["grid"], w["mats"], w["run"], w["d"], w["i"], 4)
    while srv.info(wid)["state"] < nm.REQ_DONE:
        srv.advance(4)
    srv.close()

    # served: the request loop runs on the server's thread; this thread submits at arrival times
    srv = nm.Server(analysis=nm.THREED, graph_steps=graph_steps, max_resident=max_resident)
    srv.start()
    base = srv.clock()
    wall = time.perf_counter()
    ids = []
    for r, t_arr in zip(reqs, arrivals):
        delay = t_arr - (time.perf_counter() - wall)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(r["grid"], r["mats"], r["run"], r["d"], r["i"], r["steps"]))
    for rid in ids:
        srv.wait(rid)
    srv.stop()
    infos = [srv.info(rid) for rid in ids]
    lat = np.array([inf["finished"] - base - t for inf, t in zip(infos, arrivals)])
    span = max(inf["finished"] for inf in infos) - base
    served = total_ps / span
    cnt = srv.counts()
    layouts = cnt["layouts"]
    served_out = [srv.collect(rid) for rid in ids]
    srv.close()
    print(f"# served: {served:.4e} particle-steps/s over {span:.2f} s, {layouts} layouts "
          f"({cnt['layout_seconds']:.2f} s building them, {cnt['step_seconds']:.2f} s in batch steps), latency p50 "
          f"{np.percentile(lat, 50):.2f} s p99 {np.percentile(lat, 99):.2f} s", file=sys.stderr, flush=True)

    # one at a time: a plain solve per request (create, upload, step, download), first come first served
    service = []
    equal = 0
    for r, (d_srv, i_srv) in zip(reqs, served_out):
        tc = time.perf_counter()
        sim = nm.Simulation(r["grid"], r["mats"], r["run"])
        sim.upload(r["d"], r["i"])
        sim.step(r["steps"])
        d_one, i_one = sim.particles()
        sim.close()
        service.append(time.perf_counter() - tc)
        equal += int(np.array_equal(d_one, d_srv) and np.array_equal(i_one, i_srv))
    fin = 0.0
    slat = []
    for t_arr, sv in zip(arrivals, service):
        fin = max(fin, t_arr) + sv
        slat.append(fin - t_arr)
    slat = np.array(slat)
    serial_span = fin
    serial = total_ps / serial_span
    print(f"# one at a time: {serial:.4e} particle-steps/s over {serial_span:.2f} s (busy {sum(service):.2f} s), "
          f"latency p50 {np.percentile(slat, 50):.2f} s p99 {np.percentile(slat, 99):.2f} s; "
          f"{equal}/{len(reqs)} served results equal their solo runs", file=sys.stderr, flush=True)
    summary = {
        "workload": f"served 3D stream: {requests} J2 or neo-Hookean blocks, 50k-2M particles log-uniform, "
                    f"200-2000 steps, uGIMP, Poisson arrivals at {rate:g}/s",
        "requests": requests, "particle_steps": int(total_ps),
        "aggregate": served, "unit": "particle-steps/s", "span_s": span,
        "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
        "serial_aggregate": serial, "serial_span_s": serial_span, "serial_busy_s": float(sum(service)),
        "serial_p50_s": float(np.percentile(slat, 50)), "serial_p99_s": float(np.percentile(slat, 99)),
        "speedup": served / serial, "target_speedup": 5.0, "target_met": bool(served / serial >= 5.0),
        "results_equal_alone": f"{equal}/{len(reqs)}",
        "layouts": layouts, "layout_s": cnt["layout_seconds"], "batch_step_s": cnt["step_seconds"],
        "max_resident": max_resident,
    }
    return summary, reqs


def bench_serve(a):
    """The served 3D stream as its own bench line (value = served aggregate particle-steps/s)."""
    sv, reqs = served_stream(a.requests, a.rate, a.graph_steps, a.max_resident)
    cpu = None
    if not a.no_cpu:
        pairs = [(r["prob"], (r["d"], r["i"])) for r in sorted(reqs, key=lambda r: r["n"])]
        th = cpu_threads(a.cpu_threads)
        v, done, ps, t = cpu_sample(pairs, a.cpu_seconds, steps=2, threads=th)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": th, "kind": "port",
               "sample": f"2 steps of each of the {done} smallest requests ({ps} particle-steps), "
                         f"oracle/mpm_oracle.cpp in its threaded mode on {th} threads of {cpu_model()}, {t:.1f} s"}
    return {
        "metric": METRIC, "value": sv["aggregate"], "unit": "particle-steps/s", "n_gpus": 1,
        "steps": sv["particle_steps"], "warmup": 1, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": sv["workload"], "particle_steps": sv["particle_steps"]},
        "served": sv,
        "roofline": None,
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 500, the J2 block's run length; contact 200; batch 100)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=3, help="least steps of the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (default: OMP_NUM_THREADS, else the cores, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["j2", "clamped", "contact", "batch", "serve"], default="j2",
                    help="j2: the 7.08M-particle J2 GIMP block (default); clamped: the same block clamped on "
                         "its bottom face and loaded on its top face; contact: the lCPDI contact pair; "
                         "batch: the 256-instance 2D sweep; serve: the served 3D stream")
    ap.add_argument("--requests", type=int, default=16, help="serve: number of requests")
    ap.add_argument("--served-requests", type=int, default=8,
                    help="j2/clamped: requests of the short served stream reported in the line's 'served' "
                         "object (0: none)")
    ap.add_argument("--rate", type=float, default=4.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--graph-steps", type=int, default=8, help="serve: batch steps per graph / admission check")
    ap.add_argument("--max-resident", type=int, default=None,
                    help="served stream: most solves resident at once (0: no limit; default 3 for --workload serve, "
                         "where it keeps p99 below one at a time, and no limit for the 8-request stream of the j2 line)")
    ap.add_argument("--xpic-order", type=int, default=0,
                    help="j2/clamped: XPIC(k) in every step through a PeriodicXPIC task (0: FLIP)")
    ap.add_argument("--fmpm", action="store_true", help="with --xpic-order: FMPM(k) instead of XPIC(k)")
    ap.add_argument("--normals", type=int, default=2,
                    help="contact: MultiMaterialMode normals method (2 AVGG, 5 linear and 6 logistic regression)")
    ap.add_argument("--archive", action="store_true",
                    help="j2/clamped/contact: time writing one archive of the final state (to a temporary file)")
    ap.add_argument("--tile-shape", type=int, choices=[0, 1], default=0,
                    help="j2/clamped/contact: tile shape (0: 4x4x4 cells per tile, 1: 8x4x4); the batch tunes its own")
    ap.add_argument("--instances", type=int, default=256, help="batch: number of sweep instances")
    ap.add_argument("--serial-steps", type=int, default=20, help="batch: steps of each instance run alone")
    ap.add_argument("--cpu-seconds", type=float, default=10., help="CPU baseline sample length (wall seconds)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy used to scale the HBM byte counters")
    a = ap.parse_args()
    if a.max_resident is None:
        a.max_resident = 3 if a.workload == "serve" else 0
    if a.steps is None:
        a.steps = {"j2": 500, "clamped": 500, "contact": 200}.get(a.workload, 100)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the multi-GPU slab "
              "decomposition is a design, not a built path", file=sys.stderr)
        sys.exit(2)

    if a.workload == "batch":
        print(json.dumps(bench_batch(a)), flush=True)
        return
    if a.workload == "serve":
        print(json.dumps(bench_serve(a)), flush=True)
        return

    lo, hi = a.cells // 8, a.cells - a.cells // 8
    t0 = time.time()
    if a.workload == "contact":
        # two neo-Hookean cylinders (radius 20, length 60 cells) closing at 20 m/s, lCPDI,
        # multimaterial contact with Coulomb friction 0.3, AVGG normals
        cells = a.cells if a.cells != 128 else 160
        xml = configs.contact_cylinders_3d(cells=cells, radius=cells / 8., length=cells * 3. / 8., gap=4. * cells / 160.,
                                           normals=a.normals)
        nname = {0: "MAXG", 1: "MAXV", 2: "AVGG", 3: "OWNG", 5: "linear regression", 6: "logistic regression"}
        workload = (f"3D two-body contact: neo-Hookean cylinders r={cells / 8:g} L={cells * 3 / 8:g} cells on a "
                    f"{cells}^3 grid, lCPDI, Coulomb mu=0.3, {nname.get(a.normals, a.normals)} normals, closing at 20 m/s")
        a.cells = cells
    else:
        xml = configs.j2_block_3d(cells=a.cells, lo=lo, hi=hi)
        workload = (f"3D J2 GIMP block, {a.cells}^3 cells of 1 mm, 2x2x2 particles per cell "
                    f"in [{lo},{hi}]^3 mm, halves colliding at +-50 m/s")
        if a.workload == "clamped":
            # grid velocity conditions on the bottom face's nodes (fixed in x, y, z) and a cosine
            # x velocity on one side face; a net z load and a sine x load on the top face's particles
            xml = configs.grid_bcs(xml, [
                ("BCBox", {"xmin": lo, "xmax": hi, "ymin": lo, "ymax": hi, "zmin": lo - 0.1, "zmax": lo + 0.1},
                 [{"dir": 1, "vel": 0.}, {"dir": 2, "vel": 0.}, {"dir": 3, "vel": 0.}]),
                ("BCBox", {"xmin": lo - 0.1, "xmax": lo + 0.1, "ymin": lo, "ymax": hi, "zmin": lo + 1, "zmax": hi},
                 [{"dir": 1, "style": 4, "vel": 1000., "time": 20.}])])
            xml = configs.particle_bcs(xml, [
                ("BCBox", {"xmin": lo, "xmax": hi, "ymin": lo, "ymax": hi, "zmin": hi - 0.5, "zmax": hi},
                 ["net", {"dir": 3, "load": 2.0e5}, {"dir": 1, "style": 3, "load": 1.0e4, "ftime": 20.}])])
            workload += ("; clamped on the bottom face's nodes, a cosine x velocity on one side face, "
                         "a net z load and a sine x load on the top face's particles")
    if a.xpic_order > 0 and a.workload != "contact":
        xml = configs.periodic_xpic(xml, order=a.xpic_order, fmpm=a.fmpm, steps=1)
        workload += f"; {'FMPM' if a.fmpm else 'XPIC'}({a.xpic_order}) in every step (PeriodicXPIC)"
    prob = nm.Problem.from_xml(xml)
    grid, mats, run, npart, _ = prob.describe()
    dfields, ifields = prob.particles()
    setup_s = time.time() - t0
    print(f"# setup {setup_s:.1f} s: {npart} particles, grid {grid.nx}x{grid.ny}x{grid.nz} cells, "
          f"dt {run.timestep:.4e} s", file=sys.stderr, flush=True)

    sim = nm.Simulation(grid, mats, run)
    sim.upload(dfields, ifields)
    sim.set_tile_shape(a.tile_shape)
    bcs, loads = prob.velocity_bcs(), prob.particle_loads()
    if bcs:
        sim.set_velocity_bcs(bcs)
    if loads:
        sim.set_particle_loads(loads)
    if bcs or loads:
        print(f"# {len(bcs)} velocity conditions, {len(loads)} particle loads", file=sys.stderr, flush=True)
    sim.step(a.warmup)
    sim.sync()
    mass, _, _ = sim.grid_fields()
    active_nodes = int(np.count_nonzero(mass))        # (field, node) pairs with mass

    # timed region: K steps, HIP events between kernel groups on the simulation's stream
    sim.set_profiling(True)
    sim.sync()
    t1 = time.perf_counter()
    sim.step(a.steps)
    sim.sync()
    t2 = time.perf_counter()
    times = sim.kernel_times()
    sim.set_profiling(False)
    elapsed = t2 - t1
    value = npart * a.steps / elapsed

    # up to 100 further steps replayed from the captured two-step graph (a later stretch of the run)
    replay_steps = min(a.steps, 100)
    sim.sync()
    t3 = time.perf_counter()
    sim.step(replay_steps)
    sim.sync()
    graph_value = npart * replay_steps / (time.perf_counter() - t3)
    st = sim.stats()
    print(f"# eager+events {value:.4e} particle-steps/s, graph replay {graph_value:.4e}; "
          f"left grid {st.leftGrid}, nan {st.nanParticles}", file=sys.stderr, flush=True)
    archive = None
    if a.archive:
        # one archive of the final state (every particle row the device carries), written to a
        # temporary file: device download, record packing and the file write
        import tempfile
        order = "iYYYYYNNNNNYNY"
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "bench.000001.mpm")
            sim.sync()
            ta = time.perf_counter()
            sim.write_archive(path, order, time_ms=0.)
            tsec = time.perf_counter() - ta
            nbytes = os.path.getsize(path)
        archive = {"order": order, "bytes": nbytes, "seconds": tsec, "MB_per_s": nbytes / tsec / 1e6}
        print(f"# archive: {nbytes / 1e6:.1f} MB in {tsec:.3f} s ({nbytes / tsec / 1e9:.2f} GB/s)",
              file=sys.stderr, flush=True)
    for g, (ms, n) in times.items():
        print(f"#   {g:14s} {ms / max(n, 1):9.4f} ms/launch ({n} launches)", file=sys.stderr)

    dom = max(times, key=lambda g: times[g][0])
    dom_ms = times[dom][0] / max(times[dom][1], 1)
    bp, bn = ALG_BYTES[dom]
    if a.workload == "contact" and dom not in ("k_node_mass", "k_node_force", "k_node_mom2", "binning"):
        bp += CPDI_EXTRA
    alg_bytes = bp * npart + bn * active_nodes
    achieved = alg_bytes / (dom_ms * 1e-3)
    traffic, traffic_build = measured_traffic(dom, a.cells) if a.workload == "j2" else (None, None)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": dom,
                "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": dom_ms,
                "traffic_source": (f"table profiles/hbm_traffic_j2_block_128.md (rocprofv3 FETCH_SIZE/WRITE_SIZE, "
                                   f"build {traffic_build})" if traffic is not None else None)}

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        th = cpu_threads(a.cpu_threads)
        o = orc.Oracle(grid, mats, run, dfields, ifields)
        if bcs:
            o.set_velocity_bcs(bcs)
        if loads:
            o.set_particle_loads(loads)
        o.set_threads(th)
        # steps of the same block until the sample has run about --cpu-seconds (at least one)
        cpu_steps = cpu_s = 0
        while cpu_steps < max(1, a.cpu_steps) or cpu_s < a.cpu_seconds:
            tc = time.perf_counter()
            o.step(1)
            cpu_s += time.perf_counter() - tc
            cpu_steps += 1
            if cpu_steps >= 200:
                break
        cpu = {"value": npart * cpu_steps / cpu_s, "unit": "particle-steps/s", "cores": th, "kind": "port",
               "sample": f"{cpu_steps} step(s) of the same {npart}-particle block, oracle/mpm_oracle.cpp in its "
                         f"threaded mode on {th} threads of {cpu_model()}, {cpu_s:.1f} s"}
        del o

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload,
                   "particles": npart, "grid_cells": [grid.nx, grid.ny, grid.nz], "active_nodes": active_nodes,
                   "material": "Neohookean x2 + CoulombFriction" if a.workload == "contact" else "IsoPlasticity J2, linear hardening",
                   "shape": "lCPDI" if a.workload == "contact" else "uGIMP", "method": "USAVG+ FLIP",
                   "tile_shape": a.tile_shape, "graph_replay_value": graph_value,
                   **({"xpic": {"order": a.xpic_order, "fmpm": a.fmpm}} if a.xpic_order > 0 else {}),
                   **({"normals": a.normals} if a.workload == "contact" else {}),
                   **({"archive": archive} if archive else {})},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    sim.close()
    if a.workload != "contact" and a.served_requests > 0:
        # the served half of the metric: a short served stream against the same solves one at a time
        out["served"], _ = served_stream(a.served_requests, a.rate, a.graph_steps, a.max_resident)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
