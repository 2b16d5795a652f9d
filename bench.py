#!/usr/bin/env python3
"""Benchmark of the first-order FV Euler step on one MI355X.

A step is one pass of the hot path over the whole grid: the device ghost fill plus the
fused flux/update/Dt kernel, with Dt, t and the iteration count resident on the device
(the step loop of hydro2d/main.cxx:689-815).  Default workload: the plain 2D blast on a
40000 x 20000 grid in FP64 (the reference's built-in IC scaled to the grid, periodic X,
reflective Y), 100 timed steps after a short warmup, state resident in HBM.

Prints ONE JSON line: GCUPS (1e9 cell updates per second) over the timed steps, the HBM
roofline of the step kernel (algorithmic 64 B per 2D cell update, 80 B in 3D, over the
kernel's average HIP-event duration) and the CPU oracle timed on this host's cores.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
# Published compute rate of the same workload (2D 40000 x 20000 FP64 blast) on the
# reference's hardware: 1018.004 Mc/s on two IA-840F FPGAs
# (hydro2d/reference_run/res_2000_40000_20000_2FPGAS.res:10108).
PUBLISHED_GCUPS = {"blast2d_40000x20000_fp64": 1.018004}

WORKLOADS = {
    # name: (dim, nb, C, ic_mode, radius_frac, p_in, p_out, bytes per cell update, precision)
    "blast2d_40000x20000_fp64": (2, (40000, 20000), 0.5, 0, 0.0, 10.0, 0.1, 64, 64),
    "blast3d_512_fp64": (3, (512, 512, 512), 0.4, 1, 0.1, 10.0, 0.1, 80, 64),
    "blast2d_40000x20000_fp32flux": (2, (40000, 20000), 0.5, 0, 0.0, 10.0, 0.1, 64, 3264),
    "blast3d_512_fp32flux": (3, (512, 512, 512), 0.4, 1, 0.1, 10.0, 0.1, 80, 3264),
    # The reference's USE_FLOAT build: FP32 state and arithmetic (32 / 40 B per cell).
    "blast2d_40000x20000_fp32": (2, (40000, 20000), 0.5, 0, 0.0, 10.0, 0.1, 32, 32),
    "blast3d_512_fp32": (3, (512, 512, 512), 0.4, 1, 0.1, 10.0, 0.1, 40, 32),
}
ALIASES = {"2d": "blast2d_40000x20000_fp64", "3d": "blast3d_512_fp64", "2d-fp32flux": "blast2d_40000x20000_fp32flux",
           "3d-fp32flux": "blast3d_512_fp32flux", "2d-fp32": "blast2d_40000x20000_fp32", "3d-fp32": "blast3d_512_fp32",
           "batched": "batched2d_256_fp32flux",
           "batched-fp64": "batched2d_256_fp64"}
# Batched workloads: 256 seeded 2D blasts (sides log-uniform in [64, 2048]), precision.
BATCHED = {"batched2d_256_fp32flux": 3264, "batched2d_256_fp64": 64}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads() -> int:
    """The host cores this process may use: its CPU affinity, or OMP_NUM_THREADS where
    the host sets it lower (a shared box gives each job a share of its cores)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(dim: int, budget_s: float = 12.0, serial_budget_s: float = 7.0, ranks_budget_s: float = 7.0) -> dict:
    """The CPU oracle (a restatement of the reference step in its operation order) timed
    on this host on a bounded sample of the same blast -- 4000 x 2000 in 2D, 128^3 in
    3D -- three ways: OpenMP over X rows on every core this process may use (`value`);
    the reference's two-rank shape (`mpirun -np 2 --sdx 2`: each X half stepped by one
    serial thread, the single_task of each rank, hydro2d/main.cxx:621-657); and serial
    (one thread, the single-rank build)."""
    import numpy as np

    from oracle import oracle as O

    nproc = os.cpu_count() or 1
    threads = host_threads()
    if dim == 2:
        nb = (4000, 2000)
        rhoE, mom, dt0 = O.ic2d(*nb)
        run = lambda n, th, st: O.run2d(st[0], st[1], dt0, 2.0, n, nthreads=th)  # noqa: E731
    else:
        nb = (128, 128, 128)
        rhoE, mom, dt0 = O.ic3d(*nb, ic_mode=1, radius_frac=0.1)
        run = lambda n, th, st: O.run3d(st[0], st[1], dt0, 2.0, n, nthreads=th)  # noqa: E731
    cells = int(np.prod(nb))

    def timed(th, budget):
        st = (rhoE.copy(), mom.copy())
        t0 = time.perf_counter()
        run(2, th, st)
        per_step = (time.perf_counter() - t0) / 2
        steps = int(max(2, min(2000, budget / max(per_step, 1e-6))))
        st = (rhoE.copy(), mom.copy())
        t0 = time.perf_counter()
        res = run(steps, th, st)
        el = time.perf_counter() - t0
        return cells * res["it"] / el / 1e9, res["it"], el

    par, par_it, par_s = timed(threads, budget_s)
    two, two_it, two_s = timed(2, ranks_budget_s)
    ser, ser_it, ser_s = timed(1, serial_budget_s)
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": round(par, 5),
        "unit": "GCUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{'x'.join(map(str, nb))} blast, FP64: {par_it} steps on {threads} OpenMP threads ({par_s:.1f} s); "
                  f"host nproc {nproc}, {threads} usable{', ' + cpu_model if cpu_model else ''}",
        "two_ranks": {"value": round(two, 5), "unit": "GCUPS", "cores": 2,
                      "sample": f"{two_it} steps, one serial thread per X half ({two_s:.1f} s)"},
        "serial": {"value": round(ser, 5), "unit": "GCUPS", "cores": 1,
                   "sample": f"{ser_it} steps, 1 thread ({ser_s:.1f} s)"},
    }


class EnergyMeter:
    """GPU energy over an interval from the amdsmi energy accumulator (the GPU analogue of
    the reference's energy_record.sh power sampling).  Unavailable counters leave
    `error` set and readings None."""

    def __init__(self):
        self.error = None
        self.h = None
        try:
            import amdsmi
            import torch

            self.smi = amdsmi
            amdsmi.amdsmi_init()
            handles = amdsmi.amdsmi_get_processor_handles()
            want = None
            try:
                pr = torch.cuda.get_device_properties(0)
                want = (int(pr.pci_domain_id), int(pr.pci_bus_id), int(pr.pci_device_id))
            except Exception:
                pass
            for h in handles:
                bdf = amdsmi.amdsmi_get_gpu_device_bdf(h)  # "dddd:bb:dd.f"
                dom, bus, rest = bdf.split(":")
                if want and (int(dom, 16), int(bus, 16), int(rest.split(".")[0], 16)) == want:
                    self.h = h
            if self.h is None and len(handles) == 1:
                self.h = handles[0]
            if self.h is None:
                self.error = f"cannot match the device among {len(handles)} amdsmi handles"
        except Exception as e:  # noqa: BLE001
            self.error = f"amdsmi unavailable: {e}"

    def sample(self):
        """(counter timestamp in ns, accumulated joules), or None."""
        if self.h is None:
            return None
        try:
            e = self.smi.amdsmi_get_energy_count(self.h)
            return int(e["timestamp"]), e["energy_accumulator"] * e["counter_resolution"] * 1e-6
        except Exception as e:  # noqa: BLE001
            self.error = f"energy counter: {e}"
            return None

    def busy_power(self, busy):
        """Average power while `busy()` keeps the device working (it enqueues work and
        waits for it).  The accumulator advances about once a second, so the power is
        taken between two of its own updates (joules over its own timestamps), using only
        updates that fall inside the busy period; a host-timed window would catch a
        variable number of updates and partly idle time.  Returns (watts, seconds
        spanned) or None."""
        import threading

        samples = []
        stop = threading.Event()

        def poll():
            last = None
            while not stop.is_set():
                cur = self.sample()
                if cur is None:
                    return
                if last is None or cur[0] != last[0]:
                    samples.append((time.perf_counter(), cur[0], cur[1]))
                    last = cur
                time.sleep(0.005)

        th = threading.Thread(target=poll, daemon=True)
        th.start()
        h0 = time.perf_counter()
        busy()
        h1 = time.perf_counter()
        stop.set()
        th.join()
        inside = [x for x in samples if h0 < x[0] <= h1]
        # The first update inside may close an interval that began before the work did.
        if len(inside) < 3:
            self.error = self.error or f"energy counter advanced {len(inside)} times in {h1 - h0:.1f} s of work"
            return None
        (_, t0, j0), (_, t1, j1) = inside[1], inside[-1]
        return (j1 - j0) / ((t1 - t0) * 1e-9), (t1 - t0) * 1e-9


def load_traffic(workload: str):
    p = ROOT / "profiles" / "traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get(workload, {}).get("bytes_per_launch")
        except (OSError, ValueError):
            return None
    return None


def _blast_bufs(torch, nb):
    shape = tuple(n + 2 for n in nb)
    a = [torch.zeros(shape + (2,), dtype=torch.float64, device="cuda"),
         torch.zeros(shape + (len(nb),), dtype=torch.float64, device="cuda")]
    return a, [torch.zeros_like(a[0]), torch.zeros_like(a[1])]


def _prims2(torch, rhoE, mom):
    """rho, p, u, v of the interior cells (gamma = 1.4)."""
    rho = rhoE[1:-1, 1:-1, 0]
    E = rhoE[1:-1, 1:-1, 1]
    u = mom[1:-1, 1:-1, 0] / rho
    v = mom[1:-1, 1:-1, 1] / rho
    p = 0.4 * (E - 0.5 * rho * (u * u + v * v))
    return {"rho": rho, "p": p, "u": u, "v": v}


def _batch_error_vs_fp64(hydro, torch, blasts, batch, steps):
    """Worst relative L-inf per field over all blasts of an FP32-flux batch against an
    FP64 batch of the same blasts after the same number of steps: rho and p relative to
    their own maximum, velocities relative to the FP64 solve's maximum signal speed."""
    ref_bufs = [_blast_bufs(torch, p["nb"]) for p in blasts]
    torch.cuda.synchronize()
    ref = hydro.Batch(2, hydro.FP64, len(blasts))
    for p, (a, b) in zip(blasts, ref_bufs):
        ref.add_blast(a, b, p["C"], p["radius_frac"], p["p_in"], p["p_out"], p["rho_in"], p["rho_out"],
                      max_t=1e30, max_it=10**9)
    ref.advance(steps)
    ref.sync()
    worst = {"rho": 0.0, "p": 0.0, "u": 0.0, "v": 0.0}
    own = {"u": 0.0, "v": 0.0}
    for slot in range(len(blasts)):
        if ref.status(slot)["it"] != steps or batch.status(slot)["it"] != steps:
            raise RuntimeError(f"FP64 reference of blast {slot} is not at step {steps}")
        g = _prims2(torch, *batch.current(slot))
        w = _prims2(torch, *ref.current(slot))
        vscale = float(torch.max(torch.sqrt(1.4 * w["p"].abs() / w["rho"]) + torch.sqrt(w["u"] ** 2 + w["v"] ** 2)))
        for k in worst:
            scale = vscale if k in ("u", "v") else float(w[k].abs().max())
            worst[k] = max(worst[k], float((g[k] - w[k]).abs().max()) / scale)
        for k in own:
            # against the component's own maximum, where that maximum carries flow (at
            # least 1% of the signal speed) rather than round-off
            m = float(w[k].abs().max())
            if m >= 1e-2 * vscale:
                own[k] = max(own[k], float((g[k] - w[k]).abs().max()) / m)
    ref.close()
    del ref_bufs
    return {**{k: float(f"{v:.3e}") for k, v in worst.items()},
            **{f"{k}_own_max": float(f"{v:.3e}") for k, v in own.items()}, "steps": int(steps), "bound": 1e-5,
            "metric": "worst relative L-inf over the blasts: rho and p against their own max, u and v against "
                      "the FP64 solve's max signal speed; u_own_max and v_own_max against each component's own "
                      "max where that is at least 1% of the signal speed"}


def bench_batched(args, name: str) -> dict:
    """256 independent 2D blasts of uneven sizes in one batch: autotuned tiling, steps
    replayed from a hipGraph, K timed steps.  Beside it, the same blasts solved one at a
    time (a sample of steps per blast): the device-resident single-solve loop, and the
    reference loop shape (ghost fill, synchronous launcher, Dt read back per step)."""
    import numpy as np
    import torch

    hydro = importlib.import_module("scalable-fpga-hydrodynamics_amd")
    wl = importlib.import_module("scalable-fpga-hydrodynamics_amd.workloads")
    precision = BATCHED[name]
    blasts = wl.batched_blasts(256)
    cells = sum(int(np.prod(p["nb"])) for p in blasts)
    log(f"[bench] {name}: {len(blasts)} blasts, {cells / 1e6:.1f} M cells, "
        f"{sum(wl.state_bytes(p['nb']) for p in blasts) / 1e9:.2f} GB of state")
    bufs = [_blast_bufs(torch, p["nb"]) for p in blasts]
    torch.cuda.synchronize()
    batch = hydro.Batch(2, precision, len(blasts))
    for p, (a, b) in zip(blasts, bufs):
        batch.add_blast(a, b, p["C"], p["radius_frac"], p["p_in"], p["p_out"], p["rho_in"], p["rho_out"],
                        max_t=1e30, max_it=10**9)
    t0 = time.perf_counter()
    tune_ms = batch.autotune(steps_per_trial=4, use_cache=False)
    tuning = batch.tiling()
    log(f"[bench] autotune {time.perf_counter() - t0:.2f} s: blocks/CU {tuning[0]}, tile rows {tuning[1]}, "
        f"tile windows {tuning[2]} "
        f"({tune_ms:.3f} ms/step)")
    graph_steps = 10
    batch.set_graph(graph_steps)
    W, K = args.warmup, args.steps
    batch.advance(W)
    batch.sync()
    it0 = batch.status(0)["it"]
    ext = torch.cuda.ExternalStream(batch.stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record(ext)
    batch.advance(K)
    e1.record(ext)
    batch.sync()
    wall = time.perf_counter() - w0
    ms = e0.elapsed_time(e1)
    elapsed = max(wall, ms / 1e3)
    gcups = cells * K / elapsed / 1e9
    per_step = elapsed / K
    flags = batch.done_flags(len(blasts))
    st = [batch.status(s) for s in range(len(blasts))]
    bad = [i for i, x in enumerate(st) if x["rc"] != 0 or x["it"] != it0 + K]
    if any(flags) or bad:
        raise RuntimeError(f"batched run: unexpected stop flags / status {flags.nonzero()} {bad}")
    log(f"[bench] batched: {K} steps in {elapsed * 1e3:.1f} ms, {gcups:.2f} GCUPS")
    err = None
    if precision == hydro.FP32_FLUX:
        err = _batch_error_vs_fp64(hydro, torch, blasts, batch, st[0]["it"])
        log(f"[bench] FP32-flux vs FP64 after {err['steps']} steps, worst rel L-inf: "
            + ", ".join(f"{k2} {err[k2]:.2e}" for k2 in ("rho", "p", "u", "v")))
    batch.close()

    # One solve at a time, same blasts, a sample of steps each (GCUPS is per-step).
    ks = max(1, min(K, args.plain_steps))
    k = hydro.constants
    t_res = 0.0
    t_sync = 0.0
    n_sync = min(len(blasts), args.plain_sync_blasts)
    for i, (p, (a, b)) in enumerate(zip(blasts, bufs)):
        kk = k(p["nb"], p["C"])
        dt0 = hydro.ic(a[0], a[1], kk["C_min"], ic_mode=1, radius_frac=p["radius_frac"], p_in=p["p_in"],
                       p_out=p["p_out"], rho_in=p["rho_in"], rho_out=p["rho_out"])
        run = hydro.Run(p["nb"], kk["D"], kk["C_min"], dt0, 1e30, ks, precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run.advance(a, b, ks)
        torch.cuda.synchronize()
        t_res += time.perf_counter() - t0
        run.close()
        if i < n_sync:
            # The reference loop shape (hydro2d/main.cxx:712-815): ghost fill, synchronous
            # launcher, Dt read back and DtDx / DtDy recomputed on the host every step.
            dt = hydro.ic(a[0], a[1], kk["C_min"], ic_mode=1, radius_frac=p["radius_frac"], p_in=p["p_in"],
                          p_out=p["p_out"], rho_in=p["rho_in"], rho_out=p["rho_out"])
            cur, nxt = a, b
            dtn = torch.zeros(1, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(ks):
                hydro.ghost_fill2d(cur[0], cur[1])
                hydro.launcher(cur[0], cur[1], nxt[0], nxt[1], dtn, kk["C_min"], kk["gamma"], kk["gmo"],
                               kk["divgamma"], kk["K"], dt / kk["D"][0], dt / kk["D"][1])
                dt = float(dtn.item())
                cur, nxt = nxt, cur
            t_sync += time.perf_counter() - t0
    sync_cells = sum(int(np.prod(p["nb"])) for p in blasts[:n_sync])
    plain_res = cells * ks / t_res / 1e9
    plain_sync = sync_cells * ks / t_sync / 1e9 if n_sync else None
    log(f"[bench] one solve at a time: resident {plain_res:.2f} GCUPS, synchronous launcher loop "
        f"{plain_sync if plain_sync is None else round(plain_sync, 2)} GCUPS")
    return {
        "metric": "cell-updates/s (GCUPS) + % HBM roofline",
        "value": round(gcups, 4),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(per_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == 64 else "f64 state, f32 flux",
        "data": "synthetic (256 seeded blasts, splitmix64 seed 0x5EED, IC generated on the device)",
        "config": {"workload": name, "blasts": len(blasts), "cells": cells, "sides": "log-uniform [64, 2048]",
                   "tiling": {"blocks_per_cu": tuning[0], "tile_rows": tuning[1], "tile_windows": tuning[2]},
                   "graph_steps": graph_steps,
                   "kernel": "batch2d (one launch per step: direct dispatch of tiles or a persistent grid, "
                             "as tuned) + batch finalize"},
        "roofline": {"bound": "hbm", "achieved": round(64 * cells / per_step / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(64 * cells / per_step / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(name)},
        "cpu_baseline": None,
        "plain_resident_gcups": round(plain_res, 4),
        "plain_sync_launcher_gcups": round(plain_sync, 4) if plain_sync else None,
        "speedup_vs_plain_resident": round(gcups / plain_res, 3),
        "speedup_vs_plain_sync_launcher": round(gcups / plain_sync, 3) if plain_sync else None,
        "plain_sample": f"{ks} steps per blast, all {len(blasts)} blasts (resident) / first {n_sync} (synchronous)",
        "error_vs_fp64": err,
    }


def _latency_by_dim(reqs, lat) -> dict:
    """p50 / p99 of the latencies of the 2D and of the 3D requests separately."""
    sv = importlib.import_module("scalable-fpga-hydrodynamics_amd.serve")
    out = {}
    for name, dim in (("2d", 2), ("3d", 3)):
        xs = [x for r, x in zip(reqs, lat) if len(r["nb"]) == dim]
        if xs:
            out[name] = {"requests": len(xs), "p50_s": round(sv.percentile(xs, 50), 4),
                         "p99_s": round(sv.percentile(xs, 99), 4)}
    return out


def _large_class(reqs, lat, seq_lat, dur, large_s) -> dict:
    """Latencies of the requests whose measured plain solve time is at least large_s
    (about the server's large size class, which it picks from an estimate), served and
    one at a time."""
    sel = [i for i in range(len(reqs)) if large_s > 0 and dur[i] >= large_s]
    if not sel:
        return {"requests": 0}
    return {"requests": len(sel), "served_max_s": round(max(lat[i] for i in sel), 4),
            "sequential_max_s": round(max(seq_lat[i] for i in sel), 4)}


def bench_served(args) -> dict:
    """Poisson stream of mixed solves (70% 2D sides 128..4096, 30% 3D sides 64..256,
    100..2000 steps, FP32-flux / FP64 50/50, seeded blasts) through the continuous-batching
    server.  The arrival rate is set for 80% load of a server that solves one request at a
    time on the plain device-resident loop: rate = 0.8 / mean plain solve time, from the
    measured solve time of every request.  Reported: p50/p99 latency and throughput of the
    served form at that rate, of the sequential server at the same arrivals (FIFO queueing
    of the measured solve times), and the saturation throughput of both (all requests at
    t = 0)."""
    import torch

    hydro = importlib.import_module("scalable-fpga-hydrodynamics_amd")
    wl = importlib.import_module("scalable-fpga-hydrodynamics_amd.workloads")
    sv = importlib.import_module("scalable-fpga-hydrodynamics_amd.serve")
    n = args.requests
    reqs = wl.served_requests(n, rate_per_s=1.0)
    work = sum(sv.cell_updates(r) for r in reqs)
    log(f"[bench] served: {n} requests, {work / 1e9:.1f} G cell-updates, "
        f"{sum(1 for r in reqs if len(r['nb']) == 2)} 2D / {sum(1 for r in reqs if len(r['nb']) == 3)} 3D, "
        f"{sum(1 for r in reqs if r['fp32flux'])} FP32-flux")
    prog = lambda d, t: d % 20 == 0 and log(f"[bench]   {d}/{t}")  # noqa: E731
    log("[bench] plain solves, one at a time ...")
    dur = sv.plain_solve_times(hydro, torch, reqs, progress=prog)
    torch.cuda.empty_cache()
    log("[bench] the same solves in the reference's loop shape (synchronous launcher per step) ...")
    dur_sync = sv.plain_solve_times(hydro, torch, reqs, progress=prog, mode="sync")
    torch.cuda.empty_cache()
    mean_d = sum(dur) / n
    rate = 0.8 / mean_d
    for r in reqs:
        r["arrival"] = r["arrival"] / rate
    seq_lat = sv.sequential_latencies(reqs, dur)
    seq_finish = max(r["arrival"] + l for r, l in zip(reqs, seq_lat))
    sync_lat = sv.sequential_latencies(reqs, dur_sync)
    sync_finish = max(r["arrival"] + l for r, l in zip(reqs, sync_lat))
    arena = int(args.serve_hbm_gb * 2**30)
    log(f"[bench] served at {rate:.2f} requests/s (80% of the sequential server) ...")
    srv = hydro.Server(arena, capacity=256, steps_per_tick=args.steps_per_tick)
    srv.set_limits(args.serve_cap_cells, args.steps_per_tick)
    srv.set_policy(args.serve_large_ms / 1e3)
    srv.set_backlog(args.serve_large_backlog)
    live = sv.run_trace(hydro, srv, reqs, realtime=True, progress=prog)
    srv.close()
    real_seq = None
    if not args.no_real_sequential:
        log("[bench] the one-at-a-time server run for real at the same arrivals ...")
        torch.cuda.empty_cache()
        real_seq = sv.run_sequential(hydro, torch, reqs, progress=prog)
        torch.cuda.empty_cache()
    log("[bench] served at saturation (all requests at t = 0) ...")
    srv = hydro.Server(arena, capacity=256, steps_per_tick=args.steps_per_tick)
    srv.set_limits(args.serve_cap_cells, args.steps_per_tick)
    srv.set_policy(args.serve_large_ms / 1e3)
    srv.set_backlog(args.serve_large_backlog)
    sat = sv.run_trace(hydro, srv, reqs, realtime=False, progress=prog)
    srv.close()
    seq_gcups = work / sum(dur) / 1e9
    line = {
        "metric": "cell-updates/s (GCUPS) + p50/p99 solve latency when served",
        "value": round(sat["gcups"], 4),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state; f32 flux for half the requests",
        "data": f"synthetic Poisson stream, splitmix64 seed 0x5EED, {n} requests",
        "config": {"workload": "served_mixed", "requests": n, "mix": "70% 2D 128..4096, 30% 3D 64..256",
                   "steps": "100..2000", "precision": "50% fp32-flux / 50% fp64", "hbm_arena_gb": args.serve_hbm_gb,
                   "steps_per_tick": args.steps_per_tick, "engine_cap_cells": args.serve_cap_cells,
                   "large_class_ms": args.serve_large_ms, "large_backlog": args.serve_large_backlog,
                   "arrival_rate_per_s": round(rate, 3),
                   "load": "0.8 of the one-at-a-time plain server"},
        "served": {"p50_s": round(live["p50"], 4), "p99_s": round(live["p99"], 4), "gcups": round(live["gcups"], 3),
                   "makespan_s": round(live["makespan"], 3)},
        "sequential_plain": {"p50_s": round(sv.percentile(seq_lat, 50), 4),
                             "p99_s": round(sv.percentile(seq_lat, 99), 4),
                             "gcups": round(work / (seq_finish - reqs[0]["arrival"]) / 1e9, 3),
                             "mean_solve_s": round(mean_d, 4)},
        "sequential_plain_real": ({"p50_s": round(real_seq["p50"], 4), "p99_s": round(real_seq["p99"], 4),
                                   "gcups": round(real_seq["gcups"], 3),
                                   "note": "the same one-at-a-time server run in real time, IC and buffers included"}
                                  if real_seq else None),
        "sequential_sync_launcher": {"p50_s": round(sv.percentile(sync_lat, 50), 4),
                                     "p99_s": round(sv.percentile(sync_lat, 99), 4),
                                     "gcups": round(work / (sync_finish - reqs[0]["arrival"]) / 1e9, 3),
                                     "mean_solve_s": round(sum(dur_sync) / n, 4)},
        "by_dimension": {"served": _latency_by_dim(reqs, live["latency"]),
                         "sequential_plain": _latency_by_dim(reqs, seq_lat)},
        "large_class": _large_class(reqs, live["latency"], seq_lat, dur, args.serve_large_ms / 1e3),
        "saturation": {"served_gcups": round(sat["gcups"], 3), "sequential_plain_gcups": round(seq_gcups, 3),
                       "speedup": round(sat["gcups"] / seq_gcups, 3),
                       "sequential_sync_launcher_gcups": round(work / sum(dur_sync) / 1e9, 3),
                       "speedup_vs_sync_launcher": round(sat["gcups"] / (work / sum(dur_sync) / 1e9), 3),
                       "served_p50_s": round(sat["p50"], 4), "served_p99_s": round(sat["p99"], 4)},
        "roofline": None,
        "cpu_baseline": None,
    }
    log(f"[bench] served p50 {live['p50']:.3f} s p99 {live['p99']:.3f} s | sequential p50 "
        f"{line['sequential_plain']['p50_s']} s p99 {line['sequential_plain']['p99_s']} s"
        + (f" (real run: p50 {real_seq['p50']:.3f} s p99 {real_seq['p99']:.3f} s)" if real_seq else "")
        + f" | saturation {sat['gcups']:.1f} vs {seq_gcups:.1f} GCUPS")
    return line


def served_in_child(args) -> dict:
    """The served measurement in a fresh process (python bench.py --workload served), as
    its own workload runs it, so that it shares nothing with the forms measured before
    it.  (Run after the batched form, the served p99 once came out 0.64 s instead of
    0.48 s: the streams created before had changed which of HIP's 4 hardware queues the
    server's streams shared, and an engine's ticks then waited behind a large solve
    enqueued whole.  Large solves are now fed in chunks, which removed that.)"""
    import subprocess

    cmd = [sys.executable, str(pathlib.Path(__file__).resolve()), "--workload", "served", "--requests",
           str(args.requests), "--serve-hbm-gb", str(args.serve_hbm_gb), "--steps-per-tick", str(args.steps_per_tick),
           "--serve-cap-cells", str(args.serve_cap_cells), "--serve-large-ms", str(args.serve_large_ms),
           "--serve-large-backlog", str(args.serve_large_backlog)]
    if args.no_real_sequential:
        cmd.append("--no-real-sequential")
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, check=False)
    if p.returncode != 0:
        raise RuntimeError(f"served measurement failed (exit {p.returncode})")
    return json.loads(p.stdout.strip().splitlines()[-1])


def _prims3(torch, rhoE, mom):
    """rho, p, u, v, w of the interior cells (gamma = 1.4), float64."""
    sl = (slice(1, -1),) * 3
    rho = rhoE[sl + (0,)]
    E = rhoE[sl + (1,)]
    u, v, w = (mom[sl + (c,)] / rho for c in range(3))
    p = 0.4 * (E - 0.5 * rho * (u * u + v * v + w * w))
    return {"rho": rho, "p": p, "u": u, "v": v, "w": w}


def bench_c3(args) -> dict:
    """The 512^3 blast (sphere of radius 0.1 L, pressure ratio 100, periodic X, reflective
    Y/Z) advanced 200 steps from the same IC in FP64 and in the FP32-flux form, each timed
    on its own resident loop; the FP32-flux fields' per-field relative L-inf and L2 error
    against the FP64 fields after those 200 steps, and both forms' mass and energy drift."""
    import numpy as np
    import torch

    hydro = importlib.import_module("scalable-fpga-hydrodynamics_amd")
    nb = (512, 512, 512)
    steps = 200
    cells = int(np.prod(nb))
    k = hydro.constants(nb, 0.4)
    shape = tuple(n + 2 for n in nb)
    out = {}
    for prec, tag in ((hydro.FP64, "fp64"), (hydro.FP32_FLUX, "fp32flux")):
        a = [torch.zeros(shape + (2,), dtype=torch.float64, device="cuda"),
             torch.zeros(shape + (3,), dtype=torch.float64, device="cuda")]
        b = [torch.zeros_like(a[0]), torch.zeros_like(a[1])]
        dt0 = hydro.ic(a[0], a[1], k["C_min"], ic_mode=1, radius_frac=0.1, p_in=10.0, p_out=0.1)
        m0, e0 = hydro.total_mass(a[0]), hydro.total_energy(a[0])
        run = hydro.Run(nb, k["D"], k["C_min"], dt0, 2.0, steps, prec)
        run.set_timing(True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        run.advance(a, b, steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        kms, n = run.kernel_time()
        st = run.status()
        run.close()
        if st["rc"] != 0 or st["it"] != steps:
            raise RuntimeError(f"c3 {tag} run failed: {st}")
        cur = a if st["current"] == 0 else b
        del a, b  # the other pair is freed; `cur` keeps the final state
        out[tag] = {"state": cur, "t": st["t"], "gcups": cells * steps / wall / 1e9, "kernel_ms": kms / n,
                    "mass_rel_change": abs(hydro.total_mass(cur[0]) - m0) / m0,
                    "energy_rel_change": abs(hydro.total_energy(cur[0]) - e0) / e0}
        log(f"[bench] c3 {tag}: {steps} steps, t={st['t']:.5e}, {out[tag]['gcups']:.2f} GCUPS, kernel "
            f"{out[tag]['kernel_ms']:.3f} ms")
    ref = _prims3(torch, *out["fp64"]["state"])
    got = _prims3(torch, *out["fp32flux"]["state"])
    vel = torch.sqrt(ref["u"] ** 2 + ref["v"] ** 2 + ref["w"] ** 2)
    vscale = float(torch.max(torch.sqrt(1.4 * ref["p"].abs() / ref["rho"]) + vel))
    del vel
    errors, own = {}, {}
    for key in ("rho", "p", "u", "v", "w"):
        d = (got[key] - ref[key]).abs()
        scale = vscale if key in ("u", "v", "w") else float(ref[key].abs().max())
        errors[key] = {"linf": float(f"{float(d.max()) / scale:.3e}"),
                       "l2": float(f"{float(torch.sqrt(torch.mean(d * d))) / scale:.3e}")}
        if key in ("u", "v", "w"):
            m = float(ref[key].abs().max())
            own[key] = float(f"{float(d.max()) / m:.3e}") if m >= 1e-2 * vscale else None
        del d
    f = out["fp32flux"]
    achieved = 80 * cells / (f["kernel_ms"] / 1e3) / 1e9
    log("[bench] c3 FP32-flux vs FP64 after 200 steps, relative L-inf / L2: "
        + ", ".join(f"{key} {e['linf']:.2e}/{e['l2']:.2e}" for key, e in errors.items()))
    return {
        "metric": "cell-updates/s (GCUPS) + % HBM roofline",
        "value": round(f["gcups"], 4),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": round(1e3 * cells / (f["gcups"] * 1e9), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state, f32 flux",
        "data": "synthetic (parameterised blast IC generated on the device)",
        "config": {"workload": "blast3d_512_c3", "grid": list(nb), "cells": cells, "steps": steps,
                   "ic": "sphere radius 0.1 L, p_in 10, p_out 0.1, rho 1.0 / 1.2",
                   "boundaries": "periodic X, reflective Y/Z", "kernel": "fvm3d (FP32-flux and FP64)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic("blast3d_512_fp32flux")},
        "fp64": {"gcups": round(out["fp64"]["gcups"], 4), "kernel_ms": round(out["fp64"]["kernel_ms"], 4),
                 "t": out["fp64"]["t"], "mass_rel_change": out["fp64"]["mass_rel_change"],
                 "energy_rel_change": out["fp64"]["energy_rel_change"]},
        "fp32flux": {"gcups": round(f["gcups"], 4), "kernel_ms": round(f["kernel_ms"], 4), "t": f["t"],
                     "mass_rel_change": f["mass_rel_change"], "energy_rel_change": f["energy_rel_change"]},
        "error_vs_fp64": {**errors, "velocity_own_max_linf": own, "bound": 1e-5,
                          "metric": "relative to each field's max for rho and p, to the FP64 solve's max signal "
                                    "speed for u, v, w (and, where at least 1% of it, to each component's own max)"},
    }


def bench_csv(args) -> dict:
    """The CSV output path of the command-line driver at scale: the same run with and
    without output; the difference is the cost of writing one final file of nb_x * nb_y
    rows (byte-identical to the reference's std::ostream formatting)."""
    import subprocess
    import tempfile

    exe = ROOT / "scalable-fpga-hydrodynamics_amd" / "hydro"
    nx, ny = args.csv_grid
    base = [str(exe), "-x", str(nx), "-y", str(ny), "-i", "2", "-t", "2"]
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        t0 = time.perf_counter()
        subprocess.run(base + ["--no-output"], cwd=d, check=True, capture_output=True)
        t_no = time.perf_counter() - t0
        t0 = time.perf_counter()
        subprocess.run(base + ["-o", "out"], cwd=d, check=True, capture_output=True)
        t_out = time.perf_counter() - t0
        files = sorted(pathlib.Path(d).glob("out_*.csv"))
        nbytes = sum(f.stat().st_size for f in files)
    rows = nx * ny * len(files)
    extra = max(t_out - t_no, 1e-9)
    log(f"[bench] csv: {len(files)} files, {rows / 1e6:.1f} M rows, {nbytes / 1e9:.2f} GB in {extra:.2f} s extra")
    return {"metric": "CSV rows/s (driver output path)", "value": round(rows / extra, 1), "unit": "rows/s",
            "n_gpus": 1, "higher_is_better": True, "config": {"workload": f"csv_{nx}x{ny}", "files": len(files),
                                                             "rows": rows, "bytes": nbytes},
            "gb_per_s": round(nbytes / extra / 1e9, 3), "seconds_with_output": round(t_out, 3),
            "seconds_without_output": round(t_no, 3)}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="2d")
    ap.add_argument("--rows", type=int, default=0, help="override the kernel's X rows per strip")
    ap.add_argument("--occ", type=int, default=0, help="override the 2D kernel's workgroups per CU (3..8)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-copy-reference", action="store_true", help="skip the achievable-bandwidth copy")
    ap.add_argument("--no-energy", action="store_true",
                    help="skip the energy window (about 5.5 s of further steps at the power limit), e.g. "
                         "under a profiler, so that every traced step launch is a timed one")
    ap.add_argument("--no-forms", action="store_true",
                    help="default workload only: skip the batched and served summaries")
    ap.add_argument("--plain-steps", type=int, default=20, help="steps per blast of the one-solve-at-a-time baselines")
    ap.add_argument("--plain-sync-blasts", type=int, default=256)
    ap.add_argument("--requests", type=int, default=120, help="served workload: number of requests")
    ap.add_argument("--serve-hbm-gb", type=float, default=48.0, help="served workload: state arena (GiB)")
    ap.add_argument("--steps-per-tick", type=int, default=20)
    ap.add_argument("--serve-large-ms", type=float, default=20.0,
                    help="served workload: requests whose estimated solve time alone is at least this many ms run "
                         "one at a time in arrival order beside the smaller ones (0: one class)")
    ap.add_argument("--serve-large-backlog", type=int, default=8,
                    help="served workload: while more than this many large requests wait, the oldest join the "
                         "batched engines (0: never)")
    ap.add_argument("--no-real-sequential", action="store_true",
                    help="served workload: skip the real one-at-a-time server run")
    ap.add_argument("--serve-cap-cells", type=float, default=6e6,
                    help="served workload: per-engine cap on the interior cells of running requests (0 = none)")
    ap.add_argument("--csv-grid", type=int, nargs=2, default=[6000, 5000], help="csv workload: grid of the driver run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (the reference for HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X and the "
              "solver keeps one resident domain per GPU; run with --gpus 1.", file=sys.stderr)
        return 2

    name = ALIASES.get(args.workload, args.workload)
    if name == "csv":
        print(json.dumps(bench_csv(args)), flush=True)
        return 0
    if name == "served":
        import torch

        if not torch.cuda.is_available():
            log("no GPU visible")
            return 1
        print(json.dumps(bench_served(args)), flush=True)
        return 0
    if name == "c3":
        import torch

        if not torch.cuda.is_available():
            log("no GPU visible")
            return 1
        print(json.dumps(bench_c3(args)), flush=True)
        return 0
    if name in BATCHED:
        import torch

        if not torch.cuda.is_available():
            log("no GPU visible")
            return 1
        print(json.dumps(bench_batched(args, name)), flush=True)
        return 0
    if name not in WORKLOADS:
        log(f"unknown workload {args.workload}; choose from {sorted(WORKLOADS)}")
        return 2
    dim, nb, C_cfl, ic_mode, rfrac, p_in, p_out, bpc, precision = WORKLOADS[name]

    import numpy as np
    import torch

    hydro = importlib.import_module("scalable-fpga-hydrodynamics_amd")
    if not torch.cuda.is_available():
        log("no GPU visible")
        return 1
    hydro.lib()
    if args.rows:
        hydro.tune_set("rows2d" if dim == 2 else "rows3d", args.rows)
    if args.occ and dim == 2:
        hydro.tune_set("occ2d", args.occ)

    fp32_state = precision == 32
    tdt = torch.float32 if fp32_state else torch.float64
    k = hydro.constants(nb, C_cfl, "f32" if fp32_state else "f64")
    shape = tuple(n + 2 for n in nb)
    mom_c = 2 if dim == 2 else 3
    log(f"[bench] {name}: allocating {4 * np.prod(shape) * (2 + mom_c) / 2 * (4 if fp32_state else 8) / 1e9:.1f} GB")
    a = [torch.empty(shape + (2,), dtype=tdt, device="cuda"),
         torch.empty(shape + (mom_c,), dtype=tdt, device="cuda")]
    b = [torch.empty_like(a[0]), torch.empty_like(a[1])]
    for t in a + b:
        t.zero_()
    dt0 = hydro.ic(a[0], a[1], k["C_min"], ic_mode=ic_mode, radius_frac=rfrac, p_in=p_in, p_out=p_out)
    mass0 = hydro.total_mass(a[0])
    log(f"[bench] IC done: dt0={dt0:.3e} mass0={mass0:.7e}")

    W, K = args.warmup, args.steps
    E_MAX = 4000  # cap on the energy window's steps (see below)
    run = hydro.Run(nb, k["D"], k["C_min"], dt0, 2.0, W + K + E_MAX, precision)
    run.advance(a, b, W)
    st = run.status()
    if st["rc"] != 0:
        log(f"[bench] warmup failed: {st}")
        return 1
    log(f"[bench] warmup {W} steps done, t={st['t']:.4e}")

    run.set_timing(True)
    meter = EnergyMeter()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record()
    run.advance(a, b, K)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    gpu_s = ev0.elapsed_time(ev1) / 1e3
    kernel_ms, launches = run.kernel_time()
    st = run.status()
    if st["rc"] != 0 or st["it"] != W + K:
        log(f"[bench] timed run failed: {st}")
        return 1
    cur = a if st["current"] == 0 else b
    mass1 = hydro.total_mass(cur[0])
    cells = int(np.prod(nb))
    elapsed = max(wall, gpu_s)

    # Energy: average power between updates of the accumulator (about one a second)
    # while the same resident steps keep the device busy (about 5.5 s of them enqueued),
    # times the timed step time.
    run.set_timing(False)
    bp = None
    if args.no_energy:
        meter.error = "not measured (--no-energy)"
    else:
        E = min(E_MAX, max(K, int(np.ceil(5.5 / (elapsed / K)))))
        torch.cuda.synchronize()

        def busy():
            run.advance(a, b, E)
            torch.cuda.synchronize()

        bp = meter.busy_power(busy)
        st_e = run.status()
        if st_e["rc"] != 0 or st_e["it"] != W + K + E:
            log(f"[bench] energy window failed: {st_e}")
            return 1
        cur = a if st_e["current"] == 0 else b
    gcups = cells * K / elapsed / 1e9
    avg_kernel_s = kernel_ms / 1e3 / max(launches, 1)
    achieved = bpc * cells / avg_kernel_s / 1e9
    log(f"[bench] {K} steps: wall {wall * 1e3:.1f} ms, events {gpu_s * 1e3:.1f} ms, kernel avg "
        f"{avg_kernel_s * 1e3:.3f} ms x {launches}, {gcups:.2f} GCUPS, {achieved:.0f} GB/s")

    # Achievable bandwidth: the library's stream copy over the same number of bytes one
    # step reads (and writes), on buffers of the state's size.  Reuses the state's
    # second buffer pair as the destination after the run.
    copy_gbs = copy_cached_gbs = copy_nt_gbs = read_gbs = write_gbs = pattern_gbs = None
    if not args.no_copy_reference:
        src, dst = cur[0], (b if cur is a else a)[0]
        hydro.stream_copy(dst, src)
        torch.cuda.synchronize()
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record()
        for _ in range(3):
            hydro.stream_copy(dst, src)
        c1.record()
        torch.cuda.synchronize()
        nbytes = src.numel() * src.element_size()
        copy_cached_gbs = 2 * nbytes / (c0.elapsed_time(c1) / 3 / 1e3) / 1e9
        # The same copy with non-temporal loads and stores, the faster of the two on
        # MI355X (6.5 vs 6.4 TB/s): the achievable ceiling quoted for a step that reads
        # and writes once.
        hydro.stream_copy(dst, src, nontemporal=True)
        torch.cuda.synchronize()
        c0.record()
        for _ in range(3):
            hydro.stream_copy(dst, src, nontemporal=True)
        c1.record()
        torch.cuda.synchronize()
        copy_nt_gbs = 2 * nbytes / (c0.elapsed_time(c1) / 3 / 1e3) / 1e9
        copy_gbs = max(copy_cached_gbs, copy_nt_gbs)
        # Each direction alone over the same buffers: a read-only stream (non-temporal
        # loads) and a write-only fill.
        slots = torch.zeros(256, dtype=torch.int64, device="cuda")
        hydro.stream_read(src, slots)
        torch.cuda.synchronize()
        c0.record()
        for _ in range(3):
            hydro.stream_read(src, slots)
        c1.record()
        torch.cuda.synchronize()
        read_gbs = nbytes / (c0.elapsed_time(c1) / 3 / 1e3) / 1e9
        c0.record()
        for _ in range(3):
            hydro.stream_fill(dst, 0x3FF0000000000000)
        c1.record()
        torch.cuda.synchronize()
        write_gbs = nbytes / (c0.elapsed_time(c1) / 3 / 1e3) / 1e9
        log(f"[bench] stream {nbytes / 1e9:.1f} GB: copy {copy_cached_gbs:.0f} GB/s, non-temporal copy "
            f"{copy_nt_gbs:.0f} GB/s, read {read_gbs:.0f} GB/s, write {write_gbs:.0f} GB/s")
        if name == "blast2d_40000x20000_fp64":
            # The step's own memory schedule without its arithmetic, on the state's buffers:
            # the ceiling of its access pattern (the same loads and stores, 64 B per cell).
            other = b if cur is a else a
            hydro.stream_pattern2d(cur, other)
            torch.cuda.synchronize()
            c0.record()
            for _ in range(3):
                hydro.stream_pattern2d(cur, other)
            c1.record()
            torch.cuda.synchronize()
            pattern_gbs = bpc * cells / (c0.elapsed_time(c1) / 3 / 1e3) / 1e9
            log(f"[bench] the step's access pattern without arithmetic: {pattern_gbs:.0f} GB/s")

    if args.calibrate_copy:
        # A plain copy of exactly 1 GiB (read 1 GiB, write 1 GiB): what the FETCH_SIZE /
        # WRITE_SIZE counters read for it scales the step kernel's counts.
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            hydro.stream_copy(dst, src)
        torch.cuda.synchronize()
        del src, dst

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU oracle on this host ...")
        cpu = cpu_baseline(dim)
        log(f"[bench] cpu: {cpu}")

    line = {
        "metric": "cell-updates/s (GCUPS) + % HBM roofline",
        "value": round(gcups, 4),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(elapsed / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(gcups / PUBLISHED_GCUPS[name], 2) if name in PUBLISHED_GCUPS else None,
        "vs_baseline_basis": ("published 1.018 GCUPS, 2x IA-840F FPGA, res_2000_40000_20000_2FPGAS.res:10108"
                              if name in PUBLISHED_GCUPS else None),
        "dtype": {64: "f64", 32: "f32"}.get(precision, "f64 state, f32 flux"),
        "data": "synthetic (built-in blast IC generated on the device)",
        "config": {"workload": name, "grid": list(nb), "cells": cells, "boundaries": "periodic X, reflective Y/Z",
                   "steps_timed": K, "kernel": "fvm2d" if dim == 2 else "fvm3d",
                   "rows_per_strip": hydro.tune_get("rows2d_last" if dim == 2 else "rows3d_last"),
                   "workgroups_per_cu": hydro.tune_get("occ2d_last") if dim == 2 else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(name)},
        "cpu_baseline": cpu,
        "hbm_frac_whole_step": round(bpc * cells * K / elapsed / 1e9 / HBM_PEAK_GBS, 4),
        "achievable_copy_gbs": round(copy_gbs, 1) if copy_gbs else None,
        "achievable_copy_basis": "the faster of the library's stream copies (one 4 KiB chunk per 256-thread "
                                 "block, XCD-contiguous) with cached and with non-temporal loads and stores, "
                                 "over the state's buffer",
        "achievable_copy_cached_gbs": round(copy_cached_gbs, 1) if copy_cached_gbs else None,
        "achievable_copy_nt_gbs": round(copy_nt_gbs, 1) if copy_nt_gbs else None,
        "achievable_read_gbs": round(read_gbs, 1) if read_gbs else None,
        "achievable_write_gbs": round(write_gbs, 1) if write_gbs else None,
        "frac_of_achievable": round(achieved / copy_gbs, 4) if copy_gbs else None,
        "achievable_pattern_gbs": round(pattern_gbs, 1) if pattern_gbs else None,
        "achievable_pattern_basis": "the step's loads and stores without its arithmetic (hydro_stream_pattern2d: same "
                                    "workgroups, strips, block order, halo rows and rows in flight), 64 B per cell"
                                    if pattern_gbs else None,
        "frac_of_pattern": round(achieved / pattern_gbs, 4) if pattern_gbs else None,
        "kernel_ms": round(avg_kernel_s * 1e3, 4),
        "energy": ({"avg_watts": round(bp[0], 1), "joules_per_step": round(bp[0] * elapsed / K, 4),
                    "nj_per_cell_update": round(bp[0] * elapsed / K / cells * 1e9, 4), "window_s": round(bp[1], 3),
                    "source": "amdsmi energy accumulator: average power between its own updates while the same "
                              "resident steps run, times the timed step time"}
                   if bp is not None else {"error": meter.error or "no energy reading"}),
        "mass_change_per_cell": abs(mass1 - mass0) / cells,
        "t_final": st["t"],
    }
    run.close()
    if not args.no_forms and name == "blast2d_40000x20000_fp64":
        # The batched form, the 512^3 FP64 / FP32-flux comparison and the served form,
        # summarised in the same line (each has its own workload for a full line:
        # --workload batched / c3 / served).
        a = b = cur = src = dst = None  # noqa: F841  (free the 51 GB state first)
        torch.cuda.empty_cache()
        bt = bench_batched(args, "batched2d_256_fp32flux")
        torch.cuda.empty_cache()
        c3 = bench_c3(args)
        torch.cuda.empty_cache()
        sv = served_in_child(args)
        line["forms"] = {
            "batched2d_256_fp32flux": {k: bt[k] for k in (
                "value", "unit", "ms_per_step", "plain_resident_gcups", "plain_sync_launcher_gcups",
                "speedup_vs_plain_resident", "speedup_vs_plain_sync_launcher", "plain_sample", "error_vs_fp64")}
            | {"tiling": bt["config"]["tiling"], "roofline_achieved_gbs": bt["roofline"]["achieved"]},
            "blast3d_512_c3": {k: c3[k] for k in ("value", "unit", "ms_per_step", "fp64", "fp32flux", "error_vs_fp64")}
            | {"roofline_achieved_gbs": c3["roofline"]["achieved"], "roofline_frac": c3["roofline"]["frac"],
               "note": "FP32-flux GCUPS as value; 200 steps of each form from one IC"},
            "served_mixed": {"requests": sv["config"]["requests"], "arrival_rate_per_s": sv["config"]["arrival_rate_per_s"],
                             "load": sv["config"]["load"], "served": sv["served"],
                             "sequential_plain": sv["sequential_plain"],
                             "sequential_plain_real": sv["sequential_plain_real"],
                             "sequential_sync_launcher": sv["sequential_sync_launcher"],
                             "by_dimension": sv["by_dimension"], "large_class": sv["large_class"],
                             "large_class_ms": sv["config"]["large_class_ms"],
                             "large_backlog": sv["config"]["large_backlog"], "saturation": sv["saturation"]},
        }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
