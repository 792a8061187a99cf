#!/usr/bin/env python3
"""Throughput of the MPDATA step on one MI355X.

Workload (default): the single large instance -- 512 x 512 x 256 cells, FP64,
infinite-gauge non-oscillatory MPDATA with two iterations (the reference's
DWARF3 scheme, advec.f90 legacyoptmode 4 / t_adv 1), open boundaries,
solid-body rotation about the (1,1,1) axis with maximum Courant number 0.8,
and a positive random Fourier-mode field (k <= 8).  Inputs are generated on
the GPU and resident in HBM before the timed region.

A step is one call of mpdata_step_oop (x_out <- MPDATA(x_in)); the timed loop
alternates two field buffers, which is how the library's own time loop runs.
The timed path is the fused kernel's fast build (--path fast: FMA contraction
and Newton-refined reciprocal division); before the line is printed it is
checked against the exactly rounded fused step over 50 steps to the 1e-12
normwise relative tolerance the path must meet, and the exact fused step is
checked bit for bit against the per-stage kernels and timed as well
(exact_fused).
Prints one JSON line:
  value      cell-updates/s (interior cells x steps / seconds)
  roofline   achieved = algorithmic bytes per step (56 B per cell: read x,
             u1, u2, u3, h, rhr, write x) / average step time, against the
             MI355X HBM peak of 8 TB/s; traffic = measured HBM bytes per step
             from the committed rocprofv3 counter summary, when present
  cpu_baseline  the reference's own CPU path (its Fortran advec_dwarf compiled
             into oracle/_ref, one thread) on a 10-step sample of the same
             workload, timed on this host; the C oracle where that build is
             absent
Sub-objects: unfused_stages (the per-stage kernels), ensemble (64 uneven
tracers in one persistent launch / HIP graph), multifield (eight fields on one
grid), density (the density step with the velocity predictor), configs (the
periodic-standard and deformational-flow configurations, fused kernels),
adjoint (forward+backward against one forward); the ensemble, configs and
adjoint objects carry their own cpu_baseline.  --density-only,
--configs-only, --multifield-only and --adjoint-only print that object
alone.
"""
import argparse
import faulthandler
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MPDATA cell-updates/s, % HBM roofline; adjoint fwd+bwd time vs plain fwd"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, nargs=3, default=(512, 512, 256), metavar=("NP", "MP", "LP"))
    ap.add_argument("--dtype", choices=("f64", "f32"), default="f64")
    ap.add_argument("--path", choices=("fast", "auto", "fused", "stages"), default="fast")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="steps of the CPU baseline sample (default 10 through the "
                         "reference build, 2 through the oracle; 0 disables it)")
    ap.add_argument("--no-check", action="store_true", help="skip the fused-vs-stages check")
    ap.add_argument("--no-unfused", action="store_true",
                    help="skip timing the per-stage (unfused) path beside the fused one")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched-ensemble measurement (64 tracers on uneven grids)")
    ap.add_argument("--ensemble-steps", type=int, default=100)
    ap.add_argument("--no-multifield", action="store_true",
                    help="skip the many-fields-on-one-grid (CE_advection) measurement")
    ap.add_argument("--no-density", action="store_true",
                    help="skip the density step with the velocity predictor")
    ap.add_argument("--density-only", action="store_true",
                    help="run only the density-step measurement (--steps calls) and print its "
                         "JSON object")
    ap.add_argument("--configs-only", action="store_true",
                    help="run only the periodic-standard and deformational-flow configurations "
                         "(--steps steps each) and print their JSON object")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the periodic-standard and deformational-flow configurations")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint cost measurement and print its JSON object")
    ap.add_argument("--multifield-only", action="store_true",
                    help="run only the many-field measurement (--steps steps) and print its "
                         "JSON object")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint cost measurement (128x128x64, 50 steps)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of one field (counter calibration)")
    return ap.parse_args()


def copy_probe(a):
    """A plain copy of one scalar field of the workload, 5 times: the HBM
    counters of this kernel calibrate FETCH_SIZE / WRITE_SIZE."""
    import torch
    np_, mp_, lp_ = a.size
    n = (np_ + 2) * (mp_ + 2) * (lp_ + 2)
    tdt = torch.float64 if a.dtype == "f64" else torch.float32
    src = torch.rand(n, dtype=tdt, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": src.numel() * src.element_size(),
                      "copy_probe_bytes_written": dst.numel() * dst.element_size(),
                      "launches": 5}), flush=True)


def main():
    # a crash of the process (host segfault) prints the Python stack on stderr
    faulthandler.enable()
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {a.gpus} is not supported "
              "(the step would shard by 3-D domain decomposition, one block per GPU)", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        copy_probe(a)
        return
    if a.density_only:
        print(json.dumps(density_bench(a.steps)), flush=True)
        return
    if a.configs_only:
        print(json.dumps(configs_bench(a.steps)), flush=True)
        return
    if a.adjoint_only:
        print(json.dumps(adjoint_bench()), flush=True)
        return
    if a.multifield_only:
        print(json.dumps(multifield_bench(a.steps)), flush=True)
        return
    import torch

    from advectionhpctester_amd import mpdata, testdata

    tdt = torch.float64 if a.dtype == "f64" else torch.float32
    np_, mp_, lp_ = a.size
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    case = testdata.tilted_rotation_case(np_, mp_, lp_, cmax=0.8, dtype=tdt, device=dev)
    dom = mpdata.Domain(np_, mp_, lp_, dtype=tdt, device=dev)
    paths = {"fast": mpdata.PATH_FAST, "auto": mpdata.PATH_AUTO, "fused": mpdata.PATH_FUSED,
             "stages": mpdata.PATH_STAGES}
    dom.set_path(paths[a.path])
    bufs = [case["x"], torch.empty_like(case["x"])]
    bufs[1].copy_(bufs[0])
    args = (case["u1"], case["u2"], case["u3"], case["rhr"], case["h"], case["bcx"], case["bcy"])

    def run(n, start):
        for s in range(start, start + n):
            dom.step_oop(bufs[s & 1], bufs[(s + 1) & 1], *args)
        return start + n

    def timed(s):
        """Warm up, then time a.steps steps with HIP events on the stream the
        kernels are launched on (torch's current stream)."""
        s = run(a.warmup, s)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record()
        s = run(a.steps, s)
        ev1.record()
        torch.cuda.synchronize()
        return s, ev0.elapsed_time(ev1), time.perf_counter() - t0

    s, gpu_ms, wall = timed(0)
    cells = np_ * mp_ * lp_
    ms_step = gpu_ms / max(a.steps, 1)
    value = cells * a.steps / (gpu_ms / 1e3)
    rb = 8 if a.dtype == "f64" else 4
    alg_bytes = 7 * rb * cells
    achieved = alg_bytes / (ms_step / 1e3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        key = f"{np_}x{mp_}x{lp_}_{a.dtype}"
        if key in t:
            traffic = t[key]["bytes_per_step"]

    # the exactly rounded fused step, timed the same way
    exact = None
    if a.path == "fast":
        dom.set_path(mpdata.PATH_FUSED)
        s, ms_ex, _ = timed(s)
        exact = {"ms_per_step": ms_ex / max(a.steps, 1),
                 "value": cells * a.steps / (ms_ex / 1e3),
                 "frac": alg_bytes / (ms_ex / max(a.steps, 1) / 1e3) / 1e9 / HBM_PEAK_GBS}

    # correctness of what was timed: the exact fused and per-stage paths agree
    # bit for bit on one step; the fast path stays within 1e-12 (normwise
    # relative, interior) of the exact one over 50 steps
    checked = fast_err = None
    if not a.no_check:
        x0 = bufs[s & 1]
        xf = torch.empty_like(x0)
        dom.set_path(mpdata.PATH_FUSED)
        dom.step_oop(x0, xf, *args)
        xs = x0.clone()
        dom.set_path(mpdata.PATH_STAGES)
        dom.step(xs, *args)
        torch.cuda.synchronize()
        checked = bool(torch.equal(xf, xs))
        del xf, xs
        if a.path == "fast":
            res = []
            for p in (mpdata.PATH_FAST, mpdata.PATH_FUSED):
                dom.set_path(p)
                xr = x0.clone()
                dom.timeloop(50, xr, *args)
                res.append(mpdata.interior(xr))
            torch.cuda.synchronize()
            fast_err = float((res[0] - res[1]).abs().max() / res[1].abs().max())
            del res, xr
            if not fast_err <= 1e-12:
                print(f"bench.py: fast path differs from the exact step by {fast_err:.3e} "
                      "(> 1e-12) after 50 steps", file=sys.stderr)
                sys.exit(1)

    unfused = None
    if not a.no_unfused and a.path != "stages":
        dom.set_path(mpdata.PATH_STAGES)
        s, ms_unf, _ = timed(s)
        unfused = {"ms_per_step": ms_unf / max(a.steps, 1),
                   "value": cells * a.steps / (ms_unf / 1e3),
                   "hbm_alg_GBs": alg_bytes / (ms_unf / max(a.steps, 1) / 1e3) / 1e9}

    cpu = None
    if a.cpu_steps is None or a.cpu_steps > 0:
        cpu = cpu_baseline(case, (np_, mp_, lp_, 1), a.dtype, a.cpu_steps)

    ens = None
    if not a.no_ensemble and a.dtype == "f64":
        del bufs, args, case
        dom.close()
        torch.cuda.empty_cache()
        ens = ensemble_bench(a.ensemble_steps, cpu=a.cpu_steps is None or a.cpu_steps > 0)

    multi = None
    if not a.no_multifield and a.dtype == "f64":
        torch.cuda.empty_cache()
        multi = multifield_bench(20)

    dens = None
    if not a.no_density and a.dtype == "f64":
        torch.cuda.empty_cache()
        dens = density_bench(20)

    cfgs = None
    if not a.no_configs and a.dtype == "f64":
        torch.cuda.empty_cache()
        cfgs = configs_bench(20)

    adj = None
    if not a.no_adjoint and a.dtype == "f64":
        torch.cuda.empty_cache()
        adj = adjoint_bench(cpu=a.cpu_steps is None or a.cpu_steps > 0)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.dtype,
        "data": "synthetic",
        "config": {
            "workload": f"single large instance {np_}x{mp_}x{lp_}: gauge MPDATA, 2 iterations, "
                        "non-oscillatory, open walls, solid-body rotation about (1,1,1) at "
                        "Courant 0.8, Fourier-mode field k<=8",
            "grid": [np_, mp_, lp_], "variant": "gauge", "iters": 2, "nonos": 1,
            "boundaries": "open", "path": a.path, "parallelism": "single GPU",
            "arithmetic": ("FP64, FMA contraction and Newton-refined reciprocal division; "
                           "checked within 1e-12 of the exactly rounded step"
                           if a.path == "fast" else "FP64, exactly rounded (reference order)"),
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_step": alg_bytes,
                     "duration": "HIP events around the timed steps on the launch stream; "
                                 "a step is the fused kernel plus a halo-copy kernel"},
        "exact_fused": exact,
        "max_rel_err_fast_vs_exact_50_steps": fast_err,
        "cpu_baseline": cpu,
        "unfused_stages": unfused,
        "ensemble": ens,
        "multifield": multi,
        "density": dens,
        "configs": cfgs,
        "adjoint": adj,
        "checked_fused_equals_stages": checked,
        "wall_s_timed": wall,
    }
    print(json.dumps(out), flush=True)


def ensemble_bench(steps, nmember=64, seed=4, cpu=True):
    """64 tracers with per-tracer grids drawn from 32..256 (step 16) in each
    direction, Gaussian blobs, per-tracer solid-body rotation (max Courant
    0.3-0.8), FP64 gauge MPDATA, 2 iterations, limiter on: `steps` steps as
    one HIP graph of batched persistent launches (tiling autotuned first)
    against the same steps as 64 single-domain launches per step.  The fast
    build of the fused step is timed first (as the headline), then the
    exactly rounded one (exact_*); the two differ by at most 1e-12 (relative)
    after 10 steps from the same start, checked here.  cpu: one step of the
    first members on the CPU (ensemble_cpu_baseline)."""
    import torch

    from advectionhpctester_amd import mpdata, testdata
    dev = torch.device("cuda:0")
    shapes = testdata.ensemble_shapes(nmember, seed)
    members = [testdata.ensemble_member(*sh, seed=1000 + q, device=dev)
               for q, sh in enumerate(shapes)]
    cells = sum(p * q * r for p, q, r in shapes)
    xa = [m["x"] for m in members]
    x0 = [t.clone() for t in xa]
    xb = [torch.empty_like(t) for t in xa]
    rest = [[m[k] for m in members] for k in ("u1", "u2", "u3", "rhr", "h", "bcx", "bcy")]
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {}
    with mpdata.Ensemble(shapes) as ens:
        ens.set_fields(xa, xb, *rest)
        for name, path in (("fast", mpdata.PATH_FAST), ("exact", mpdata.PATH_FUSED)):
            ens.set_path(path)
            cfg, clen, _ = ens.autotune(trial_steps=2)
            ens.run(steps, graph=True)           # records the graph
            torch.cuda.synchronize()
            ev0.record()
            ens.run(steps, graph=True)
            ev1.record()
            torch.cuda.synchronize()
            out[name] = (ev0.elapsed_time(ev1), cfg, clen, ens.info()[0])
        ends = {}
        for name, path in (("fast", mpdata.PATH_FAST), ("exact", mpdata.PATH_FUSED)):
            ens.set_path(path)
            for a, b in zip(xa, x0):
                a.copy_(b)
            ens.run(10, graph=False)
            torch.cuda.synchronize()
            ends[name] = [mpdata.interior(t).clone() for t in xa]
        err = max(float((f - e).abs().max() / e.abs().max())
                  for f, e in zip(ends["fast"], ends["exact"]))
        del ends
    if not err <= 1e-12:
        raise SystemExit(f"ensemble: fast build differs by {err:.3e} from the exact build "
                         "after 10 steps (bound 1e-12)")
    doms = [mpdata.Domain(*sh) for sh in shapes]
    fields = [(m["u1"], m["u2"], m["u3"], m["rhr"], m["h"], m["bcx"], m["bcy"]) for m in members]

    def sequential(n):
        for s in range(n):
            for d, f, a, b in zip(doms, fields, xa, xb):
                if s & 1:
                    d.step_oop(b, a, *f)
                else:
                    d.step_oop(a, b, *f)

    seq = {}
    for name, path in (("fast", mpdata.PATH_FAST), ("exact", mpdata.PATH_AUTO)):
        for d in doms:
            d.set_path(path)
        sequential(2)
        torch.cuda.synchronize()
        ev0.record()
        sequential(steps)
        ev1.record()
        torch.cuda.synchronize()
        seq[name] = ev0.elapsed_time(ev1)
    for d in doms:
        d.close()
    cpu_entry = ensemble_cpu_baseline(x0, rest, shapes) if cpu else None
    alg = 56 * cells
    ms_graph, cfg, clen, nitem = out["fast"]
    ms_ex, cfg_ex, clen_ex, _ = out["exact"]
    return {"members": nmember, "cells": cells, "steps": steps, "path": "fast",
            "tiling": mpdata.Ensemble.configs()[cfg], "k_chunk": clen, "work_items": nitem,
            "graph_ms_per_step": ms_graph / steps,
            "graph_value": cells * steps / (ms_graph / 1e3),
            "graph_hbm_alg_GBs": alg / (ms_graph / steps / 1e3) / 1e9,
            "graph_frac": alg / (ms_graph / steps / 1e3) / 8e12,
            "sequential_ms_per_step": seq["fast"] / steps,
            "sequential_value": cells * steps / (seq["fast"] / 1e3),
            "speedup_graph_vs_sequential": seq["fast"] / ms_graph,
            "max_rel_err_fast_vs_exact_10_steps": err,
            "cpu_baseline": cpu_entry,
            "exact": {"tiling": mpdata.Ensemble.configs()[cfg_ex], "k_chunk": clen_ex,
                      "graph_ms_per_step": ms_ex / steps,
                      "graph_value": cells * steps / (ms_ex / 1e3),
                      "graph_frac": alg / (ms_ex / steps / 1e3) / 8e12,
                      "sequential_ms_per_step": seq["exact"] / steps,
                      "speedup_graph_vs_sequential": seq["exact"] / ms_ex}}


def ensemble_cpu_baseline(x0, rest, shapes, max_cells=25_000_000):
    """One step of the first members (up to max_cells cells) from their
    initial fields with the C oracle (oracle/liboracle_mpdata.so, one
    thread), a member at a time.  Not the reference build: its y-wall
    zeroing (defines.inc:208-216) writes row np+1 of the y-face array, past
    the array when np > mp+1, which most members are, so it is run on square
    planes only."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle

    orc = Oracle(np.float64)
    cells, dt, n = 0, 0.0, 0
    for q, (np_, mp_, lp_) in enumerate(shapes):
        if cells >= max_cells:
            break
        x = np.ascontiguousarray(x0[q].cpu().numpy())
        u1, u2, u3, rhr, h, bcx, bcy = (np.ascontiguousarray(r[q].cpu().numpy()) for r in rest)
        t0 = time.perf_counter()
        orc.step(1, 2, 1, x, u1, u2, u3, rhr, h, bcx, bcy, (0, 0, 0), (np_, mp_, lp_, 1))
        dt += time.perf_counter() - t0
        cells += np_ * mp_ * lp_
        n += 1
    return {"value": cells / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"one step of each of the first {n} of the {len(shapes)} members ({cells} "
                      f"cells) with oracle/liboracle_mpdata.so (plain C restatement, -O2, one "
                      f"thread), {dt:.1f} s"}


def multifield_bench(steps, size=(256, 256, 128), nfields=8):
    """Eight fields on one 256x256x128 grid with one velocity (the
    reference's CE_advection with th, u, v, w, tht, pstr, qv, qc): Fourier
    fields of different seeds, the headline's tilted rotation at Courant 0.8.
    All fields gauge (mpdata_alg_choice 1), out of place with two buffer
    sets: one ce_advection call per step against eight single-field steps;
    then the reference's default mix (alg_choice 0: five gauge, three
    standard fields), in place; exactly rounded builds, then the same three
    on the fast builds ("fast")."""
    import torch

    from advectionhpctester_amd import mpdata, testdata
    dev = torch.device("cuda:0")
    np_, mp_, lp_ = size
    case = testdata.tilted_rotation_case(np_, mp_, lp_, cmax=0.8, device=dev)
    vel = (case["u1"], case["u2"], case["u3"], case["rhr"], case["h"])
    roles = [0, 1, 1, 1, 1, 1, 0, 0]
    xa = [testdata.fourier_field(np_, mp_, lp_, seed=40 + f, device=dev)
          for f in range(nfields)]
    xb = [torch.empty_like(t) for t in xa]
    bx = [case["bcx"]] * nfields
    by = [case["bcy"]] * nfields
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    cells = np_ * mp_ * lp_ * nfields

    def timed(fn):
        fn(2)
        torch.cuda.synchronize()
        ev0.record()
        fn(steps)
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / steps

    with mpdata.Domain(np_, mp_, lp_) as d:
        def together(n):
            for s in range(n):
                src, dst = (xa, xb) if s % 2 == 0 else (xb, xa)
                d.ce_advection(src, bx, by, roles, *vel, alg_choice=1, out=dst)

        def one_by_one(n):
            for s in range(n):
                for f in range(nfields):
                    src, dst = (xa[f], xb[f]) if s % 2 == 0 else (xb[f], xa[f])
                    d.step_oop(src, dst, *vel, bx[f], by[f])

        def mixed(n):
            for _ in range(n):
                d.ce_advection(xa, bx, by, roles, *vel, alg_choice=0)

        ms_together = timed(together)
        ms_single = timed(one_by_one)
        ms_mixed = timed(mixed)
        # the same on the fast builds of both variants
        d.set_path(mpdata.PATH_FAST)
        fast = {"gauge_together_ms_per_step": timed(together),
                "gauge_one_by_one_ms_per_step": timed(one_by_one),
                "ce_mix_in_place_ms_per_step": timed(mixed)}
        fast["gauge_together_value"] = cells / (fast["gauge_together_ms_per_step"] / 1e3)
        fast["speedup_together_vs_one_by_one"] = (fast["gauge_one_by_one_ms_per_step"] /
                                                   fast["gauge_together_ms_per_step"])
    return {"grid": list(size), "fields": nfields, "steps": steps,
            "gauge_together_ms_per_step": ms_together,
            "gauge_together_value": cells / (ms_together / 1e3),
            "gauge_one_by_one_ms_per_step": ms_single,
            "speedup_together_vs_one_by_one": ms_single / ms_together,
            "ce_mix_in_place_ms_per_step": ms_mixed,
            "ce_mix": "alg_choice 0: th, qv, qc standard (fused standard kernel, one at a time), "
                      "u, v, w, tht, pstr gauge (one batched launch)",
            "fast": fast}


# Arrays one call of the density step with a fixed time step must move per
# cell (FP64 words): reads ox,oy,oz at levels n and n-1, h, u,v,w at levels n
# and n-1, x_in, x_incomp, pexe (16); writes ox,oy,oz at n+1/2 and n-1, u,v,w
# at n+1/2 and n-1, the mass fluxes u1,u2,u3, x_in, x_incomp, rhoadv, rhr,
# pexr1, pexr2 (21).
DENSITY_ARRAYS = 37


def configs_bench(steps):
    """The periodic-standard and deformational-flow configurations.
      periodic: 64x64x32, periodic in all directions, standard scheme, 2
        iterations, limiter off, solid-body rotation about z at Courant 0.5,
        Gaussian sphere (testdata.periodic_rotation_case); the reference's
        time loop entry (mpdata_c_timeloop) over 100 steps on the fast build
        of the fused standard kernel (path "fast"), then on the exactly
        rounded one (exact_ms_per_step), the two fields after the 102 steps
        within 1e-12 of each other (checked); stages_ms_per_step the same
        steps on the per-stage kernels; 56 B/cell.
      deformational: 256x256x128, time-varying deformational flow at Courant
        0.7 with the velocity predictor every step, gauge, 3 iterations,
        limiter on, slotted cylinders, FP64 and FP32.  A step is the one-pass
        predictor to n+1/2 (velprd_c, the time levels rotated by swapping the
        level arrays) and the three-iteration fused step; 80 B/cell in FP64
        (read x, h, rhr and the A-grid velocity at n and n-1, write x), 40 B in
        FP32.  ms_per_step excludes the harness's analytic velocity update
        (level n = U*cos(pi t/T), three elementwise products), which
        ms_per_step_with_velocity_update includes; stages_ms_per_step is the
        same step on the per-stage kernels with stored levels and the
        reference's level rotation.  FP64 times the fast build of the fused
        step (FMA contraction, Newton-refined reciprocals; path "fast"), as
        the headline does, with the exactly rounded build's time beside it
        (exact_ms_per_step) and the relative difference of the two after 20
        steps from the same start, which must stay within 1e-12; FP32 times
        the exactly rounded build.
    Each entry carries roofline (algorithmic bytes / step time against
    8 TB/s) and cpu_baseline: the oracle (plain C restatement, one thread;
    the reference cannot run these options) on a bounded sample of the same
    workload.  Timed with events on torch's current stream after 2 warm-up
    steps."""
    import numpy as np
    import torch

    from advectionhpctester_amd import mpdata, testdata
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    dev = torch.device("cuda:0")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, n):
        fn(2)
        torch.cuda.synchronize()
        ev0.record()
        fn(n)
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / n

    def roof(cells, ms, bpc):
        gbs = cells * bpc / (ms / 1e3) / 1e9
        return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs / HBM_PEAK_GBS, "bytes_per_cell": bpc}

    out = {}
    dims = (64, 64, 32)
    cells1 = dims[0] * dims[1] * dims[2]
    c = testdata.periodic_rotation_case(*dims, seed=11)
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in c.items()}
    with mpdata.Domain(*dims, ibc=(1, 1, 1)) as d:
        def periodic(n):
            d.timeloop(n, t["x"], t["u1"], t["u2"], t["u3"], t["rhr"], t["h"],
                       variant=mpdata.STANDARD, iters=2, nonos=0)
        # the fast build timed as the headline is, then the exactly rounded
        # one from the same start; their fields after the 102 steps compared
        x_start = t["x"].clone()
        c1_ms, c1_end = {}, {}
        for name, path in (("fast", mpdata.PATH_FAST), ("exact", mpdata.PATH_FUSED)):
            d.set_path(path)
            t["x"].copy_(x_start)
            c1_ms[name] = timed(periodic, 100)
            c1_end[name] = t["x"].clone()
        c1_err = float((c1_end["fast"] - c1_end["exact"]).abs().max() /
                       c1_end["exact"].abs().max())
        if not c1_err <= 1e-12:
            raise SystemExit(f"periodic standard: fast build differs by {c1_err:.3e} from the "
                             "exact build after 102 steps (bound 1e-12)")
        ms = c1_ms["fast"]
        d.set_path(mpdata.PATH_STAGES)
        ms_stages = timed(periodic, 100)
    orc = Oracle(np.float64)
    x = c["x"].copy()
    n_cpu = 100
    t0 = time.perf_counter()
    for _ in range(n_cpu):
        orc.step(2, 2, 0, x, c["u1"], c["u2"], c["u3"], c["rhr"], c["h"], c["bcx"], c["bcy"],
                 (1, 1, 1), dims + (1,))
    dt_cpu = time.perf_counter() - t0
    out["periodic_64x64x32_standard"] = {
        "ms_per_step": ms, "value": cells1 / (ms / 1e3), "path": "fast",
        "exact_ms_per_step": c1_ms["exact"], "max_rel_err_fast_vs_exact_102_steps": c1_err,
        "stages_ms_per_step": ms_stages, "roofline": roof(cells1, ms, 56),
        "cpu_baseline": {"value": cells1 * n_cpu / dt_cpu, "unit": "cell-updates/s", "cores": 1,
                         "kind": "port",
                         "sample": f"{n_cpu} steps of the same 64x64x32 workload with "
                                   f"oracle/liboracle_mpdata.so (one thread), {dt_cpu:.1f} s"}}

    np_, mp_, lp_ = 256, 256, 128
    cells2 = np_ * mp_ * lp_
    period = 200
    vel = testdata.deformational_velocity(np_, mp_, lp_, cmax=0.7)
    x0 = testdata.slotted_cylinders(np_, mp_, lp_)
    for tdt, name, bpc in ((torch.float64, "f64", 80), (torch.float32, "f32", 40)):
        U = [torch.from_numpy(np.ascontiguousarray(a)).to(dev, tdt) for a in vel]
        xs = [torch.from_numpy(np.ascontiguousarray(x0)).to(dev, tdt)]
        xs.append(torch.empty_like(xs[0]))
        h = torch.ones_like(xs[0])
        bcx = torch.zeros((2, lp_, mp_), dtype=tdt, device=dev)
        bcy = torch.zeros((2, lp_, np_), dtype=tdt, device=dev)
        lev = [[a.clone() for a in U], [a.clone() for a in U]]  # level n, level n-1
        with mpdata.Domain(np_, mp_, lp_, dtype=tdt) as d:
            d.set_path(mpdata.PATH_FUSED)
            u1, u2, u3 = (d.zeros(k) for k in (1, 2, 3))
            state = {"n": 0}

            def deform(n, update=True):
                for _ in range(n):
                    k = state["n"]
                    cur, prev = lev[k & 1], lev[(k + 1) & 1]
                    if update:
                        f = testdata.reversal_factor(k, period)
                        for o, a in zip(cur, U):
                            torch.mul(a, f, out=o)
                    d.velprd_c(*cur, *prev, h, 1.0, 1.0, 1.0, 1.0, 1.0, u1, u2, u3)
                    d.step_oop(xs[k & 1], xs[(k + 1) & 1], u1, u2, u3, h, h, bcx, bcy,
                               variant=mpdata.GAUGE, iters=3, nonos=1)
                    state["n"] += 1
            ms_upd = timed(deform, steps)
            ms = timed(lambda n: deform(n, update=False), steps)
            fast = None
            if name == "f64":
                # the fast build: timed the same way, and its result after 20
                # steps from the start against the exactly rounded build's
                def run_from(path, n):
                    d.set_path(path)
                    state["n"] = 0
                    xs[0].copy_(torch.from_numpy(np.ascontiguousarray(x0)).to(dev, tdt))
                    for lv in lev:
                        for o, a in zip(lv, U):
                            o.copy_(a)
                    deform(n)
                    torch.cuda.synchronize()
                    return mpdata.interior(xs[state["n"] & 1]).clone()
                a_ = run_from(mpdata.PATH_FUSED, 20)
                b_ = run_from(mpdata.PATH_FAST, 20)
                err = float((a_ - b_).abs().max() / a_.abs().max())
                del a_, b_
                d.set_path(mpdata.PATH_FAST)
                fast = {"ms_upd": timed(deform, steps),
                        "ms": timed(lambda n: deform(n, update=False), steps), "err": err}
                if not err <= 1e-12:
                    raise SystemExit(f"deformational flow: fast build differs by {err:.3e} "
                                     "from the exact build after 20 steps (bound 1e-12)")

            # the per-stage path with stored levels and the reference's rotation
            d.set_path(mpdata.PATH_STAGES)
            L = [torch.zeros((3,) + tuple(h.shape), dtype=tdt, device=dev) for _ in range(6)]
            ox, oy, oz, uu, vv, ww = L
            for o, a in zip((ox, oy, oz), U):
                o[0] = a
                o[2] = a * h
            xst = xs[0].clone()

            def stages(n):
                for _ in range(n):
                    d.velprd_traj0(ox, oy, oz, u1, u2, u3, h, 1.0, 1.0, 1.0, 1.0, 1.0)
                    d.step(xst, u1, u2, u3, h, h, bcx, bcy, variant=mpdata.GAUGE, iters=3, nonos=1)
                    d.velprd_rotate(uu, vv, ww, ox, oy, oz, h, 1.0)
            ms_st = timed(stages, max(steps // 2, 2))
            del L, ox, oy, oz, uu, vv, ww, xst
        if fast is not None:
            entry = {"ms_per_step": fast["ms"], "value": cells2 / (fast["ms"] / 1e3),
                     "path": "fast", "ms_per_step_with_velocity_update": fast["ms_upd"],
                     "exact_ms_per_step": ms, "exact_ms_per_step_with_velocity_update": ms_upd,
                     "max_rel_err_fast_vs_exact_20_steps": fast["err"],
                     "stages_ms_per_step": ms_st, "roofline": roof(cells2, fast["ms"], bpc)}
        else:
            entry = {"ms_per_step": ms, "value": cells2 / (ms / 1e3), "path": "fused",
                     "ms_per_step_with_velocity_update": ms_upd, "stages_ms_per_step": ms_st,
                     "roofline": roof(cells2, ms, bpc)}
        if name == "f64":
            npdt = np.float64
            orc = Oracle(npdt)
            host = {"x": np.ascontiguousarray(x0.astype(npdt))}
            S = host["x"].shape
            Lh = [np.zeros((3,) + S, npdt) for _ in range(6)]
            uh = [np.zeros(shape_c, npdt) for shape_c in
                  ((lp_ + 2, mp_ + 2, np_ + 3), (lp_ + 2, mp_ + 3, np_ + 2), (lp_ + 3, mp_ + 2, np_ + 2))]
            hh = np.ones(S, npdt)
            bx, by = np.zeros((2, lp_, mp_), npdt), np.zeros((2, lp_, np_), npdt)
            for o, a in zip(Lh[:3], vel):
                o[0] = a
                o[2] = a * hh
            n_cpu = 2
            t0 = time.perf_counter()
            for _ in range(n_cpu):
                orc.velprd_traj0(*Lh[:3], *uh, hh, 1.0, 1.0, 1.0, 1.0, 1.0, (0, 0, 0),
                                 (np_, mp_, lp_, 1))
                orc.step(1, 3, 1, host["x"], *uh, hh, hh, bx, by, (0, 0, 0), (np_, mp_, lp_, 1))
                orc.velprd_rotate(*Lh[3:], *Lh[:3], hh, 1.0, (np_, mp_, lp_, 1))
            dt_cpu = time.perf_counter() - t0
            entry["cpu_baseline"] = {
                "value": cells2 * n_cpu / dt_cpu, "unit": "cell-updates/s", "cores": 1,
                "kind": "port",
                "sample": f"{n_cpu} step(s) of the same 256x256x128 FP64 workload (predictor, "
                          f"3-iteration gauge step, level rotation) with oracle/liboracle_mpdata.so "
                          f"(one thread), {dt_cpu:.1f} s"}
        out[f"deformational_256x256x128_gauge3_{name}"] = entry
        del U, xs, h, lev
        torch.cuda.empty_cache()
    out["steps"] = steps
    return out


def density_bench(steps, size=(256, 256, 128)):
    """The density step with the velocity predictor (the reference's
    mpdata3d_rho_gauge_gpubc) on one 256x256x128 FP64 grid
    (testdata.density_workload: A-grid solid-body rotation about (1,1,1) at
    Courant 0.8, dx = dt = 1, steady levels; density x_incomp a positive
    Fourier field, h = 1 + 0.05 (Fourier field - 1), pexe = 1, zero
    inflow).  Fixed time step: one call = extrapolation to n+1/2, A->C
    averaging, level rotation and the density MPDATA step, all queued
    without a host read-back; the same on the fast build of the fused
    density kernel (fixed_fast).  Then with substepping at Courant 1.6 (two
    substeps; each call reads the Courant maxima back to the host)."""
    import torch

    from advectionhpctester_amd import mpdata, testdata
    dev = torch.device("cuda:0")
    np_, mp_, lp_ = size
    f = testdata.density_workload(np_, mp_, lp_, device=dev)
    cells = np_ * mp_ * lp_
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {"grid": list(size), "steps": steps, "alg_arrays_per_cell": DENSITY_ARRAYS}
    with mpdata.Domain(np_, mp_, lp_) as d:
        def call(state, **flags):
            d.rho_gauge_gpubc(state, f["ox"], f["oy"], f["oz"], f["u"], f["v"], f["w"], f["u1"],
                              f["u2"], f["u3"], f["x_in"], f["x_incomp"], f["h"], f["rhoadv"],
                              f["rhr"], f["pexe"], f["pexr1"], f["pexr2"], f["bcx"], f["bcy"],
                              1.0, 1.0, 1.0, 1.0, crmaxlim=0.9, dtmax=1.0, **flags)

        for mode, flags, scale, path in (("fixed", {}, 1.0, mpdata.PATH_AUTO),
                                         ("fixed_fast", {}, 1.0, mpdata.PATH_FAST),
                                         ("substep", {"lsubstepping": 1}, 2.0, mpdata.PATH_AUTO)):
            d.set_path(path)
            for name in ("ox", "oy", "oz"):
                f[name][0] *= scale
                f[name][2] *= scale
            st = {"nsubsteps": 1, "nsubsteps_o": 1, "dtn": 1.0, "dtnold": 1.0}
            for _ in range(3):
                call(st, **flags)
            torch.cuda.synchronize()
            ev0.record()
            for _ in range(steps):
                call(st, **flags)
            ev1.record()
            torch.cuda.synchronize()
            ms = ev0.elapsed_time(ev1) / steps
            gbs = DENSITY_ARRAYS * 8 * cells / (ms / 1e3) / 1e9
            out[mode] = {"ms_per_call": ms, "value": cells / (ms / 1e3),
                         "alg_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
                         "nsubsteps": st["nsubsteps"]}
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh).get(f"density_{np_}x{mp_}x{lp_}_f64")
        if t:
            out["traffic_bytes_per_call"] = t["bytes_per_call"]
    return out


def adjoint_bench(nsteps=50, ckpt=10, reps=3, cpu=True):
    """Gradient of J = 0.5 |x_T - obs|^2 after `nsteps` standard
    non-oscillatory steps at 128x128x64 under the reversing deformational
    flow, with respect to x0 and the face velocities, checkpoints every
    `ckpt` steps: time of forward + backward against one plain forward of
    the same steps (both include forming each step's velocity).  The same
    measurement over gauge steps goes under "gauge".  cpu: the same gradient
    by reverse-mode differentiation on the CPU (adjoint_cpu_baseline)."""
    out = _adjoint_cost("standard", nsteps, ckpt, reps)
    out["target_ratio"] = 4.0
    g = _adjoint_cost("gauge", nsteps, ckpt, reps)
    out["gauge"] = {k: g[k] for k in ("forward_ms", "forward_backward_ms", "ratio")}
    out["cpu_baseline"] = adjoint_cpu_baseline(nsteps) if cpu else None
    return out


def adjoint_cpu_baseline(nsteps, sample=2):
    """The reference has no adjoint; the CPU path for the gradient is
    torch.autograd over oracle/mpdata_torch.py (the oracle's standard step
    in vectorised torch operations, every intermediate kept, no checkpoints)
    on one thread: forward and backward of the first `sample` steps of the
    same problem, scaled to `nsteps` steps."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mpdata_torch

    from advectionhpctester_amd import testdata
    c = testdata.adjoint_case(128, 128, 64, nsteps=nsteps)
    f = c.pop("factors")
    obs = torch.from_numpy(c.pop("obs"))
    t = {k: torch.from_numpy(v) for k, v in c.items()}
    threads = torch.get_num_threads()
    torch.set_num_threads(1)
    try:
        leaves = [t[k].clone().requires_grad_(True) for k in ("x", "u1", "u2", "u3")]
        t0 = time.perf_counter()
        y = leaves[0]
        for n in range(sample):
            y = mpdata_torch.step(y, leaves[1] * f[n], leaves[2] * f[n], leaves[3] * f[n],
                                  t["rhr"], t["h"], t["bcx"], t["bcy"])
        r = (y - obs)[1:-1, 1:-1, 1:-1]
        (0.5 * (r * r).sum()).backward()
        dt = time.perf_counter() - t0
    finally:
        torch.set_num_threads(threads)
    return {"value": dt * 1e3 * nsteps / sample, "unit": f"ms per forward+backward of {nsteps} steps",
            "cores": 1, "kind": "port",
            "sample": f"{sample} steps forward + backward with torch.autograd on the CPU over "
                      f"oracle/mpdata_torch.py (the oracle's standard step in torch operations, "
                      f"one thread), {dt:.1f} s, scaled to {nsteps} steps"}


def _adjoint_cost(variant_name, nsteps, ckpt, reps):
    import torch

    from advectionhpctester_amd import adjoint, mpdata, testdata
    c = testdata.adjoint_case(128, 128, 64, nsteps=nsteps)
    f = c.pop("factors")
    dev = torch.device("cuda:0")
    obs = torch.from_numpy(c.pop("obs")).to(dev)
    d = {k: torch.from_numpy(v).to(dev) for k, v in c.items()}
    rest = (d["rhr"], d["h"], d["bcx"], d["bcy"])
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    variant = {"standard": mpdata.STANDARD, "gauge": mpdata.GAUGE}[variant_name]
    with mpdata.Domain(128, 128, 64) as dom:
        def plain():
            x = d["x"].clone()
            for n in range(nsteps):
                dom.step(x, d["u1"] * f[n], d["u2"] * f[n], d["u3"] * f[n], *rest,
                         variant=variant, iters=2, nonos=1)
            return x

        def fwd_bwd():
            leaves = [d[k].clone().requires_grad_(True) for k in ("x", "u1", "u2", "u3")]
            xT = adjoint.advect(*leaves, dom, *rest, nsteps, ckpt=ckpt, factors=f,
                                variant=variant)
            r = mpdata.interior(xT - obs)
            (0.5 * (r * r).sum()).backward()
            return leaves

        def timed(fn):
            fn()
            torch.cuda.synchronize()
            best = None
            for _ in range(reps):
                ev0.record()
                fn()
                ev1.record()
                torch.cuda.synchronize()
                t = ev0.elapsed_time(ev1)
                best = t if best is None else min(best, t)
            return best

        t_fwd = timed(plain)
        t_fb = timed(fwd_bwd)
    return {"grid": [128, 128, 64], "steps": nsteps, "checkpoint_every": ckpt,
            "variant": variant_name, "iters": 2, "nonos": 1,
            "forward_ms": t_fwd, "forward_backward_ms": t_fb, "ratio": t_fb / t_fwd}


def cpu_baseline(case, dims, dtype, steps):
    """The reference's own CPU path (its Fortran routines compiled at -O2 into
    oracle/_ref/libmpdata_ref.so, one thread: advec_dwarf, mode 4, gauge) on
    `steps` steps of the same workload; the oracle (plain C restatement) when
    that build is absent or the run is FP32.  steps=None: 10 steps of the
    reference or 2 of the oracle (about 10-20 s of CPU work at 512x512x256)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, Reference

    npdt = np.float64 if dtype == "f64" else np.float32
    host = {k: np.ascontiguousarray(v.cpu().numpy()) for k, v in case.items()}
    x = host["x"].copy()
    cells = dims[0] * dims[1] * dims[2]
    size = f"{dims[0]}x{dims[1]}x{dims[2]} {dtype}"
    if dtype == "f64" and Reference.available():
        n = steps or 10
        ref = Reference()
        ref.setup(dims[0], dims[1], dims[2], dims[3])
        t0 = time.perf_counter()
        for _ in range(n):
            ref.advec_dwarf(4, 1, x, host["bcx"], host["bcy"], host["u1"], host["u2"], host["u3"],
                            host["rhr"], host["h"])
        dt = time.perf_counter() - t0
        return {"value": cells * n / dt, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                "sample": f"{n} step(s) of the same {size} workload through the reference's "
                          f"advec_dwarf (its Fortran CPU path compiled -O2 into "
                          f"oracle/_ref/libmpdata_ref.so, one thread), {dt:.1f} s"}
    n = steps or 2
    orc = Oracle(npdt)
    t0 = time.perf_counter()
    for _ in range(n):
        orc.step(1, 2, 1, x, host["u1"], host["u2"], host["u3"], host["rhr"], host["h"],
                 host["bcx"], host["bcy"], (0, 0, 0), dims)
    dt = time.perf_counter() - t0
    return {"value": cells * n / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{n} step(s) of the same {size} workload with oracle/liboracle_mpdata.so "
                      f"(plain C restatement, -O2, one thread), {dt:.1f} s"}


if __name__ == "__main__":
    main()
