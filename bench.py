#!/usr/bin/env python
"""Benchmark of the FP64 residual + JVP path on one MI355X.

Workload (one "step"): the per-Newton-iteration work of the reference driver
on the 2-D cylinder O-mesh of 512 (azimuthal) x 384 (radial) cells
(M = 0.1, Re_D = 50, free stream + 1e-4 seeded noise): the boundary-condition
sequence, one residual, and the full colored Jacobian -- 490 one-hot seeds in
two i-zones, each equivalent to its linearised BCs and tangent residual
(cylinder.py:895-935) -- pruned into the CSR the reference hands to PETSc
(BROADCAST_npz.py:1018-1157).  The timed step is an iteration after the
first: the kept pattern of every row is checked against the previous
iteration's and the values are rewritten in place.  The reference's own
route (its COO layout, then the prune + CSR) and the first iteration's full
build are timed beside it.

value = Mcells/s = cells per step / step time.  Inputs are resident in HBM
before the timed region.  The dominant kernel's roofline figure comes from HIP
events recorded around each of its launches on the stream the library uses.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "Mcells/s residual+JVP FP64 (%HBM peak); Newton time-to-1e-10 learned vs plain"

# algorithmic bytes (per unit) used for the roofline figures
RESIDUAL_B_PER_CELL = 136          # w 40 + nx,ny 32 + vol 8 + volf 16 + residu 40
JVP_B_PER_SEED_CELL = 80           # wd in 40 + resd out 40
JVP_SHARED_B_PER_CELL = 96         # w 40 + metrics 56, once per launch


def kernel_ms(flow, name):
    from broadcast_amd import _lib

    tot, n = C.c_double(), C.c_long()
    _lib.check(_lib.lib().bca_profile_read(flow.ctx.h, name.encode(), C.byref(tot), C.byref(n)),
               "bca_profile_read")
    return tot.value, n.value


def _seed_worker(args, keep=False):
    """One process of the all-core CPU baseline: the oracle's seed passes
    (testvector[_partial] + tangent BCs + JVP + scatter) of `seeds` on the
    case, until `budget_s` of work; returns (passes done, seconds) and, with
    keep, the COO they filled."""
    import numpy as np
    import oracle
    from oracle import driver as D
    from broadcast_amd.cases import cylinder_case

    im, jm, seed, seeds, budget_s = args
    case = cylinder_case(im, jm, seed=seed)
    lib = oracle.Oracle()
    g = D.geometry(lib, case)
    w = np.array(case.w, order="F", copy=True)
    D.apply_bcs(lib, case, g, w)
    gh = case.gh
    split = D.two_zone_split(case)
    zones = [(0, im - 1)] if split is None else [(0, split - 1), (split, im - 1)]
    nb = im * jm * 49 * 25 * len(zones)
    jac, ia, ja = np.zeros(nb), np.zeros(nb, np.int32), np.zeros(nb, np.int32)
    cd = np.full((im, jm), 1e-2, order="F")
    wd = np.zeros_like(w, order="F")
    done, t = 0, 0.0
    for (n, m, l, k) in seeds:
        t1 = time.perf_counter()
        if split is None:
            lib.testvector(wd, m, l, k, gh, im, jm)
        else:
            lib.testvector_partial(wd, m, l, k, gh, im, jm, zones[n][0], zones[n][1], 0, jm - 1)
        _, resd = D.jvp(lib, case, g, w, wd)
        if split is None:
            lib.computejacobianfromjv(jac, ia, ja, resd, m, l, k, gh, im, jm, coefdiag=cd, withjn=case.kind == "cyl")
        else:
            lib.computejacobianfromjv_relaxed_withjnandcheck(jac, ia, ja, resd, m, l, k, gh, im, jm, cd, 2e-16, n)
        t += time.perf_counter() - t1
        done += 1
        if t > budget_s:
            break
    if keep:
        return done, t, (jac, ia, ja)
    return done, t


def cpu_baseline(case, budget_s=20.0, seed=1000):
    """The oracle (the C++ restatement of the reference, g++ -O2) on the host:
    one BC + residual pass, a bounded sample of the seed passes of the same
    coloring (two i-zones for an O-mesh whose im is not a multiple of 7), and
    the prune + scipy CSR of the entries they produced, scaled to a full
    step; on one core, and with the seed passes spread over every host core
    the box gives this process (at most 16, one process each).  The flang
    build of the reference Fortran computes the same numbers but runs its
    residual ~20x slower than the restatement, so it is not the timed
    baseline.  Runs before the GPU is touched (the worker processes are
    forked)."""
    import multiprocessing as mp
    import numpy as np
    import scipy.sparse as sp
    import oracle
    from oracle import driver as D

    lib = oracle.Oracle()
    g = D.geometry(lib, case)
    w = np.array(case.w, order="F", copy=True)
    t0 = time.perf_counter()
    D.apply_bcs(lib, case, g, w)
    D.residual(lib, case, g, w)
    t_res = time.perf_counter() - t0
    im, jm, gh = case.im, case.jm, case.gh
    split = D.two_zone_split(case)
    nz = 1 if split is None else 2
    seeds = [(n, m, l, k) for n in range(nz) for m in range(5) for l in range(7) for k in range(7)]
    # one core
    done1, t_seed1, (jac, ia, ja) = _seed_worker((im, jm, seed, seeds, budget_s * 0.5), keep=True)
    per1 = t_seed1 / done1
    # prune + CSR of the slots the sampled seeds filled, scaled by the slot fraction
    per = im * jm * 5
    filled = np.zeros(jac.size, bool)
    for (n, m, l, k) in seeds[:done1]:
        s0 = k * per + l * per * 7 + m * per * 49 + n * per * 245
        filled[s0:s0 + per] = True
    t2 = time.perf_counter()
    sub = np.flatnonzero(filled)
    keep = np.abs(jac[sub]) > 2e-16
    sp.csr_matrix((jac[sub][keep], (ia[sub][keep], ja[sub][keep])), shape=(per, per))
    t_csr = (time.perf_counter() - t2) * len(seeds) / done1
    del jac, ia, ja, filled, sub, keep
    step1 = t_res + per1 * len(seeds) + t_csr
    # all cores: each process a share of the seed list, the same work budget
    ncores = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
    shares = [seeds[q::ncores] for q in range(ncores)]
    ctx = mp.get_context("fork")
    t3 = time.perf_counter()
    with ctx.Pool(ncores) as pool:
        res = pool.map(_seed_worker, [(im, jm, seed, sh, budget_s * 0.5) for sh in shares])
    wall = time.perf_counter() - t3
    doneN = sum(d for d, _ in res)
    # aggregate seed-pass rate with every process busy (each ran until its budget)
    rateN = sum(d / t for d, t in res)
    stepN = t_res + len(seeds) / rateN + t_csr
    return {
        "value": im * jm / stepN / 1e6,
        "unit": "Mcells/s",
        "cores": ncores,
        "kind": "port",
        "sample": f"{im}x{jm} O-mesh, {len(seeds)}-seed step: 1 BC+residual ({t_res:.2f} s), {doneN} seed passes "
                  f"(testvector + tangent BCs + JVP + scatter) over {ncores} processes in {wall:.1f} s "
                  f"(aggregate {rateN:.2f} passes/s), prune + scipy CSR scaled from {done1} passes "
                  f"({t_csr:.2f} s); scaled to a full step of {stepN:.1f} s",
        "single_core": {"value": im * jm / step1 / 1e6, "unit": "Mcells/s", "cores": 1,
                        "sample": f"{done1} of {len(seeds)} seed passes, {per1:.2f} s each; full step {step1:.1f} s"},
        "host_cpu": _cpu_model(),
        "nproc": os.cpu_count(),
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--im", type=int, default=512)
    ap.add_argument("--jm", type=int, default=384)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the Newton (256x128 boundary layer) and J(beta) (512x192, 16 beta) figures")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 512 MiB device-to-device copy first (counter calibration under rocprofv3)")
    ap.add_argument("--figure", choices=("residual_jvp", "vjp"), default=None,
                    help="run one side figure alone and print its JSON (for counter runs under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures the single-GPU path only: every box here has one MI355X; "
              "run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from broadcast_amd import _lib
    from broadcast_amd.cases import cylinder_case
    from broadcast_amd.flow import Flow, to_device

    cpu = None
    if not args.no_cpu_baseline:
        # before the GPU is touched: the all-core leg forks worker processes
        print("bench: CPU baseline (oracle, 1 core and all cores)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(cylinder_case(args.im, args.jm, seed=1000 * 1 + 0), args.cpu_budget)
        print(f"bench: CPU baseline {cpu['value']:.4f} Mcells/s on {cpu['cores']} cores", file=sys.stderr, flush=True)
    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.figure is not None:
        fig = residual_jvp_figure() if args.figure == "residual_jvp" else vjp_figure()
        print(json.dumps({"figure": args.figure, **fig}), flush=True)
        return
    case = cylinder_case(args.im, args.jm, seed=1000 * 1 + 0)
    cells = case.im * case.jm
    f = Flow(case)
    w = to_device(case.w)
    res = f.new_state()
    bufs = f.jacobian_buffers()

    # relaxation term on the diagonal, as the Newton driver passes it
    # (BROADCAST_npz.py:1063-1066; its values do not change the work)
    cd = torch.full((case.im * case.jm,), 1e-2, dtype=torch.float64, device="cuda")

    nzones = 1 if f.two_zone_split is None else 2
    nseeds = 245 * nzones
    # launch shapes of the residual faces, face blocks and assembly for this mesh
    f.apply_bcs(w)
    tuned = f.autotune(w, jac=bufs[0])
    print(f"bench: autotuned {tuned}", file=sys.stderr, flush=True)

    def step():
        # the Newton loop's Jacobian: straight into the pruned CSR; the kept
        # pattern of every row is checked and, while it holds, only the
        # values are rewritten (bca_jacobian_colored_csr_refresh)
        f.apply_bcs(w)
        f.residual(w, res)
        return f.jacobian_csr(w, coefdiag=cd, refresh=True)

    def step_coo(first=False):
        # the reference's own route: the COO of its layout, then the prune + CSR
        f.apply_bcs(w)
        f.residual(w, res)
        f.jacobian_colored(w, coefdiag=cd, bufs=bufs, write_index=first)
        return f.csr(bufs[0])

    def timed(fn, n, names):
        fn()
        torch.cuda.synchronize()
        _lib.check(_lib.lib().bca_profile_enable(f.ctx.h, 1), "bca_profile_enable")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ks = {}
        for name in names:
            tot, cnt = kernel_ms(f, name)
            ks[name] = {"total_ms": tot, "launches": cnt, "avg_ms": tot / max(cnt, 1)}
        return e0.elapsed_time(e1) / n, ks

    print(f"bench: {case.name} cells={cells} warmup={args.warmup} steps={args.steps}", file=sys.stderr,
          flush=True)
    # the COO route (figure beside the headline), then release its buffers
    nnz_coo = int(step_coo(first=True)[1].numel())
    coo_ms, coo_kernels = timed(step_coo, max(args.steps, 1),
                                ("k_iface_blocks", "k_jface_blocks", "ghost_tangents", "k_assemble_colored",
                                 "k_assemble_ghosts", "k_csr_count", "k_csr_fill"))
    from broadcast_amd.graph import JacobianGraph

    G = JacobianGraph(f, w, coefdiag=cd, bufs=bufs)
    G.replay()
    f.csr(bufs[0])
    torch.cuda.synchronize()
    eg0, eg1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    eg0.record()
    for _ in range(max(args.steps, 1)):
        G.replay()
        f.csr(bufs[0])
    eg1.record()
    torch.cuda.synchronize()
    graph_ms = eg0.elapsed_time(eg1) / max(args.steps, 1)
    del G, bufs
    torch.cuda.empty_cache()

    # the full build without the COO (count, scan, fill), as a first Newton
    # iteration does it
    def step_build():
        f.apply_bcs(w)
        f.residual(w, res)
        return f.jacobian_csr(w, coefdiag=cd)

    build_ms, build_kernels = timed(step_build, max(args.steps // 2, 1),
                                    ("k_jac_csr_count", "k_jac_csr_fill", "k_ghost_part"))

    # residual alone (same kernels as inside the step), for its own figure
    res_ms, res_kernels = timed(lambda: (f.apply_bcs(w), f.residual(w, res)), 20, ("k_face<f64>",))
    face64 = res_kernels["k_face<f64>"]

    # timed region: the headline step
    names = ("k_iface_blocks", "k_jface_blocks", "ghost_tangents", "k_ghost_part", "k_jac_csr_refresh")
    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    _lib.check(_lib.lib().bca_profile_enable(f.ctx.h, 1), "bca_profile_enable")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        indptr, indices, data = step()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / args.steps
    assert not f.csr_rebuilt, "the kept pattern changed inside the timed region"
    nnz = int(data.numel())
    assert nnz == nnz_coo, (nnz, nnz_coo)
    kernels = {}
    for name in names:
        tot, n = kernel_ms(f, name)
        kernels[name] = {"total_ms": tot, "launches": n, "avg_ms": tot / max(n, 1)}
    _lib.check(_lib.lib().bca_profile_enable(f.ctx.h, 0), "bca_profile_enable")
    print(f"bench: {ms_step:.3f} ms/step (wall {wall / args.steps * 1e3:.3f}) kernels {kernels}",
          file=sys.stderr, flush=True)

    # Algorithmic bytes of the step for the colored Jacobian: 8 B per CSR
    # value written + 96 B per cell of shared inputs (w and metrics); the
    # face blocks are scratch and the CSR indices a per-mesh pattern.  The
    # roofline line gives these bytes to the dominant kernel.
    step_alg = 8 * nnz + JVP_SHARED_B_PER_CELL * cells
    nfi, nfj = (case.im + 1) * case.jm, case.im * (case.jm + 1)
    # bytes each kernel moves by design (face blocks are written once, read by
    # the two cells beside each face)
    blocks = 8 * 24 * 25 * (nfi + nfj)
    by_design = {
        "k_iface_blocks": JVP_SHARED_B_PER_CELL * cells + 8 * 24 * 25 * nfi,
        "k_jface_blocks": JVP_SHARED_B_PER_CELL * cells + 8 * 24 * 25 * nfj,
        "k_jac_csr_refresh": blocks + 8 * nnz + 64 * 5 * cells,
        "ghost_tangents": 0,
        "k_ghost_part": 0,
    }
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    avg_dom = kernels[dom]["avg_ms"]
    ach_dom = step_alg / (avg_dom * 1e-3) / 1e9
    ach_dom_design = by_design[dom] / (avg_dom * 1e-3) / 1e9
    ach_res = RESIDUAL_B_PER_CELL * cells / (res_ms * 1e-3) / 1e9
    ach_step = step_alg / (ms_step * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": cells / (ms_step * 1e-3) / 1e6,
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"cylinder O-mesh {case.im}x{case.jm} (M=0.1, Re_D=50, seeded free stream): "
                        f"BCs + residual + full colored Jacobian ({nseeds} seeds"
                        f"{', two i-zones' if nzones == 2 else ''}) in the pruned CSR, as a Newton "
                        "iteration after the first (kept pattern checked per row, values rewritten)",
            "cells": cells,
            "seeds": nseeds,
            "csr_nnz": nnz,
            "launch_shapes": tuned,
        },
        "roofline": {
            "kernel": dom + (" (k_jac_csr_rows<2,2,false> + <2,2,true>)" if dom == "k_jac_csr_refresh" else ""),
            "bound": "hbm",
            "achieved": ach_dom,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ach_dom / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom, case),
            "avg_launch_ms": avg_dom,
            "launches": kernels[dom]["launches"],
            "algorithmic_bytes_per_launch": step_alg,
            "algorithmic_bytes": "8 B per CSR value + 96 B per cell (w and metrics) of the step",
            "by_design": {"bytes_per_launch": by_design[dom], "achieved_gbs": ach_dom_design,
                          "frac": ach_dom_design / HBM_PEAK_GBS,
                          "note": "the face blocks read, the CSR values and kept masks written"},
        },
        "step_kernels": kernels,
        "step_achieved_gbs": ach_step,
        "coo_route": {
            "ms_per_step": coo_ms,
            "kernels": coo_kernels,
            "note": "the reference's route: colored assembly into its COO layout (16 B per slot, "
                    f"{16 * f.nbentry / 1e9:.2f} GB), then the device prune + CSR; same CSR bit for bit",
        },
        "graph_replay_coo_route": {
            "ms_per_step": graph_ms,
            "note": "BCs + residual + colored Jacobian into the COO captured once in a HIP graph and replayed, "
                    "then the pruned CSR",
        },
        "first_iteration_build": {
            "ms_per_step": build_ms,
            "kernels": build_kernels,
            "note": "the full CSR build without a previous pattern: count pass, scan, fill pass",
        },
        "residual": {
            "ms": res_ms,
            "mcells_per_s": cells / (res_ms * 1e-3) / 1e6,
            "achieved_gbs": ach_res,
            "frac": ach_res / HBM_PEAK_GBS,
            "face_kernel_avg_ms": face64["avg_ms"],
        },
        "jvp_seed_mcells_per_s": nseeds * cells / (ms_step * 1e-3) / 1e6,
    }
    if not args.no_extras:
        out["newton_bl256x128"] = newton_figure()
        out["jbeta_bl512x192"] = jbeta_figure()
        out["uneven_batch_bl8"] = uneven_batch_figure()
        out["learned_newton"] = learned_figure()
        out["residual_jvp_bl2048x1024"] = residual_jvp_figure()
        out["vjp_bl1024x256"] = vjp_figure()
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def measured_traffic(kernel, case):
    """HBM bytes per launch of `kernel` read by rocprofv3 FETCH_SIZE and
    WRITE_SIZE on this workload (profiles/hbm_traffic_cyl512x384.json, from the
    counter runs its summary lists), or None when not measured."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_cyl{case.im}x{case.jm}.json")
    try:
        with open(path) as fh:
            return json.load(fh)["traffic"].get(kernel)
    except (OSError, KeyError, ValueError):
        return None


def newton_figure():
    """Plain Newton (no learned guess) on the 256x128 boundary layer (M 0.5,
    Re_L 1e5, Blasius-type state + 1e-3 noise): wall time to
    max_e ||R_e||_2 < 1e-10, the per-iteration norms and times."""
    import torch
    from broadcast_amd.cases import bl_case
    from broadcast_amd.flow import Flow, to_device
    from broadcast_amd.newton import newton

    case = bl_case(256, 128, seed=0)
    f = Flow(case)
    w = to_device(case.w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    log = newton(f, w, ite=40, cfl=1e3, tol=1e-10)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"bench: newton 256x128 {wall:.2f} s, {len(log.norms) - 1} steps, converged {log.converged}",
          file=sys.stderr, flush=True)
    return {"seconds_to_1e-10": wall, "converged": log.converged, "newton_steps": len(log.norms) - 1,
            "norms": log.norms, "seconds_per_iteration": log.seconds, "cfl": 1e3,
            "start": "Blasius-type state + 1e-3 noise (seed 0), plain Newton"}


def uneven_batch_figure(reps=20):
    """BCs + residual + JVP (one dense tangent) of the 8 boundary-layer meshes
    of log-uniform size (im 128..1024, jm 64..256, seeds 2000..2007): the
    one-launch batch against the same work as a per-mesh loop of the
    single-mesh entry points.  HBM figure at the dense-JVP algorithmic bytes,
    216 B per cell (wd in + residud out 80 B, shared state and metrics 96 B,
    residu out 40 B)."""
    import numpy as np
    import torch
    from broadcast_amd.batch import UnevenBatch
    from broadcast_amd.cases import uneven_bl_cases
    from broadcast_amd.flow import Flow, to_device

    flows = [Flow(c) for c in uneven_bl_cases()]
    rng = np.random.default_rng(2100)
    ws = [to_device(f.case.w) for f in flows]
    wds = [to_device(rng.standard_normal(f.shape_w)) for f in flows]
    res = [f.new_state() for f in flows]
    resd = [f.new_state() for f in flows]
    b = UnevenBatch(flows)
    cells = b.cells

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms_batch = timed(lambda: b.run(ws, wds, res, resd))

    def per_mesh():
        for k, f in enumerate(flows):
            f.jvp(ws[k], wds[k], res[k], resd[k])

    ms_loop = timed(per_mesh)
    gbs = 216 * cells / (ms_batch * 1e-3) / 1e9
    print(f"bench: uneven batch {cells} cells {ms_batch:.3f} ms (per-mesh loop {ms_loop:.3f} ms)",
          file=sys.stderr, flush=True)
    return {"meshes": [[f.im, f.jm] for f in flows], "cells": cells, "ms_batch": ms_batch,
            "mcells_per_s": cells / (ms_batch * 1e-3) / 1e6, "ms_per_mesh_loop": ms_loop,
            "achieved_gbs_216B_per_cell": gbs, "frac": gbs / HBM_PEAK_GBS}


def residual_jvp_figure(im=2048, jm=1024, reps=10):
    """BCs + residual and BCs + dense JVP (one tangent, BROADCAST_npz.py:
    1079-1085) on a 2048x1024 boundary layer (2.1M cells, 285 MB of
    algorithmic traffic per residual), where HBM rather than launch latency
    sets the pace: Mcells/s and achieved GB/s at 136 B/cell (residual) and
    216 B/cell (JVP with the primal residual)."""
    import numpy as np
    import torch
    from broadcast_amd.cases import bl_case
    from broadcast_amd.flow import Flow, to_device

    case = bl_case(im, jm, seed=3000)
    f = Flow(case)
    w = to_device(case.w)
    wd = to_device(np.random.default_rng(3001).standard_normal(f.shape_w))
    res, resd = f.new_state(), f.new_state()
    f.apply_bcs(w)
    cells = im * jm

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def resid():
        f.apply_bcs(w)
        f.residual(w, res)

    ms_r = timed(resid)
    ms_j = timed(lambda: f.jvp(w, wd, res, resd))
    out = {"cells": cells,
           "residual": {"ms": ms_r, "mcells_per_s": cells / (ms_r * 1e-3) / 1e6,
                        "achieved_gbs": RESIDUAL_B_PER_CELL * cells / (ms_r * 1e-3) / 1e9},
           "jvp": {"ms": ms_j, "mcells_per_s": cells / (ms_j * 1e-3) / 1e6,
                   "achieved_gbs": 216 * cells / (ms_j * 1e-3) / 1e9}}
    for k in ("residual", "jvp"):
        out[k]["frac"] = out[k]["achieved_gbs"] / HBM_PEAK_GBS
    print(f"bench: BL {im}x{jm} residual {ms_r:.3f} ms ({out['residual']['achieved_gbs']:.0f} GB/s), "
          f"JVP {ms_j:.3f} ms ({out['jvp']['achieved_gbs']:.0f} GB/s)", file=sys.stderr, flush=True)
    return out


def vjp_figure(im=1024, jm=256, reps=10):
    """The adjoint wb = (dR/dw)^T resb (flux_num_dnc5_b, cylinder.py:1141) and
    the same followed by the BC sequence's adjoint, on the 1024x256 boundary
    layer (M = 0.5, Re_L = 3e5, perturbed Blasius state): ms per call and
    achieved GB/s at 176 B/cell (w, metrics, resb in; wb out).  Each call
    consumes resb's interior, as the reference's, so the timed loop copies
    resb in first (a 10.8 MB device copy, inside the timing)."""
    import numpy as np
    import torch
    from broadcast_amd.cases import bl_case
    from broadcast_amd.flow import Flow, to_device

    case = bl_case(im, jm, mach=0.5, reynolds=3e5, seed=3000)
    f = Flow(case)
    w = to_device(case.w)
    f.apply_bcs(w)
    resb0 = to_device(np.random.default_rng(3002).standard_normal(f.shape_w))
    rb, wb = f.new_state(), f.new_state()
    cells = im * jm

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def vjp():
        rb.copy_(resb0)
        f.vjp(w, rb, wb, consume=True)

    def vjp_bcs():
        rb.copy_(resb0)
        f.vjp(w, rb, wb, consume=True)
        f.adjoint_bcs(w, wb)

    ms_v = timed(vjp)
    ms_vb = timed(vjp_bcs)
    gbs = 176 * cells / (ms_v * 1e-3) / 1e9
    print(f"bench: BL {im}x{jm} VJP {ms_v:.3f} ms ({gbs:.0f} GB/s), with BC adjoints {ms_vb:.3f} ms",
          file=sys.stderr, flush=True)
    return {"cells": cells, "vjp_ms": ms_v, "mcells_per_s": cells / (ms_v * 1e-3) / 1e6, "achieved_gbs": gbs,
            "frac": gbs / HBM_PEAK_GBS, "vjp_with_bc_adjoints_ms": ms_vb,
            "workload": "flux adjoint (a reverse sweep per face, the wall row by 24 seeded passes contracted "
                        "with the face adjoint, fixed-order gather); then the BC adjoints in reverse order"}


def learned_figure():
    """Newton time to ||R|| < 1e-10, learned guess vs free-stream start, on a
    short protocol that fits the bench: cylinder O-mesh 96x64, 16 training
    baseflows and 4 held-out Sobol points (the measured full run, 192x128
    with 96 + 16, is broadcast_amd.learned's own command)."""
    from broadcast_amd import learned as L

    res = L.experiment(96, 64, n_train=16, n_test=4, width=64, modes=8, steps=1500, log=lambda s: None)
    print(f"bench: learned guess 96x64: over {res['converged_points']} converged points plain "
          f"{res['plain_seconds_converged']:.2f} s, learned {res['learned_seconds_converged']:.2f} s "
          f"(x{res['speedup_wall_converged']:.2f})", file=sys.stderr, flush=True)
    keep = ("mesh", "train", "held_out", "network", "training_s", "converged_points", "not_converged",
            "plain_seconds_converged", "learned_seconds_converged", "speedup_wall_converged",
            "speedup_wall_median_converged", "newton_steps_mean_plain", "newton_steps_mean_learned",
            "all_converged", "fallbacks")
    return {k: res[k] for k in keep}


def jbeta_figure():
    """J(beta) = Jac/vol - beta^2 Dzz + i beta Dz for beta_k = k/15, k = 0..15
    on the 512x192 boundary layer: Jacobian, Dz/Dzz and the 16 complex CSR
    value arrays on the device; HBM figure of the J(beta) fill (the values
    written, 16 B per complex entry per beta, plus the three COOs read)."""
    import torch
    from broadcast_amd import _lib
    from broadcast_amd.cases import bl_case
    from broadcast_amd.flow import Flow, to_device

    case = bl_case(512, 192, seed=2000)
    f = Flow(case)
    w = to_device(case.w)
    f.apply_bcs(w)
    cd = torch.full((case.im * case.jm,), 1e-2, dtype=torch.float64, device="cuda")
    betas = [k / 15.0 for k in range(16)]
    bufs = f.jacobian_buffers()
    dzb = None

    def run():
        nonlocal dzb
        jac, _, _ = f.jacobian_colored(w, coefdiag=cd, bufs=bufs, write_index=dzb is None)
        jdz, jdz2, ia, ja = f.dz_colored(w, bufs=dzb, write_index=dzb is None)
        dzb = (jdz, jdz2, ia, ja)
        return f.jbeta(jac, jdz, jdz2, betas)

    run()
    torch.cuda.synchronize()
    _lib.check(_lib.lib().bca_profile_enable(f.ctx.h, 1), "bca_profile_enable")
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        indptr, indices, vals = run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    tot, n = kernel_ms(f, "k_jbeta_values")
    tot_p, n_p = kernel_ms(f, "k_jbeta_parts")
    dz_tot, dz_n = kernel_ms(f, "k_dz_colored")
    _lib.check(_lib.lib().bca_profile_enable(f.ctx.h, 0), "bca_profile_enable")
    nnz = indices.numel()
    fill_ms = tot / max(n, 1)
    # k_jbeta_values: the 16 complex values written and the 3 components read
    alg = 16 * len(betas) * nnz + 24 * nnz
    print(f"bench: J(beta) 512x192 x16 {ms:.2f} ms per batch, values {fill_ms:.3f} ms", file=sys.stderr, flush=True)
    return {"ms_per_batch": ms, "betas": len(betas), "nnz": nnz, "cells": case.im * case.jm,
            "k_jbeta_values_ms": fill_ms, "k_jbeta_parts_ms": tot_p / max(n_p, 1),
            "k_dz_colored_ms": dz_tot / max(dz_n, 1),
            "k_jbeta_values_gbs": alg / (fill_ms * 1e-3) / 1e9,
            "k_jbeta_values_frac": alg / (fill_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "workload": "Jacobian + Dz/Dzz seed loops + J(beta) CSR values for 16 wavenumbers"}


if __name__ == "__main__":
    main()
