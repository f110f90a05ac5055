#!/usr/bin/env python3
"""Volume-warp benchmark on one MI355X: warpMesh + warpDeriv, fast (kd-tree)
mode at errTol=1e-4, on a star-shaped body with 80 stretched shells of volume
nodes (cube-sphere n=205: 19,975,840 volume nodes, 249,698 unique surface
nodes).  A step is one warpMesh plus one warpDeriv (one adjoint seed, N(0,1)
on a random 10% of the volume nodes) with every input resident in HBM.

Prints one JSON line: ms per step in mixed precision (FP32 far field, FP64
near wall and accumulators) as ``value``, the FP64 kernel and the CPU
oracle (timed on a bounded sample of the same workload, scaled to the full
mesh) beside it, the mixed-vs-FP64 error, and the roofline of the dominant
kernel (algorithmic flops from the device work count / its mean HIP-event
duration over the timed steps).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (= FP64 matrix) peak
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix peak

# algorithmic flops per unit of fast-mode work (FMA = 2, sqrt/div = 1), per
# (V internal visit, A accepted cluster, L bucket point pair, E error check):
WALK_FLOPS = (10.0, 0.0, 0.0, 7.0)   # distance test, error-table interpolation
FWD_FLOPS = (0.0, 32.0, 42.0, 0.0)   # weights and displacement sums
REV_FLOPS = (0.0, 31.0, 41.0, 0.0)   # weights and the 12 adjoint sums


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=205, help="cube-sphere nodes per edge")
    ap.add_argument("--shells", type=int, default=80)
    ap.add_argument("--errtol", type=float, default=1e-4)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile parameters (no autotuner search)")
    ap.add_argument("--only", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--workload", choices=["star-fast", "ellipsoid-exact", "batch"], default="star-fast",
                    help="star-fast: 20M-node star body, fast mode (the headline); ellipsoid-exact: 3.2M-node "
                         "10:1:1 ellipsoid O-mesh, exact mode, interactions/s and adjoint checks; batch: four "
                         "meshes of 0.5M-20M nodes, 16 surfaces and 8 adjoint seeds each, one graph launch")
    ap.add_argument("--fd-directions", type=int, default=20, help="ellipsoid-exact: central-FD directions")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 1 GiB (reference for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def copy_calibration():
    """One device-to-device copy kernel of exactly 2**30 bytes read and
    2**30 bytes written, run 3 times: the per-dispatch FETCH_SIZE and
    WRITE_SIZE it reports convert the counters of the warp kernels to bytes."""
    import torch

    a = torch.ones(2 ** 27, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 2 ** 30, "copy_bytes_written": 2 ** 30}))


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(m, xs, dxv, errtol, target_s):
    """The oracle (C restatement of the reference loops) on a random sample
    of the volume nodes, on all host cores; set-up time is added once and the
    volume loops are scaled by N / sample.  Also returns the last sample's
    results (node indices, warped nodes, and the reverse product of the seed
    restricted to the sample), the checker for the GPU legs."""
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0))
    O.set_threads(cores)
    nvol = len(m["xv0"])
    rng = np.random.default_rng(7)

    def run(k):
        idx = np.sort(rng.choice(nvol, size=k, replace=False))
        om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"][idx], evalMode="fast", errTol=errtol)
        om.setSurfaceCoordinates(xs)
        xv = om.warpMesh()
        sw, lw = O.last_times()
        dxs = om.warpDeriv(dxv.reshape(-1, 3)[idx].ravel())
        sd, ld = O.last_times()
        return sw + sd, lw + ld, dict(idx=idx, xv=xv.reshape(-1, 3), dxs=dxs, ldef=om.Ldef0)

    # grow the sample until its volume loops take about target_s (small
    # samples overstate the per-node cost: thread start-up, cold caches)
    k = min(nvol, 2000)
    setup, loop, chk = run(k)
    for _ in range(4):
        if loop >= 0.6 * target_s or k >= nvol:
            break
        k = int(min(nvol, max(2 * k, k * target_s / max(loop, 1e-3))))
        setup, loop, chk = run(k)
    ms = 1e3 * (setup + loop * nvol / k)
    return dict(value=round(ms, 1), unit="ms per warpMesh+warpDeriv (full mesh, extrapolated)", cores=cores,
                kind="port",
                sample="%d of %d volume nodes (seed 7), fast mode errTol=%g; volume loops %.2fs scaled x%.0f, "
                       "surface set-up %.2fs added once" % (k, nvol, errtol, loop, nvol / k, setup)), chk


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (volume nodes would be sharded across "
              "GPUs, with one all-reduce of dXs); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ellipsoid-exact":
        return ellipsoid_exact(args)
    if args.workload == "batch":
        return batched(args)
    import torch
    from idwarp_amd import USMesh, _lib, synthetic

    t0 = time.time()
    m = synthetic.star_shells(n=args.n, nshell=args.shells)
    xs = m["perturb"](m["pts"])
    nvol, npts = len(m["xv0"]), len(m["pts"])
    dxv = synthetic.random_dxv(nvol, seed=5, frac=0.1)
    log("mesh %s: %d volume nodes, %d surface points (%.1fs)" % (m["name"], nvol, npts, time.time() - t0))

    t0 = time.time()
    mesh = USMesh(options=dict(gridFile=m["xv0"], evalMode="fast", errTol=args.errtol, precision="fp64"))
    mesh.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    ldef = mesh.getLdef0()
    log("init (tree, error tables, denominator estimate): %.1fs" % (time.time() - t0))
    allc = mesh.getFastCounts()
    counts = allc[:4]  # V, A, L, E summed over volume nodes
    # The reverse adds nothing for a node whose seed dXv is zero, and skips
    # such nodes: its algorithmic work is that of the seeded nodes, whose
    # counts come from a mesh of those nodes alone (a node's counts depend
    # only on its position, the tree and errTol).
    seeded = np.flatnonzero(np.abs(dxv.reshape(-1, 3)).sum(1) != 0.0)
    ms_ = USMesh(options=dict(gridFile=m["xv0"][seeded], evalMode="fast", errTol=args.errtol, precision="fp64"))
    ms_.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    scounts = ms_.getFastCounts()[:4]
    del ms_
    kflops = {"walk": float(np.dot(WALK_FLOPS, counts)), "fwd": float(np.dot(FWD_FLOPS, counts)),
              "rev": float(np.dot(REV_FLOPS, scounts)), "rev_dense": float(np.dot(REV_FLOPS, counts))}
    log("seeded nodes %d of %d; reverse flops %.4g (dense seed %.4g)" % (len(seeded), nvol, kflops["rev"],
                                                                          kflops["rev_dense"]))
    util = {"internal_visits": counts[0] / (64 * allc[4]), "accepts": counts[1] / (64 * allc[5]),
            "bucket_pairs": counts[2] / (64 * allc[7])}
    log("work: visits %.3g accepts %.3g leaf pairs %.3g checks %.3g; wave-level %.3g %.3g %.3g %.3g" % tuple(allc))
    log("SIMD lane utilisation of the walk: %s" % {k: round(v, 3) for k, v in util.items()})

    dev = torch.device("cuda:0")
    d_xs = torch.tensor(xs.ravel(), dtype=torch.float64, device=dev)
    d_dxv = torch.tensor(dxv, dtype=torch.float64, device=dev)
    # the same step with a dense seed (dXv ~ N(0,1) on every node), for the
    # reverse's cost when no node can be skipped
    d_dxv_dense = torch.tensor(synthetic.random_dxv(nvol, seed=5, frac=1.0), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    L, ctx = mesh._L, mesh.ctx

    def vp(t):
        return ctypes.c_void_p(t.data_ptr())

    out, tuning = {}, {}
    legs = {"both": ["fp64", "mixed", "mixed_reuse", "mixed_det", "mixed_dense"], "fp64": ["fp64"],
            "mixed": ["mixed"]}[args.only]

    def set_leg(leg):
        mesh.setOption("precision", leg.split("_")[0])
        mesh.setOption("reusePlan", leg.endswith("_reuse"))
        # bit-reproducible reverse: per-entry sums reduced per record in plan order
        mesh.setOption("deterministic", leg.endswith("_det"))
        mesh._setMeshOptions()

    for leg in legs:
        prec = leg.split("_")[0]
        set_leg(leg)
        # the deterministic and dense-seed legs keep the mixed leg's choice
        if not args.no_tune and not leg.endswith("_det") and not leg.endswith("_dense"):
            # the tile parameters for this mesh and precision, searched once
            # before the timed steps (kept for the same options)
            t0 = time.time()
            mesh.tune()
            chosen, table, _ = mesh.getTuning()
            tuning[prec] = {"chosen": chosen, "ms_per_launch": {k: v for k, v in table.items() if v}}
            log("%s tuning (%.1fs): %s %s" % (leg, time.time() - t0, chosen, tuning[prec]["ms_per_launch"]))
        d_xv = torch.empty(3 * nvol, dtype=torch.float64, device=dev)
        d_dxs = torch.empty(3 * npts, dtype=torch.float64, device=dev)
        sp = ctypes.c_void_p(stream.cuda_stream)
        seed_t = d_dxv_dense if leg.endswith("_dense") else d_dxv

        def step():
            _lib.check(L.idw_warp_mesh_dev(ctx, vp(d_xs), vp(d_xv), sp))
            _lib.check(L.idw_warp_deriv_dev(ctx, vp(d_xs), vp(seed_t), vp(d_dxs), sp))

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t1) * 1e3 / max(args.steps, 1)
        # the phases over as many steps again: reading a step's HIP events
        # waits for the step, which the timed loop above must not do
        phases = np.zeros(6)
        for _ in range(args.steps):
            step()
            phases += mesh.getTimings()
        torch.cuda.synchronize()
        phases /= max(args.steps, 1)
        out[leg] = dict(ms=wall, phases=phases, xv=d_xv, dxs=d_dxs)
        log("%s: %.3f ms/step  [nodal %.3f | walk %.3f | fwd %.3f | rev setup %.3f | rev %.3f | nodal rev %.3f]" %
            ((leg, wall) + tuple(phases)))

    head = out.get("mixed", out.get("fp64"))
    head_prec = "mixed" if "mixed" in out else "fp64"
    res = {
        "metric": "warp+warpDeriv ms at 20M nodes: mixed vs FP64 GPU vs CPU, err<=1e-6*Ldef",
        "value": round(head["ms"], 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head_prec == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "star-shaped body, cube-sphere n=%d, %d shells: %d volume nodes, %d unique surface "
                               "nodes; fast mode errTol=%g; warpMesh+warpDeriv, device-resident" %
                               (args.n, args.shells, nvol, _unique(mesh), args.errtol),
                   "precision": head_prec, "n_vol": nvol, "n_surface_points": npts},
    }
    if "fp64" in out:
        res["fp64_ms_per_step"] = round(out["fp64"]["ms"], 3)
    if "mixed_reuse" in out:
        res["mixed_plan_reused_ms_per_step"] = round(out["mixed_reuse"]["ms"], 3)
    if "mixed_det" in out:
        res["mixed_deterministic_ms_per_step"] = round(out["mixed_det"]["ms"], 3)
    if "mixed_dense" in out:
        res["mixed_dense_seed_ms_per_step"] = round(out["mixed_dense"]["ms"], 3)
    if "fp64" in out and "mixed" in out:
        a, b = out["mixed"], out["fp64"]
        res["mixed_speedup_vs_fp64"] = round(b["ms"] / a["ms"], 3)
        res["mixed_max_err_over_Ldef"] = float((a["xv"] - b["xv"]).abs().max().item() / ldef)
        res["mixed_adjoint_rel_err"] = float(((a["dxs"] - b["dxs"]).norm() / b["dxs"].norm()).item())
    # roofline of the dominant kernel launch of the headline leg, the longest
    # of the walk (phase 1), the forward (phase 2) and the reverse (phase 4),
    # each bracketed by HIP events on the context stream.  Kernel names as
    # rocprofv3 reports them; the template argument 0 is the exponent mode of
    # the default exponents (3, 5), compiled as products.
    ph = head["phases"]
    chosen = tuning.get(head_prec, {}).get("chosen", {})
    if head_prec == "mixed":
        # the mixed reverse phase is two launches: k_fast_rev_group takes the groups of four walk waves with
        # sparse seeds (all of them with this seed), k_fast_rev_pair the others' pairs (here it returns at once)
        rev = ("idw::k_fast_rev_group<0, 4>", FP32_PEAK_TFLOPS,
               "FP32 terms (the bulk): FP32 weights on VALU and the sums over the seeded volume nodes of four walk "
               "waves on v_mfma_f32_16x16x4f32, each term of their union added once with FP64 atomics; FP64 terms "
               "near the wall: FP64 weights and v_mfma_f64_16x16x4f64 per wave.  Priced against the 157.3 TFLOP/s "
               "dense FP32 roof; algorithmic work of the seeded nodes only")
        rev_dense = "idw::k_fast_rev_pair<0, %d>" % chosen.get("mergeWindow", 32)
        fwd = ("idw::k_fast_fwd<1, 0, %d>" % chosen.get("forwardChunk", 32), FP32_PEAK_TFLOPS,
               "FP32 weights on VALU and the sums over each wave's 64 volume nodes on v_mfma_f32_16x16x4f32 (FP64 "
               "near the wall and far out).  Priced against the 157.3 TFLOP/s dense FP32 roof")
    else:
        rev_dense = "idw::k_fast_rev<0, 0>"
        rev = ("idw::k_fast_rev<0, 0>", FP64_PEAK_TFLOPS,
               "FP64 weights on VALU, the sums over each wave's seeded volume nodes on v_mfma_f64_16x16x4f64; the "
               "78.6 TFLOP/s FP64 vector roof equals the FP64 matrix roof; algorithmic work of the seeded nodes only")
        fwd = ("idw::k_fast_fwd<0, 0, %d>" % chosen.get("forwardChunk", 32), FP64_PEAK_TFLOPS,
               "FP64 weights on VALU, the sums on v_mfma_f64_16x16x4f64")
    walk = ("idw::k_plan_walk<false, %d>" % chosen.get("walkBuffer", 128), FP64_PEAK_TFLOPS,
            "the reference's accept/recurse tests in FP64 on VALU (distance, square root, error-table "
            "interpolation), decisions bit for bit; priced against the 78.6 TFLOP/s FP64 vector roof")
    cands = {1: walk + (kflops["walk"],), 2: fwd + (kflops["fwd"],), 4: rev + (kflops["rev"],)}
    phase = max(cands, key=lambda i: ph[i])
    kname, peak, compute, kfl = cands[phase]
    kms = ph[phase]
    achieved = kfl / (kms * 1e-3) / 1e12
    res["roofline"] = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                       "frac": round(achieved / peak, 4),
                       "traffic": _traffic_record(kname)["bytes"] if _traffic_record(kname) else None,
                       "kernel": kname, "kernel_ms": round(float(kms), 3),
                       "algorithmic_flops_per_launch": kfl,
                       "compute": compute}
    tr = _traffic_record(kname)
    if tr:
        res["roofline"]["traffic_detail"] = {"fetch_bytes": tr["fetch_bytes"], "write_bytes": tr["write_bytes"],
                                             "hbm_GBps": round(tr["bytes"] / (kms * 1e-3) / 1e9, 1),
                                             "source": "profiles/traffic.json"}
    # every launch of the headline step against its roof, and the reverse with
    # a dense seed (all nodes' work)
    res["kernels"] = {n: {"kernel": c[0], "ms": round(float(ph[i]), 3), "algorithmic_flops": c[3],
                          "TFLOPs": round(c[3] / (ph[i] * 1e-3) / 1e12, 3),
                          "frac": round(c[3] / (ph[i] * 1e-3) / 1e12 / c[1], 4)}
                      for n, i, c in (("walk", 1, cands[1]), ("fwd", 2, cands[2]), ("rev", 4, cands[4]))}
    if "mixed_dense" in out and head_prec == "mixed":
        dms = out["mixed_dense"]["phases"][4]
        res["kernels"]["rev_dense_seed"] = {"kernel": rev_dense, "ms": round(float(dms), 3),
                                            "algorithmic_flops": kflops["rev_dense"],
                                            "TFLOPs": round(kflops["rev_dense"] / (dms * 1e-3) / 1e12, 3),
                                            "frac": round(kflops["rev_dense"] / (dms * 1e-3) / 1e12 / rev[1], 4)}
    if tuning:
        res["tuning"] = tuning
    res["phases_ms"] = {p: dict(zip(("nodal", "walk", "fwd", "rev_setup", "rev", "nodal_rev"),
                                    [round(float(x), 3) for x in out[p]["phases"]])) for p in out}
    res["walk_lane_utilisation"] = {k: round(float(v), 3) for k, v in util.items()}
    if not args.no_cpu:
        t0 = time.time()
        res["cpu_baseline"], chk = cpu_baseline(m, xs, dxv, args.errtol, args.cpu_seconds)
        res["cpu_baseline"]["speedup_gpu_vs_cpu"] = round(res["cpu_baseline"]["value"] / head["ms"], 1)
        log("cpu baseline: %s (%.1fs)" % (res["cpu_baseline"], time.time() - t0))
        # the baseline's sample as the checker of the GPU legs: the warped
        # nodes of the sample, and the reverse product of the seed restricted
        # to it (one more reverse per precision, outside the timed steps)
        idx = torch.as_tensor(chk["idx"], device=dev)
        seed = torch.zeros(nvol, 3, dtype=torch.float64, device=dev)
        seed[idx] = d_dxv.view(-1, 3)[idx]
        vs = {}
        for leg in [k for k in out if not k.endswith("_dense")]:  # the dense-seed leg runs the mixed leg's kernels
            set_leg(leg)
            got = out[leg]["xv"].view(-1, 3)[idx].cpu().numpy()
            d_dxs = torch.empty(3 * npts, dtype=torch.float64, device=dev)
            _lib.check(L.idw_warp_mesh_dev(ctx, vp(d_xs), vp(out[leg]["xv"]), ctypes.c_void_p(stream.cuda_stream)))
            _lib.check(L.idw_warp_deriv_dev(ctx, vp(d_xs), vp(seed), vp(d_dxs), ctypes.c_void_p(stream.cuda_stream)))
            torch.cuda.synchronize()
            vs[leg] = {"max_err_over_Ldef": float(np.abs(got - chk["xv"]).max() / chk["ldef"]),
                       "adjoint_rel_err": _rel(d_dxs.cpu().numpy(), chk["dxs"])}
        res["vs_oracle"] = dict(vs, sample="the CPU baseline's %d sampled volume nodes; reverse product of the "
                                           "step's seed restricted to them" % len(chk["idx"]))
        log("vs oracle: %s" % res["vs_oracle"])
    print(json.dumps(res), flush=True)


def ellipsoid_exact(args):
    """Exact mode on the 10:1:1 ellipsoid O-mesh 257x129x97 (3,215,841 volume
    nodes, 32,514 unique surface nodes): warpMesh + warpDeriv in FP64 and
    mixed precision, interactions/s = pairs evaluated / t (the warp's
    N_vol * N_surf, the reverse's N_seeded * N_surf: it skips the nodes
    with a zero seed; its time with a dense seed is reported beside), the mixed error
    against FP64 (mode for mode), the FP64 error against the oracle on a
    sample of volume nodes, and the adjoint checks: dXv ~ N(0,1) on a random
    10% of the nodes (seed 5), central differences (h = 1e-6) along random
    unit surface directions (seed 6), and the oracle's reverse for a seed
    carried by a 2000-node sample."""
    import torch
    from idwarp_amd import USMesh, _lib, synthetic
    from oracle import oracle as O

    t0 = time.time()
    m = synthetic.ellipsoid_omesh(257, 129, 97)
    xs = m["perturb"](m["pts"])
    nvol, npts = len(m["xv0"]), len(m["pts"])
    dxv = synthetic.random_dxv(nvol, seed=5, frac=0.1)
    log("mesh %s: %d volume nodes, %d surface points (%.1fs)" % (m["name"], nvol, npts, time.time() - t0))
    mesh = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="exact", precision="fp64"))
    mesh.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
    nsurf = _unique(mesh)
    ldef = mesh.getLdef0()
    dev = torch.device("cuda:0")
    d_xs = torch.tensor(xs.ravel(), dtype=torch.float64, device=dev)
    d_dxv = torch.tensor(dxv, dtype=torch.float64, device=dev)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    L, ctx = mesh._L, mesh.ctx

    def vp(t):
        return ctypes.c_void_p(t.data_ptr())

    out = {}
    for prec in ("fp64", "mixed"):
        mesh.setOption("precision", prec)
        mesh._setMeshOptions()
        d_xv = torch.empty(3 * nvol, dtype=torch.float64, device=dev)
        d_dxs = torch.empty(3 * npts, dtype=torch.float64, device=dev)

        def step():
            _lib.check(L.idw_warp_mesh_dev(ctx, vp(d_xs), vp(d_xv), sp))
            _lib.check(L.idw_warp_deriv_dev(ctx, vp(d_xs), vp(d_dxv), vp(d_dxs), sp))

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t1) * 1e3 / max(args.steps, 1)
        phases = np.zeros(6)  # over as many steps again (reading the events waits for each step)
        for _ in range(args.steps):
            step()
            phases += mesh.getTimings()
        torch.cuda.synchronize()
        phases /= max(args.steps, 1)
        out[prec] = dict(ms=wall, phases=phases, xv=d_xv.clone(), dxs=d_dxs.clone())
        log("%s: %.3f ms/step [nodal %.3f | fwd %.3f | rev setup %.3f | rev %.3f | nodal rev %.3f]" %
            ((prec, wall) + tuple(phases[[0, 2, 3, 4, 5]])))
    # the reverse with a dense seed (dXv ~ N(0,1) on every node), mixed precision
    mesh.setOption("precision", "mixed")
    mesh._setMeshOptions()
    d_dense = torch.tensor(synthetic.random_dxv(nvol, seed=5, frac=1.0), dtype=torch.float64, device=dev)
    _lib.check(L.idw_warp_mesh_dev(ctx, vp(d_xs), vp(d_xv), sp))
    rev_dense = []
    for it in range(args.warmup + args.steps):
        _lib.check(L.idw_warp_deriv_dev(ctx, vp(d_xs), vp(d_dense), vp(d_dxs), sp))
        if it >= args.warmup:
            rev_dense.append(mesh.getTimings()[4])
    rev_dense = float(np.mean(rev_dense))
    log("mixed reverse with a dense seed: %.3f ms" % rev_dense)
    pairs = float(nvol) * nsurf
    nseed = int(np.count_nonzero(np.any(dxv.reshape(-1, 3) != 0.0, axis=1)))
    rpairs = float(nseed) * nsurf
    a, b = out["mixed"], out["fp64"]
    dx = (a["xv"] - b["xv"]).abs()
    res = {
        "metric": "exact-mode interactions/s (volume x surface node pairs evaluated per second, warpMesh + warpDeriv)",
        "value": round((pairs + rpairs) / (a["ms"] * 1e-3), 1),
        "unit": "pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(a["ms"], 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "10:1:1 ellipsoid O-mesh 257x129x97: %d volume nodes, %d unique surface nodes; "
                               "exact mode; warpMesh+warpDeriv, device-resident" % (nvol, nsurf),
                   "n_vol": nvol, "n_surface_unique": nsurf, "n_seeded": nseed},
        "fp64_ms_per_step": round(b["ms"], 3),
        "pairs": {"warp": pairs, "reverse": rpairs,
                  "note": "the reverse evaluates the pairs of the %d nodes with a nonzero seed" % nseed},
        "fwd_interactions_per_s": {p: round(pairs / (out[p]["phases"][2] * 1e-3), 1) for p in out},
        "rev_interactions_per_s": {p: round(rpairs / (out[p]["phases"][4] * 1e-3), 1) for p in out},
        "mixed_dense_seed_rev_ms": round(rev_dense, 3),
        "mixed_dense_seed_rev_interactions_per_s": round(pairs / (rev_dense * 1e-3), 1),
        "mixed_speedup_vs_fp64": round(b["ms"] / a["ms"], 3),
        "mixed_vs_fp64": {"max_err_over_Ldef": float(dx.max().item() / ldef),
                          "rms_err_over_Ldef": float(dx.pow(2).mean().sqrt().item() / ldef),
                          "adjoint_rel_err": float(((a["dxs"] - b["dxs"]).norm() / b["dxs"].norm()).item())},
    }
    kms = {p: (out[p]["phases"][2], out[p]["phases"][4]) for p in out}
    fwd_tf = 42.0 * pairs / (kms["mixed"][0] * 1e-3) / 1e12
    rev_tf = 41.0 * rpairs / (kms["mixed"][1] * 1e-3) / 1e12
    mesh.setOption("precision", "mixed")
    mesh._setMeshOptions()
    ech = mesh.getTuning()[0]["exactGroup"]  # points per chunk of the mixed warp
    kname, kt = (("idw::k_exact_rev<1, 0>", rev_tf) if kms["mixed"][1] > kms["mixed"][0]
                 else ("idw::k_exact_fwd_mc<0, %d>" % ech, fwd_tf))
    res["roofline"] = {"bound": "mfma", "achieved": round(kt, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(kt / FP64_PEAK_TFLOPS, 4),
                       "traffic": _traffic_record(kname)["bytes"] if _traffic_record(kname) else None,
                       "kernel": kname, "algorithmic_flops_per_pair": {"fwd": 42, "rev": 41},
                       "fp64_TFLOPs": {"fwd": round(42.0 * pairs / (kms["fp64"][0] * 1e-3) / 1e12, 3),
                                       "rev": round(41.0 * rpairs / (kms["fp64"][1] * 1e-3) / 1e12, 3)},
                       "mixed_TFLOPs": {"fwd": round(fwd_tf, 3), "rev": round(rev_tf, 3),
                                        "rev_dense_seed": round(41.0 * pairs / (rev_dense * 1e-3) / 1e12, 3)},
                       "compute": "mixed warp: FP32 weights on VALU, displacement sums on v_mfma_f32_16x16x4f32 "
                                  "(FP64 sums per chunk); FP64 warp and both reverses on VALU; priced against the "
                                  "78.6 TFLOP/s FP64 vector roof"}
    # central differences of the FP64 warp on the device along unit surface
    # directions (the quotient formed from the difference of the two grids)
    mesh.setOption("precision", "fp64")
    mesh.setSurfaceCoordinates(xs)
    mesh.warpMesh()
    mesh.warpDeriv(dxv, solverVec=False)
    dxs = mesh.getdXs().ravel()
    rng = np.random.default_rng(6)
    worst = 0.0
    for _ in range(args.fd_directions):
        p = rng.standard_normal(xs.size)
        p /= np.linalg.norm(p)
        h = 1e-6
        grids = []
        for sgn in (1.0, -1.0):
            mesh.setSurfaceCoordinates((xs.ravel() + sgn * h * p).reshape(-1, 3))
            mesh.warpMesh()
            grids.append(mesh.getWarpGrid())
        fd = dxv @ (grids[0] - grids[1]) / (2 * h)
        worst = max(worst, abs(fd - dxs @ p) / max(abs(fd), 1e-300))
    res["adjoint_fd"] = {"directions": args.fd_directions, "h": 1e-6, "max_rel_diff": worst,
                         "seed": "dXv ~ N(0,1) on a random 10% of the volume nodes (seed 5)"}
    log("FD check: %s" % res["adjoint_fd"])
    res["stored_weights"] = stored_weights(args)
    if not args.no_cpu:
        # the CPU baseline leg: the oracle on a sample of volume nodes, which
        # also checks the FP64 forward on it and the reverse for a seed carried
        # by it
        t0 = time.time()
        O.set_threads(len(os.sched_getaffinity(0)))
        rng = np.random.default_rng(7)
        idx = np.sort(rng.choice(nvol, size=2000, replace=False))
        om = O.OracleMesh(m["pts"], m["conn"], m["faceSizes"], m["xv0"][idx], evalMode="exact")
        om.setSurfaceCoordinates(xs)
        ref = om.warpMesh().reshape(-1, 3)
        got = b["xv"].view(-1, 3)[torch.as_tensor(idx, device=dev)].cpu().numpy()
        res["fp64_vs_oracle"] = {"sample": "2000 volume nodes (seed 7)",
                                 "max_err_over_Ldef": float(np.abs(got - ref).max() / ldef)}
        seed = np.zeros(3 * nvol)
        g = np.random.default_rng(5).standard_normal((2000, 3))
        seed.reshape(-1, 3)[idx] = g
        mesh.setOption("precision", "fp64")
        mesh.setSurfaceCoordinates(xs)
        mesh.warpMesh()
        mesh.warpDeriv(seed, solverVec=False)
        res["fp64_vs_oracle"]["adjoint_rel_err"] = float(_rel(mesh.getdXs().ravel(), om.warpDeriv(g.ravel())))
        log("oracle checks %.1fs: %s" % (time.time() - t0, res["fp64_vs_oracle"]))
        t0 = time.time()
        om.setSurfaceCoordinates(xs)
        om.warpMesh()
        sw, lw = O.last_times()
        om.warpDeriv(np.ones(3 * len(idx)))
        sd, ld = O.last_times()
        ms = 1e3 * (sw + sd + (lw + ld) * nvol / len(idx))
        res["cpu_baseline"] = {"value": round(ms, 1), "unit": "ms per warpMesh+warpDeriv (full mesh, extrapolated)",
                               "cores": len(os.sched_getaffinity(0)), "kind": "port",
                               "sample": "2000 of %d volume nodes (seed 7), exact mode; volume loops %.2fs scaled "
                                         "x%.0f" % (nvol, lw + ld, nvol / len(idx))}
        log("cpu baseline %.1fs" % (time.time() - t0))
    print(json.dumps(res), flush=True)


def stored_weights(args):
    """Recompute against stored weights (option storeWeights) for the
    reverse product, in exact mode on the CPU-sized sphere O-mesh 49x33x33
    (the 3.2M-node ellipsoid would need 836 GB of weights): device memory of
    the context, the reverse kernels' time (HIP events) and the whole
    warpDeriv, FP64, dXv ~ N(0,1) on a random 10% of the nodes (seed 5)."""
    import torch
    from idwarp_amd import USMesh, _lib, synthetic

    m = synthetic.sphere_omesh(49, 33, 33)
    xs = m["perturb"](m["pts"])
    nvol, npts = len(m["xv0"]), len(m["pts"])
    dev = torch.device("cuda:0")
    d_xs = torch.tensor(xs.ravel(), dtype=torch.float64, device=dev)
    d_dxv = torch.tensor(synthetic.random_dxv(nvol, seed=5, frac=0.1), dtype=torch.float64, device=dev)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    out = {}
    for mode in ("recompute", "stored"):
        g = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="exact", precision="fp64",
                                storeWeights=mode == "stored"))
        g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
        g._setMeshOptions()
        d_xv = torch.empty(3 * nvol, dtype=torch.float64, device=dev)
        d_dxs = torch.empty(3 * npts, dtype=torch.float64, device=dev)
        _lib.check(g._L.idw_warp_mesh_dev(g.ctx, ctypes.c_void_p(d_xs.data_ptr()), ctypes.c_void_p(d_xv.data_ptr()),
                                          sp))
        reps = max(args.steps, 1) * 4
        ker = 0.0
        for r in range(reps + 2):
            if r == 2:
                torch.cuda.synchronize()
                t1 = time.perf_counter()
            _lib.check(g._L.idw_warp_deriv_dev(g.ctx, ctypes.c_void_p(d_xs.data_ptr()),
                                               ctypes.c_void_p(d_dxv.data_ptr()), ctypes.c_void_p(d_dxs.data_ptr()),
                                               sp))
            if r >= 2:
                ker += g.getTimings()[4]
        torch.cuda.synchronize()
        out[mode] = {"device_bytes": g.getDeviceMemory(),
                     "warpDeriv_ms": round((time.perf_counter() - t1) * 1e3 / reps, 4),
                     "reverse_kernels_ms": round(ker / reps, 4), "dxs": d_dxs.cpu().numpy()}
    a, b = out["stored"], out["recompute"]
    rec = {"workload": "sphere O-mesh 49x33x33: %d volume nodes, %d unique surface nodes, exact mode, FP64" %
                       (nvol, _unique(g)),
           "weights_bytes": a["device_bytes"] - b["device_bytes"],
           "adjoint_rel_diff": _rel(a["dxs"], b["dxs"])}
    for k in ("recompute", "stored"):
        rec[k] = {kk: vv for kk, vv in out[k].items() if kk != "dxs"}
    log("stored weights: %s" % rec)
    return rec


def batched(args):
    """Four meshes of uneven size (ellipsoid O-meshes 129x65x61, 193x97x107,
    257x129x151 and the 20M-node star body), fast mode errTol 1e-4, mixed
    precision; per mesh 16 surfaces Xs0 + a_p (Xs1 - Xs0), a_p ~ U(0.5, 1.5)
    (seeds 100+p), and 8 reverse products with dXv ~ N(0,1) on 10% of the
    nodes (seeds 200+s).  All 96 solves are recorded into one HIP graph and
    replayed; reported beside the same solves as sequential FP64 calls that
    walk the tree every time."""
    import torch
    from idwarp_amd import USMesh, _lib, synthetic
    from idwarp_amd.batch import Batch

    dev = torch.device("cuda:0")
    specs = [("ellipsoid 129x65x61", lambda: synthetic.ellipsoid_omesh(129, 65, 61)),
             ("ellipsoid 193x97x107", lambda: synthetic.ellipsoid_omesh(193, 97, 107)),
             ("ellipsoid 257x129x151", lambda: synthetic.ellipsoid_omesh(257, 129, 151)),
             ("star n=205, 80 shells", lambda: synthetic.star_shells(205, 80))]
    meshes = []
    for name, make in specs:
        t0 = time.time()
        m = make()
        g = USMesh(options=dict(gridFile=m["xv0"], fileType="memory", evalMode="fast", errTol=1e-4,
                                precision="mixed"))
        g.setSurfaceDefinition(m["pts"], m["conn"], m["faceSizes"])
        base = m["pts"]
        delta = m["perturb"](base) - base
        xs = [torch.tensor((base + np.random.default_rng(100 + p).uniform(0.5, 1.5) * delta).ravel(),
                           dtype=torch.float64, device=dev) for p in range(16)]
        dxv = [torch.tensor(synthetic.random_dxv(len(m["xv0"]), seed=200 + q, frac=0.1), dtype=torch.float64,
                            device=dev) for q in range(8)]
        xv = [torch.empty(3 * g.nVol, dtype=torch.float64, device=dev) for _ in range(16)]
        dxs = [torch.empty(3 * g.nSurf, dtype=torch.float64, device=dev) for _ in range(8)]
        meshes.append(dict(name=name, g=g, xs=xs, dxv=dxv, xv=xv, dxs=dxs, nvol=g.nVol))
        log("%s: %d volume nodes (%.1fs)" % (name, g.nVol, time.time() - t0))

    def fill(b, ms):
        for m in ms:
            for x, o in zip(m["xs"], m["xv"]):
                b.add_warp(m["g"], x, o)
            for d, o in zip(m["dxv"], m["dxs"]):
                b.add_warp_deriv(m["g"], m["xs"][0], d, o)

    def timed(run, reps):
        run()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(reps):
            run()
        torch.cuda.synchronize()
        return (time.perf_counter() - t1) / reps

    b = Batch()
    fill(b, meshes)
    t0 = time.time()
    b.run()
    torch.cuda.synchronize()
    layout, lms = b.info()
    log("recorded (warm-up, two layouts timed: %s ms, layout %d) in %.1fs" % (lms, layout, time.time() - t0))
    reps = max(1, args.steps // 2)
    t_all = timed(b.run, reps)
    solves = sum(24 for _ in meshes)
    per_mesh = {}
    for m in meshes:
        bm = Batch()
        fill(bm, [m])
        t_graph = timed(bm.run, reps)
        # the same solves as sequential FP64 calls, walking the tree each time
        g = m["g"]
        g.setOption("precision", "fp64")
        g.setOption("reusePlan", False)
        g._setMeshOptions()
        L, ctx = g._L, g.ctx
        sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)

        def seq():
            for x, o in zip(m["xs"], m["xv"]):
                _lib.check(L.idw_warp_mesh_dev(ctx, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(o.data_ptr()), sp))
            for d, o in zip(m["dxv"], m["dxs"]):
                _lib.check(L.idw_warp_deriv_dev(ctx, ctypes.c_void_p(m["xs"][0].data_ptr()),
                                                ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(o.data_ptr()), sp))

        t_seq = timed(seq, 1)
        g.setOption("precision", "mixed")
        g._setMeshOptions()
        chosen, table, _ = g.getTuning()
        per_mesh[m["name"]] = {"n_vol": m["nvol"], "tuning": {"chosen": chosen,
                                                               "ms_per_launch": {k: v for k, v in table.items() if v}},
                               "graph_ms": round(t_graph * 1e3, 2),
                               "sequential_fp64_ms": round(t_seq * 1e3, 2),
                               "speedup": round(t_seq / t_graph, 2)}
        log("%s: graph %.1f ms, sequential FP64 %.1f ms" % (m["name"], t_graph * 1e3, t_seq * 1e3))
        del bm
    seq_total = sum(v["sequential_fp64_ms"] for v in per_mesh.values())
    res = {
        "metric": "batched solves/s: 4 meshes (0.5M-20M volume nodes) x (16 warps + 8 reverse products), one graph",
        "value": round(solves / t_all, 2), "unit": "solves/s", "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": round(t_all * 1e3, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "ellipsoid O-meshes 129x65x61, 193x97x107, 257x129x151 and the 20M-node star body; "
                               "fast mode errTol=1e-4, mixed; 16 surfaces + 8 reverse products per mesh, recorded "
                               "into one HIP graph",
                   "graph_layout": ["branch per mesh", "chain"][layout], "autotune_replay_ms": lms},
        "per_mesh": per_mesh,
        "sequential_fp64_ms": round(seq_total, 2),
        "speedup_vs_sequential_fp64": round(seq_total / (t_all * 1e3), 2),
    }
    print(json.dumps(res), flush=True)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _unique(mesh):
    n = ctypes.c_int64(0)
    mesh._L.idw_get_sizes(mesh.ctx, None, None, ctypes.byref(n), None)
    return n.value


def _traffic_record(kname):
    """HBM bytes per launch of a kernel measured with rocprofv3 --pmc
    (FETCH_SIZE, then WRITE_SIZE, each scaled by what the counter reads for a
    1 GiB device copy), as recorded in profiles/traffic.json, or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(kname)
    except (OSError, ValueError):
        return None
    return rec if isinstance(rec, dict) and "bytes" in rec else None


if __name__ == "__main__":
    main()
