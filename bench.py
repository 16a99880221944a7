"""Benchmark: near-quadrature pairs/s (getnearquad_helm_comb_dir on the GPU) and GMRES
solve time to 1e-6 relative error (plain FP64 vs learned order selection + FP32-inner /
FP64-refined GMRES + batched right-hand sides).

Workload (default, the largest configuration): Helmholtz combined field, zk = 10,
alpha = i zk, beta = 1, on a sphere deformed radially by 1 + 0.2 sin(3 theta) cos(2 phi),
cube-projected at 32x32 cells per face (12288 triangles, order 6, 344064 points),
eps = 1e-6.  A step is one full near-quadrature generation over every near (target,
patch) pair with all inputs resident in HBM.  The same JSON line carries the applies of
that configuration and the solves of the smaller configurations (Laplace sphere,
Helmholtz sphere, 64 batched plane waves, wavy torus with learned order selection);
--combined adds the combined solve of the large configuration (16 plane waves, plain
FP64 against learned + mixed + batched), which takes about ten minutes.

Prints one JSON line (progress lines go to stderr).  --gpus N with N > 1 exits: this
build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

# the CPUs this process may run on, read before any OpenMP runtime can bind the main
# thread to one of them
_AFFINITY0 = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

F_K_HELM = 40.0
F_K_LAP = 22.0
PEAK_FP64 = 78.6e12  # MI355X dense FP64 (vector and matrix) FLOP/s
PEAK_HBM = 8.0e12


_T0 = time.time()


def _progress(msg):
    """One line per section on stderr (the JSON line stays alone on stdout)."""
    print("[bench %6.1f s] %s" % (time.time() - _T0, msg), file=sys.stderr, flush=True)


def far_flops_per_pair(m):
    """Far sum, flops per (target, source) pair for m right-hand sides: 23 for the pair
    (difference 3, r^2 5, 1/r 1, r 1, n.(x-y) 5, phase 1, sin and cos 2, e^{ikr}/r 2,
    dipole factor 3) and 16 per density (q + m t(1-ikr): complex multiply-add 8; times
    e^{ikr}/r: complex multiply 6; accumulate 2)."""
    return 23.0 + 16.0 * m


def workload(cfg):
    from fmm3dbie_amd import geometry as G
    if cfg == 0:
        g = G.get_sphere_npat(1.0, 4, (0, 0, 0), 3)
        return g, False, np.array([1.0, 1.0]), "laplace_sphere_4x4_o3", None
    if cfg == 1:
        g = G.get_sphere_npat(1.0, 12, (0, 0, 0), 4)
        zk = 1.1
        return g, True, np.array([zk, 1j * zk, 1.0]), "helmholtz_zk1.1_sphere_12x12_o4", zk
    if cfg == 3:
        g = G.get_wavy_torus(48, 16, 6)
        zk = 3.0
        return g, True, np.array([zk, 1j * zk, 1.0]), "helmholtz_zk3_wavytorus_48x16_o6", zk
    if cfg == 4:
        g = G.get_deformed_sphere(32, 6)
        zk = 10.0
        return g, True, np.array([zk, 1j * zk, 1.0]), "helmholtz_zk10_deformedsphere_32x32_o6", zk
    raise SystemExit("config %d not wired into bench.py" % cfg)


def measured_traffic(kernel, wname):
    """HBM bytes per launch of `kernel` on workload `wname` from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size), or None."""
    import glob
    for f in glob.glob(os.path.join(HERE, "profiles", "*.txt")):
        for line in open(f):
            if line.startswith("TRAFFIC_JSON "):
                d = json.loads(line[len("TRAFFIC_JSON "):])
                if d.get("kernel") == kernel and d.get("workload") == wname:
                    return d
    return None


def pair_flops(stats, p, kq, nf, fk):
    """Algorithmic flops per launch of each quadrature kernel (adaptive: kq(1+4R)(F_K+4p)
    + 4p^2 per pair with R refinements; fixed: nf(F_K+4p) + 4p^2; self: n_s(F_K+22p+15)
    + 4p^2 with the actual self-rule size)."""
    nref = stats["nref"][stats["cls"] == 1].astype(np.float64)
    adap = np.sum(kq * (1 + 4 * nref) * (fk + 4 * p) + 4 * p * p)
    fixed = stats["nfixed"] * (nf * (fk + 4 * p) + 4 * p * p)
    return adap, fixed, nref


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-learned", action="store_true", help="skip the torus learned+mixed solve")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-plane-wave batched solve")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the configuration's own sections (no solves of the smaller configurations)")
    ap.add_argument("--combined", action="store_true",
                    help="with --config 4: the combined solve of 16 plane waves (plain FP64 against learned + "
                         "mixed + batched; about ten minutes, so not part of the default run)")
    ap.add_argument("--combined-fp64-rhs", type=int, default=2,
                    help="right-hand sides of the 344k-point combined run solved by plain FP64 GMRES (timed, "
                         "scaled to 16) and by the converged FP64 comparator")
    ap.add_argument("--batched-sequential", type=int, default=8,
                    help="FP64 single-RHS solves timed for the batched comparison")
    ap.add_argument("--apply-only", action="store_true",
                    help="only the near quadrature once, then 5 FP64 (add/subtract), 5 FP32 (masked) and 5 16-RHS FP32 applies "
                         "of the workload (for the apply's rocprofv3 counters)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device read + 1 GiB write (torch.neg) for HBM counter scaling")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"read_bytes": a.numel() * 4, "write_bytes": b.numel() * 4,
                                               "launches": 3}}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus %d" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    from fmm3dbie_amd import problem, layer

    dev = torch.device("cuda:0")
    g, helm, pars, wname, zk = workload(args.config)
    _progress("workload %s: %d points" % (wname, len(g["srcvals"])))
    eps = 1e-6
    # set-up as the reference's solver does it (near list, quadrature offsets, far orders,
    # oversampled geometry), the near list and far orders on the GPU; on the smaller
    # surfaces the host set-up is timed beside it and must give the same lists and orders
    problem.setup(g, eps, helm, pars, gpu_near=True)
    t0 = time.time()
    P = problem.setup(g, eps, helm, pars, gpu_near=True)
    t_setup_gpu = time.time() - t0
    t_setup_host, setup_same = None, None
    if args.config != 4:
        t0 = time.time()
        Ph = problem.setup(g, eps, helm, pars)
        t_setup_host = time.time() - t0
        setup_same = bool(np.array_equal(Ph["col_ind"], P["col_ind"]) and np.array_equal(Ph["novers"], P["novers"]))
        del Ph
    N = len(g["srcvals"])
    p = int(np.diff(g["ixyzs"])[0])
    nq = layer.NearQuadDevice(g, helm, pars, eps, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=dev)
    if args.apply_only:
        print(json.dumps(apply_only(g, helm, pars, P, nq, dev)))
        return
    stream = torch.cuda.current_stream()
    _progress("set-up done; %d warm-up and %d timed generations" % (args.warmup, args.steps))
    for _ in range(args.warmup):
        nq.run()
    torch.cuda.synchronize()
    # timed steps
    ms_adap = []
    ms_fixed = []
    ms_self = []
    ms_prep = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nq.run()
        ms_adap.append(nq.stats["ms_adap"])
        ms_fixed.append(nq.stats["ms_fixed"])
        ms_self.append(nq.stats["ms_self"])
        ms_prep.append(nq.stats["ms_prep"])
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    nnz = P["nnz"]
    value = nnz / dt
    # one detailed run for the work counts (refinements per pair)
    nq.run(want_detail=True)
    st = nq.stats
    # the same kernels alone: every kernel in order on one stream (in the timed steps the
    # fixed-rule and self kernels run beside the adaptive one and share the GPU with it)
    layer.set_side_stream(False)
    iso = {"adaptive": [], "fixed": [], "self": []}
    for _ in range(max(3, args.steps // 4)):
        nq.run()
        iso["adaptive"].append(nq.stats["ms_adap"])
        iso["fixed"].append(nq.stats["ms_fixed"])
        iso["self"].append(nq.stats["ms_self"])
    layer.set_side_stream(True)
    iso = {k: float(np.mean(v)) for k, v in iso.items()}
    fk = F_K_HELM if helm else F_K_LAP
    adap_fl, fixed_fl, nref = pair_flops(st, p, st["kq"], st["nf"], fk)
    ka = float(np.mean(ms_adap)) * 1e-3
    kf = float(np.mean(ms_fixed)) * 1e-3
    kself = float(np.mean(ms_self)) * 1e-3
    dominant = max([("adaptive", ka, adap_fl), ("fixed", kf, fixed_fl), ("self", kself, None)], key=lambda x: x[1])
    name, kt, fl = dominant
    if fl is None:
        fl = adap_fl
        name, kt = "adaptive", ka
    achieved = fl / kt
    out = {
        "metric": "near-quad pairs/s",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if helm else "f64",
        "data": "synthetic (analytic geometry generated in-process)",
        "config": {"workload": wname, "npts": N, "npatches": len(g["norders"]), "norder": p and int(g["norders"][0]),
                   "eps": eps, "nnz": nnz, "nquad": P["nquad"], "pairs_self": st["nself"],
                   "pairs_adaptive": st["nadap"], "pairs_fixed": st["nfixed"],
                   "refinements_mean": float(nref.mean()) if nref.size else 0.0},
        "kernels_ms": {"prep": float(np.mean(ms_prep)), "fixed": kf * 1e3, "adaptive": ka * 1e3,
                       "self": kself * 1e3},
        "kernels_note": "timed steps: fixed-rule and self kernels on a second stream beside the adaptive one; "
                        "kernels_ms_alone: each kernel alone, in order on one stream",
        "kernels_ms_alone": iso,
        "roofline": {"bound": "valu", "pipe": "fp64 VALU (no MFMA in this kernel; MI355X FP64 vector peak = "
                                              "FP64 matrix peak = 78.6 TFLOP/s)",
                     "kernel": "k_adaptive", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": None,
                     "alone": {"ms": iso["adaptive"], "achieved": fl / (iso["adaptive"] * 1e-3) / 1e12,
                               "frac": fl / (iso["adaptive"] * 1e-3) / PEAK_FP64}},
        "setup_host_s": t_setup_host,
        "setup_gpu_near_s": t_setup_gpu,
        "setup_gpu_near_identical": setup_same,
    }
    tr = measured_traffic("k_adaptive", wname) if name == "adaptive" else None
    if tr is not None:
        out["roofline"]["traffic"] = tr["traffic_bytes"]
        out["roofline"]["traffic_source"] = "profiles/ PMC summary (FETCH_SIZE x %.2f + WRITE_SIZE x %.2f)" % (
            tr["fetch_scale"], tr["write_scale"])
    # GMRES solves
    if args.config == 4:
        _progress("applies")
        out["apply"] = large_apply(g, pars, P, nq.wnear, dev)
        from fmm3dbie_amd import near
        rfac, _ = near.get_rfacs(near.norder_avg(g))
        near.findnear_gpu(P["cms"], P["rads"] * rfac, g["srcvals"][:, :3])
        t0 = time.perf_counter()
        rpg, cig = near.findnear_gpu(P["cms"], P["rads"] * rfac, g["srcvals"][:, :3])
        tg = time.perf_counter() - t0
        out["findnear"] = {"gpu_s_incl_copies": tg, "same_as_setup": bool(np.array_equal(rpg, P["row_ptr"])
                                                                            and np.array_equal(cig, P["col_ind"]))}
        if args.combined:
            _progress("combined solve")
            out["combined_solve"] = combined_solve(g, pars, P, nq, dev, dt, args.combined_fp64_rhs)
    elif not args.no_solve and helm and args.config == 1:
        out["solve"] = sphere_solve(dev, (g, pars, P, nq.wnear))
    del nq
    if not args.no_cpu:
        _progress("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(g, helm, pars, eps, P, args.cpu_seconds)
    if not args.no_solve and not args.no_extra:
        _progress("laplace solve")
        out["laplace_solve"] = laplace_solve(dev)
        if args.config != 1:
            _progress("sphere solve")
            out["sphere_solve"] = sphere_solve(dev)
        if not args.no_learned:
            _progress("torus learned + mixed solve")
            out["learned_mixed_solve"] = learned_mixed_solve(dev)
        if not args.no_batched:
            _progress("batched solve")
            out["batched_solve"] = batched_solve(dev, args.batched_sequential)
    print(json.dumps(out))


# Residual tolerance of the timed solves.  With the combined-field operator's condition
# number a relative residual of 3e-7 keeps every solution within 1e-6 (weighted relative
# l2) of the converged FP64 solution, which each section reports.
RTOL = 3e-7
RTOL_REF = 1e-10  # the converged FP64 comparator (untimed)


def _wrel(x, ref, wts):
    """Relative error in the quadrature-weighted l2 norm (max over right-hand sides)."""
    x = np.atleast_2d(x)
    ref = np.atleast_2d(ref)
    return float(max(np.sqrt((np.abs(a - b) ** 2 * wts).sum() / (np.abs(b) ** 2 * wts).sum()) for a, b in zip(x, ref)))


def sphere_solve(dev, case=None):
    """Helmholtz zk = 1.1, alpha = i zk, beta = 1, cube sphere 12x12 cells per face, order 4
    (25920 points), point source at (0.1, 0.2, -0.3): plain FP64 GMRES against FP32 inner
    GMRES with FP64 iterative refinement, both to a relative residual of RTOL; each
    solution's error against a converged FP64 solve, the exterior field at 10 points of
    radius 3 against the exact field, and the per-phase apply times."""
    import torch
    from fmm3dbie_amd import geometry as G, layer, problem
    if case is None:
        g = G.get_sphere_npat(1.0, 12, (0, 0, 0), 4)
        pars = np.array([1.1, 1.1j, 1.0])
        P = problem.setup(g, 1e-6, True, pars, gpu_near=True)
        nq = layer.NearQuadDevice(g, True, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=dev)
        wnear = nq.run()
    else:
        g, pars, P, wnear = case
    zk = float(np.real(pars[0]))
    N = len(g["srcvals"])
    x0 = np.array([0.1, 0.2, -0.3])
    rhs = torch.as_tensor(problem.point_source(g["srcvals"][:, :3], x0, zk), device=dev)
    wts = torch.as_tensor(P["wts"], device=dev)
    plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], wnear, P["novers"],
                           P["ixyzso"], P["srcover"], P["whtsover"])
    npts_, nptso, _, _ = plan.info()
    tm, t64, t32 = np.zeros(4), np.zeros(4), np.zeros(4)
    s32 = rhs.to(torch.complex64)
    plan.apply(rhs, 64, 0)
    plan.apply(s32, 32, 1)
    for _ in range(3):  # per-phase HIP-event times, mean of 3 warm applies each
        plan.apply(rhs, 64, 0, times=tm)
        t64 += tm / 3
        plan.apply(s32, 32, 1, times=tm)
        t32 += tm / 3
    xref, nref_it, _, rref, _ = layer.gmres(plan, rhs, 0.5, wts, numit=200, eps_gmres=RTOL_REF)
    layer.gmres(plan, rhs, 0.5, wts, numit=200, eps_gmres=RTOL)
    (x64, nit, _, rres64, _), ts64 = _timed(lambda: layer.gmres(plan, rhs, 0.5, wts, numit=200, eps_gmres=RTOL))
    layer.gmres_ir(plan, rhs, 0.5, wts, numit=200, eps_inner=RTOL / 2, tol=RTOL)
    (xir, outer, inner, rresir, n32, n64), tsir = _timed(
        lambda: layer.gmres_ir(plan, rhs, 0.5, wts, numit=200, eps_inner=RTOL / 2, tol=RTOL))
    targ = 3.0 * problem.plane_wave_dirs(10, 0)
    exact = problem.point_source(targ, x0, zk)
    xr = xref.cpu().numpy()
    f64 = _exterior_field(g, P["wts"], x64.cpu().numpy(), targ, pars, zk)[0]
    fir = _exterior_field(g, P["wts"], xir.cpu().numpy(), targ, pars, zk)[0]
    res = {"workload": "helmholtz_zk1.1_sphere_12x12_o4_pointsource", "npts": N, "nnz": P["nnz"],
           "rtol": RTOL, "gmres_fp64_s": ts64, "gmres_fp64_iters": int(nit[0]), "gmres_fp64_rres": float(rres64[0]),
           "ir_mixed_s": tsir, "ir_outer": outer, "ir_inner_iters": inner, "ir_fp32_applies": n32,
           "ir_fp64_applies": n64, "ir_rres": float(rresir[0]), "speedup_mixed_vs_fp64": ts64 / tsir,
           "reference_solve": {"rres": float(rref[0]), "iters": int(nref_it[0])},
           "err_vs_converged_fp64": {"plain_fp64": _wrel(x64.cpu().numpy(), xr, P["wts"]),
                                     "mixed": _wrel(xir.cpu().numpy(), xr, P["wts"])},
           "field_err_fp64": float(np.abs(f64 - exact).max() / np.abs(exact).max()),
           "field_err_mixed": float(np.abs(fir - exact).max() / np.abs(exact).max()),
           "nptso": nptso,
           "far_fp64_tflops": npts_ * nptso * far_flops_per_pair(1) / (t64[1] * 1e-3) / 1e12,
           "far_fp32_tflops": npts_ * nptso * far_flops_per_pair(1) / (t32[1] * 1e-3) / 1e12,
           # near correction (fused with the far reduction): wnear + CSR index arrays + row
           # pointers + sigma and pot (the partial sums read are not counted)
           "csr_fp64_gbs": (16.0 * P["nquad"] + 16.0 * P["nnz"] + 8.0 * (N + 1) + 32.0 * N) / (t64[2] * 1e-3) / 1e9,
           "csr_fp32_gbs": (8.0 * P["nquad"] + 16.0 * P["nnz"] + 8.0 * (N + 1) + 16.0 * N) / (t32[2] * 1e-3) / 1e9,
           "apply_fp64_ms": {"oversample": t64[0], "far": t64[1], "csr": t64[2], "subtract": t64[3]},
           "apply_fp32_ms": {"oversample": t32[0], "far": t32[1], "csr": t32[2]}}
    plan.close()
    return res


def _exterior_field(g, wts, x, targ, pars, zk):
    """(alpha S + beta D) x at off-surface targets by the smooth rule (targets far away)."""
    d = targ[:, None, :] - g["srcvals"][None, :, :3]
    r = np.linalg.norm(d, axis=-1)
    dn = (d * g["srcvals"][None, :, 9:12]).sum(-1)
    ker = (pars[2] * dn * (1 - 1j * zk * r) / r ** 3 + pars[1] / r) * np.exp(1j * zk * r) / (4 * np.pi)
    return (np.atleast_2d(x) * wts[None, :]) @ ker.T


def _timed(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0


def apply_only(g, helm, pars, P, nq, dev):
    """5 FP64 applies in the reference's add/subtract algebra, 5 FP32 masked applies of a
    random density and 5 FP32 masked applies of 16 random densities (the far sum on the
    matrix cores), with the per-phase HIP-event times (ms): oversampling, far sum, split
    sum + CSR correction, subtraction."""
    import torch
    from fmm3dbie_amd import layer
    w = nq.run()
    plan = layer.ApplyPlan(g, helm, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                           P["srcover"], P["whtsover"])
    npts, nptso, _, _ = plan.info()
    rng = np.random.default_rng(0)
    sig = torch.as_tensor(rng.standard_normal(npts) + 1j * rng.standard_normal(npts), device=dev)
    out = {"workload_points": npts, "oversampled_sources": nptso, "nnz": P["nnz"], "nquad": P["nquad"]}
    sig16 = torch.as_tensor(rng.standard_normal((16, npts)) + 1j * rng.standard_normal((16, npts)),
                            device=dev).to(torch.complex64)
    for name, prec, mode, s in (("fp64_addsub", 64, 0, sig), ("fp32_masked", 32, 1, sig.to(torch.complex64)),
                                ("fp32_masked_16rhs", 32, 1, sig16)):
        tm = np.zeros(4)
        acc = []
        plan.apply(s, prec, mode)
        for _ in range(5):
            plan.apply(s, prec, mode, times=tm)
            acc.append(tm.copy())
        torch.cuda.synchronize()
        out[name + "_ms"] = dict(zip(("oversample", "far", "csr", "subtract"), np.mean(acc, axis=0).tolist()))
    plan.close()
    return out


def large_apply(g, pars, P, wnear, dev):
    """Apply timings at the 344k-point configuration (no solve: one FP64 matvec costs
    N x nptso pair interactions): FP64 and FP32 single right-hand side, FP32 with 16
    right-hand sides after autotuning; the FP32 result against FP64."""
    import torch
    from fmm3dbie_amd import layer, problem
    plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], wnear, P["novers"], P["ixyzso"],
                           P["srcover"], P["whtsover"])
    npts, nptso, _, _ = plan.info()
    dirs = problem.plane_wave_dirs(16, 4)
    zk = float(pars[0].real)
    S = torch.as_tensor(-problem.plane_waves(g["srcvals"][:, :3], dirs, zk), device=dev)
    t64, t32, t32b = np.zeros(4), np.zeros(4), np.zeros(4)
    p64 = plan.apply(S[0], 64, 1, times=t64)
    p32 = plan.apply(S[0].to(torch.complex64), 32, 1, times=t32)
    rel = float((p32.to(torch.complex128) - p64).abs().max() / p64.abs().max())
    tune = plan.tune(32, 16)
    plan.apply(S.to(torch.complex64), 32, 1, times=t32b)
    plan.close()
    fl = npts * nptso * far_flops_per_pair(1)
    return {"npts": npts, "nptso": nptso, "far_fp64_ms": t64[1], "far_fp32_ms": t32[1],
            "far_fp32_16rhs_ms": t32b[1], "far_fp64_tflops": fl / (t64[1] * 1e-3) / 1e12,
            "far_fp32_tflops": fl / (t32[1] * 1e-3) / 1e12,
            "far_fp32_16rhs_tflops": npts * nptso * far_flops_per_pair(16) / (t32b[1] * 1e-3) / 1e12,
            "csr_fp64_ms": t64[2], "fp32_vs_fp64_rel": rel,
            "tuned": min(tune, key=lambda t: t[2]) if tune else None}


def laplace_solve(dev):
    """Laplace combined field (alpha = beta = 1) on the 4x4 cube sphere, order 3 (1920
    points): near-quadrature rate and the exterior Dirichlet solve (GMRES tol 1e-7) for a
    unit charge at (0.1, 0.2, -0.3), field checked at 10 points of radius 3."""
    import torch
    from fmm3dbie_amd import geometry as G, layer, problem
    pars = np.array([1.0, 1.0])
    g = G.get_sphere_npat(1.0, 4, (0, 0, 0), 3)
    P = problem.setup(g, 1e-6, False, pars)
    nq = layer.NearQuadDevice(g, False, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=dev)
    for _ in range(3):
        nq.run()
    _, tq = _timed(lambda: [nq.run() for _ in range(20)])
    x0 = np.array([0.1, 0.2, -0.3])
    rhs = torch.as_tensor(problem.point_source(g["srcvals"][:, :3], x0).astype(complex), device=dev)
    wts = torch.as_tensor(P["wts"], device=dev)
    plan = layer.ApplyPlan(g, False, pars, P["row_ptr"], P["col_ind"], P["iquad"], nq.wnear, P["novers"],
                           P["ixyzso"], P["srcover"], P["whtsover"])
    layer.gmres(plan, rhs, 0.5, wts, numit=100, eps_gmres=1e-7)
    (x, nit, _, rres, _), ts = _timed(lambda: layer.gmres(plan, rhs, 0.5, wts, numit=100, eps_gmres=1e-7))
    targ = 3.0 * problem.plane_wave_dirs(10, 0)
    d = targ[:, None, :] - g["srcvals"][None, :, :3]
    r = np.linalg.norm(d, axis=-1)
    dn = (d * g["srcvals"][None, :, 9:12]).sum(-1)
    ker = (dn / r ** 3 + 1.0 / r) / (4 * np.pi)
    u = ker @ (x.cpu().numpy().real * P["wts"])
    exact = problem.point_source(targ, x0)
    plan.close()
    return {"workload": "laplace_sphere_4x4_o3_pointcharge", "npts": len(g["srcvals"]), "nnz": P["nnz"],
            "pairs_per_s": P["nnz"] / (tq / 20), "quad_ms": tq / 20 * 1e3, "gmres_s": ts, "iters": int(nit[0]),
            "rres": float(rres[0]), "field_err": float(np.abs(u - exact).max() / np.abs(exact).max())}


def learned_mixed_solve(dev):
    """Wavy torus 48x16 cells, order 6, zk=3, interior point source at (2, 0, 0.05):
    plain FP64 GPU form (adaptive quadrature, FP64 GMRES) against learned + mixed (selector
    routed quadrature, FP32 inner GMRES with FP64 iterative refinement), both to 1e-6."""
    import torch
    from fmm3dbie_amd import geometry as G, layer, problem, selector as S
    zk = 3.0
    pars = np.array([zk, 1j * zk, 1.0])
    g = G.get_wavy_torus(48, 16, 6)
    P = problem.setup(g, 1e-6, True, pars)
    x0 = np.array([2.0, 0.0, 0.05])
    rhs = torch.as_tensor(problem.point_source(g["srcvals"][:, :3], x0, zk), device=dev)
    wts = torch.as_tensor(P["wts"], device=dev)
    nq = layer.NearQuadDevice(g, True, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=dev)
    # labels and features from one plain run, GPU training on 2e5 sampled pairs (seed 3)
    (X, y, idx, linfo), tcol = _timed(lambda: S.collect(nq))
    params, info = S.train(X, y, n_train=200000, epochs=30, seed=3, device=dev)
    lmax = S.level_caps(y, linfo)

    def plain():
        w = nq.run()
        plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                               P["srcover"], P["whtsover"])
        t0 = time.perf_counter()
        x, nit, _, rres, _ = layer.gmres(plan, rhs, 0.5, wts, numit=300, eps_gmres=RTOL)
        torch.cuda.synchronize()
        ts = time.perf_counter() - t0
        plan.close()
        return x, int(nit[0]), float(rres[0]), ts

    def learned():
        w = nq.run(selector=params, route_lmax=lmax)
        plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                               P["srcover"], P["whtsover"])
        # one right-hand side: the default far-sum shape (tuning costs more than it saves
        # in a single solve)
        t0 = time.perf_counter()
        x, outer, inner, rr, n32, n64 = layer.gmres_ir(plan, rhs, 0.5, wts, numit=300, eps_inner=RTOL / 2,
                                                       tol=RTOL)
        torch.cuda.synchronize()
        ts = time.perf_counter() - t0
        plan.close()
        return x, outer, inner, float(rr[0]), ts

    def converged():
        """FP64 solve of the plain operator to RTOL_REF (untimed comparator)."""
        w = nq.run()
        plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                               P["srcover"], P["whtsover"])
        x, nit, _, rres, _ = layer.gmres(plan, rhs, 0.5, wts, numit=300, eps_gmres=RTOL_REF)
        plan.close()
        return x.cpu().numpy(), float(rres[0]), int(nit[0])

    plain()
    (x64, nit, rr64, ts64), tt64 = _timed(plain)
    learned()
    (xl, outer, inner, rrl, tsl), ttl = _timed(learned)
    # per-kernel times of each kernel alone (every kernel in order on one stream)
    layer.set_side_stream(False)
    nq.run()
    q64 = {k: nq.stats[k] for k in ("ms_tree", "ms_adap", "ms_fixed", "ms_self")}
    nq.run(selector=params, route_lmax=lmax)
    ql = {k: nq.stats[k] for k in ("ms_select", "ms_tree", "ms_tree_routed", "ms_adap", "ms_fixed", "ms_self",
                                   "nfallback", "ncls")}
    layer.set_side_stream(True)
    targ = np.array([[4.0, 0.0, 1.0], [0.0, 4.0, -1.0], [-3.5, 1.0, 2.0], [0.5, -3.8, 0.5], [0.0, 0.0, 3.0],
                     [2.0, 2.0, 2.5], [-2.0, 2.5, -2.0], [3.0, -3.0, 0.0], [0.0, 0.0, -3.5], [-4.0, -1.0, 0.0]])
    exact = problem.point_source(targ, x0, zk)
    e64 = float(np.abs(_exterior_field(g, P["wts"], x64.cpu().numpy(), targ, pars, zk) - exact).max()
                / np.abs(exact).max())
    el = float(np.abs(_exterior_field(g, P["wts"], xl.cpu().numpy(), targ, pars, zk) - exact).max()
               / np.abs(exact).max())
    xc, rrc, itc = converged()
    return {"workload": "helmholtz_zk3_wavytorus_48x16_o6_pointsource", "npts": len(g["srcvals"]), "nnz": P["nnz"],
            "rtol": RTOL, "reference_solve": {"rres": rrc, "iters": itc},
            "plain_fp64": {"setup_plus_solve_s": tt64, "solve_s": ts64, "iters": nit, "rres": rr64,
                           "quad_ms": q64, "field_err": e64,
                           "err_vs_converged_fp64": _wrel(x64.cpu().numpy(), xc, P["wts"])},
            "learned_mixed": {"setup_plus_solve_s": ttl, "solve_s": tsl, "ir_outer": outer, "ir_inner_iters": inner,
                              "rres": rrl, "quad_ms": ql, "field_err": el,
                              "err_vs_converged_fp64": _wrel(xl.cpu().numpy(), xc, P["wts"])},
            "speedup_solve": ts64 / tsl, "speedup_setup_plus_solve": tt64 / ttl,
            "adaptive_part_ms": {"plain": q64["ms_tree"] + q64["ms_adap"],
                                 "learned": ql["ms_select"] + ql["ms_tree"] + ql["ms_tree_routed"] + ql["ms_adap"]},
            "selector": {"train_pairs": info["n_train"], "heldout_pairs": info["n_heldout"],
                         "acc_train": info["acc_train"], "acc_heldout": info["acc_heldout"],
                         "train_ms": info["ms_train"], "loss_first_last": [info["loss"][0], info["loss"][-1]]}}


def batched_solve(dev, nseq):
    """Sphere 12x12 per face, order 4, zk=5, 64 plane waves (directions from seed 2),
    sound-soft data f = -u_inc: 64 batched systems with the autotuned FP32 far sum and
    FP64 iterative refinement against FP64 GMRES one right-hand side at a time (timed on
    nseq of them and scaled to 64)."""
    import torch
    from fmm3dbie_amd import geometry as G, layer, problem
    zk = 5.0
    pars = np.array([zk, 1j * zk, 1.0])
    g = G.get_sphere_npat(1.0, 12, (0, 0, 0), 4)
    P = problem.setup(g, 1e-6, True, pars)
    dirs = problem.plane_wave_dirs(64, 2)
    rhs = torch.as_tensor(-problem.plane_waves(g["srcvals"][:, :3], dirs, zk), device=dev)
    wts = torch.as_tensor(P["wts"], device=dev)
    nq = layer.NearQuadDevice(g, True, pars, 1e-6, P["row_ptr"], P["col_ind"], P["iquad"], P["rfac0"], device=dev)
    w = nq.run()
    plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], w, P["novers"], P["ixyzso"],
                           P["srcover"], P["whtsover"])
    layer.gmres(plan, rhs[0], 0.5, wts, numit=300, eps_gmres=RTOL)
    xs, iters = [], []
    _, tseq = _timed(lambda: [xs.append(layer.gmres(plan, rhs[r], 0.5, wts, numit=300, eps_gmres=RTOL))
                              for r in range(nseq)])
    iters = [int(x[1][0]) for x in xs]
    x64 = torch.stack([x[0] for x in xs]).cpu().numpy()
    tune, ttune = _timed(lambda: plan.tune(32, 64))
    best = min(tune, key=lambda t: t[2])
    layer.gmres_ir(plan, rhs, 0.5, wts, numit=300, eps_inner=RTOL / 2, tol=RTOL)
    (xb, outer, inner, rr, n32, n64), tb = _timed(
        lambda: layer.gmres_ir(plan, rhs, 0.5, wts, numit=300, eps_inner=RTOL / 2, tol=RTOL))
    xb = xb.cpu().numpy()
    # converged FP64 comparator on the sequentially solved right-hand sides (untimed)
    xc, _, _, rrc, _ = layer.gmres(plan, rhs[:nseq].contiguous(), 0.5, wts, numit=300, eps_gmres=RTOL_REF)
    xc = xc.cpu().numpy()
    targ = 3.0 * problem.plane_wave_dirs(10, 0)
    exact = problem.mie_sound_soft(targ, dirs[:8], zk)
    u = _exterior_field(g, P["wts"], xb[:8], targ, pars, zk)
    emie = float(np.abs(u - exact).max() / np.abs(exact).max())
    tm = np.zeros(4)
    plan.apply(rhs.to(torch.complex64), 32, 1, times=tm)
    plan.close()
    return {"workload": "helmholtz_zk5_sphere_12x12_o4_64planewaves", "npts": len(g["srcvals"]), "nrhs": 64,
            "fp64_sequential_s_per_rhs": tseq / nseq, "fp64_sequential_s_64_scaled": tseq / nseq * 64,
            "fp64_sequential_timed_rhs": nseq, "fp64_iters": iters,
            "batched_mixed_s": tb, "ir_outer": outer, "ir_inner_iters": inner, "ir_rres_max": float(np.max(rr)),
            "speedup_vs_fp64_sequential": tseq / nseq * 64 / tb, "rtol": RTOL,
            "reference_solve": {"rhs": nseq, "rres_max": float(np.max(rrc))},
            "err_vs_converged_fp64_max": {"plain_fp64": _wrel(x64, xc, P["wts"]),
                                          "batched_mixed": _wrel(xb[:nseq], xc, P["wts"])},
            "mie_field_err_8rhs": emie, "tuned_far": {"group": best[0], "wg_target": best[1], "ms": best[2]},
            "tune_s": ttune, "apply_fp32_64rhs_ms": {"oversample": tm[0], "far": tm[1], "reduce_csr": tm[2]}}


# CPU share of one GPU on the benchmark machines (one process, one GPU): 16 cores
CPU_SHARE = 16


def _host_cores():
    """(cores used, nproc, CPU model): the process's affinity mask at start-up, capped by a
    cgroup CPU quota when one is set (cpu.max) and by the 16-core CPU share of one GPU."""
    nproc = len(_AFFINITY0) if _AFFINITY0 is not None else (os.cpu_count() or 1)
    usable = min(nproc, CPU_SHARE)
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            usable = max(1, min(nproc, int(int(q) // int(per))))
    except (OSError, ValueError):
        pass
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return usable, nproc, model


def cpu_baseline(g, helm, pars, eps, P, budget_s):
    """The C oracle (FP64 restatement of the reference algorithm, OpenMP over target rows,
    threads bound close to their cores) on every core this process may use, over a
    bounded sample of evenly spaced target rows of the same workload; rate in near-quad
    pairs/s (pairs of the sample / time)."""
    usable, nproc, model = _host_cores()
    # the oracle's OpenMP runtime reads its places when it is loaded: give it the start-up
    # affinity (another runtime in the process may have bound the main thread to one core)
    # and keep its threads next to their cores
    if _AFFINITY0 is not None:
        os.sched_setaffinity(0, _AFFINITY0)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    from oracle import oracle as O
    from fmm3dbie_amd.layer import SELF_NR, SELF_NT, SELF_TAU
    threads = O.set_num_threads(usable)
    N = len(g["srcvals"])
    row_ptr = P["row_ptr"]
    kind = O.HELM_COMB if helm else O.LAP_COMB
    take = max(64, N // 512)
    rate = None
    used = 0
    for _ in range(8):
        rows = np.linspace(0, N - 1, take).astype(np.int64)
        sub_ptr = np.concatenate([[1], 1 + np.cumsum(np.diff(row_ptr)[rows])]).astype(np.int64)
        cols = np.concatenate([P["col_ind"][row_ptr[r] - 1: row_ptr[r + 1] - 1] for r in rows])
        iq = np.ones(len(cols) + 1, np.int64)
        iq[1:] += np.cumsum(np.diff(g["ixyzs"])[cols - 1])
        t0 = time.perf_counter()
        O.getnearquad(g, g["srcvals"][rows], P["ipatch_id"][rows], P["uvs_targ"][rows], eps, kind, pars, sub_ptr,
                      cols, iq, P["rfac0"], O.SELF_POLAR, SELF_NR, SELF_NT, self_tau=SELF_TAU)
        dt = time.perf_counter() - t0
        rate = len(cols) / dt
        used = len(cols)
        if dt > budget_s / 3 or take >= N:
            break
        take = min(N, int(take * max(2.0, budget_s / 2 / max(dt, 1e-3))))
    return {"value": rate, "unit": "pairs/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": model, "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
            "sample": "%d near pairs of %d evenly spaced target rows (of %d) of the same workload, C oracle "
                      "(same self, adaptive and fixed rules as the GPU) on %d threads (%d CPUs visible, CPU share "
                      "of one GPU %d)" % (used, take, N, threads, nproc, CPU_SHARE)}


def combined_solve(g, pars, P, nq, dev, t_quad, nfp):
    """The 344k-point configuration (deformed sphere 32x32 cells per face, order 6, zk =
    10) with 16 plane waves (directions from seed 4, sound-soft data f = -u_inc):

    * plain FP64 GPU form: the plain adaptive quadrature (the benchmark step) and FP64
      GMRES one right-hand side at a time to a relative residual of RTOL, timed on nfp
      right-hand sides and scaled to 16;
    * learned + mixed + batched form: the learned selector trained on this surface's own
      adaptive pairs (training time reported apart), the routed quadrature, and
      FP32-inner / FP64-refined GMRES over all 16 right-hand sides in lockstep (one far-sum
      launch per step for 8 right-hand sides at a time) to the same RTOL;
    * a converged FP64 solve (RTOL_REF, untimed) of the first nfp right-hand sides: the
      error of both forms against it."""
    import torch
    from fmm3dbie_amd import layer, problem, selector as S
    layer.set_gmres_print(10)  # progress on stderr: the solves take minutes
    zk = float(np.real(pars[0]))
    dirs = problem.plane_wave_dirs(16, 4)
    rhs = torch.as_tensor(-problem.plane_waves(g["srcvals"][:, :3], dirs, zk), device=dev)
    wts = torch.as_tensor(P["wts"], device=dev)
    w_plain = nq.run().clone()
    plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], w_plain, P["novers"],
                           P["ixyzso"], P["srcover"], P["whtsover"])
    plan.apply(rhs[0], 64, 0)
    xs = []

    def one(r):
        _progress("combined: plain FP64 GMRES, right-hand side %d of %d" % (r + 1, nfp))
        return layer.gmres(plan, rhs[r], 0.5, wts, numit=300, eps_gmres=RTOL)

    _, tseq = _timed(lambda: [xs.append(one(r)) for r in range(nfp)])
    x64 = torch.stack([x[0] for x in xs]).cpu().numpy()
    it64 = [int(x[1][0]) for x in xs]
    _progress("combined: converged FP64 comparator")
    xc, itc, _, rrc, _ = layer.gmres(plan, rhs[:nfp].contiguous(), 0.5, wts, numit=400, eps_gmres=RTOL_REF)
    xc = xc.cpu().numpy()
    plan.close()
    del w_plain
    # learned + mixed + batched
    _progress("combined: labels and training")
    (X, y, idx, linfo), tcol = _timed(lambda: S.collect(nq))
    params, info = S.train(X, y, n_train=200000, epochs=30, seed=3, device=dev)
    lmax = S.level_caps(y, linfo)
    nq.run(selector=params, route_lmax=lmax)
    _, tq_l = _timed(lambda: nq.run(selector=params, route_lmax=lmax))
    st = dict(nq.stats)
    plan = layer.ApplyPlan(g, True, pars, P["row_ptr"], P["col_ind"], P["iquad"], nq.wnear, P["novers"],
                           P["ixyzso"], P["srcover"], P["whtsover"])
    # the default far-sum shape (8 right-hand sides per group, 32768-target workgroups) is
    # within 1% of the tuned one at this size, and tuning would time 20 shapes of 0.7-1.9 s
    _progress("combined: mixed-precision GMRES of 16 right-hand sides")
    ttune = 0.0
    plan.apply(rhs.to(torch.complex64), 32, 1)
    (xb, outer, inner, rr, n32, n64), tb = _timed(
        lambda: layer.gmres_ir(plan, rhs, 0.5, wts, numit=300, eps_inner=RTOL / 2, tol=RTOL))
    xb = xb.cpu().numpy()
    plan.close()
    plain_total = t_quad + tseq / nfp * 16
    learned_total = tq_l + ttune + tb
    return {"workload": "helmholtz_zk10_deformedsphere_32x32_o6_16planewaves", "npts": len(g["srcvals"]),
            "nrhs": 16, "rtol": RTOL,
            "plain_fp64": {"quad_s": t_quad, "solve_s_per_rhs": tseq / nfp, "solve_s_16_scaled": tseq / nfp * 16,
                           "timed_rhs": nfp, "iters": it64, "setup_plus_solve_s": plain_total,
                           "err_vs_converged_fp64_max": _wrel(x64, xc, P["wts"])},
            "learned_mixed_batched": {"quad_s": tq_l, "tune_s": ttune, "solve_s": tb, "ir_outer": outer,
                                      "ir_inner_iters": inner, "ir_fp32_applies": n32, "ir_fp64_applies": n64,
                                      "rres_max": float(np.max(rr)), "setup_plus_solve_s": learned_total,
                                      "err_vs_converged_fp64_max": _wrel(xb[:nfp], xc, P["wts"]),
                                      "quad_stats": {k: st[k] for k in ("ms_select", "ms_tree", "ms_tree_routed",
                                                                        "ms_adap", "ms_fixed", "ms_self")},
                                      "selector": {"label_collection_s": tcol, "train_ms": info["ms_train"],
                                                   "acc_heldout": info["acc_heldout"],
                                                   "train_pairs": info["n_train"]}},
            "reference_solve": {"rhs": nfp, "iters": [int(v) for v in itc], "rres_max": float(np.max(rrc))},
            "speedup_solve": (tseq / nfp * 16) / tb, "speedup_setup_plus_solve": plain_total / learned_total}

if __name__ == "__main__":
    main()
