"""Benchmark: MG-PCG time to a 1e-8 relative residual and CSR SpMV HBM GB/s on
the 3D Q1 Poisson problem, level 7 (128^3 cells, 2,146,689 dofs, 57,066,625
nonzeros), on one MI355X.

One step = one complete PCG(MultiGrid-V(4,4), Richardson(Jacobi 0.5)) solve of
A x = b from x = 0 to |r| <= 1e-8 |r0| (coarse level 1, exact coarse solve),
inputs resident in HBM.  `value` is the time to solution in seconds (lower is
better).  The roofline object is for the SpMV kernel r = A*x in the format the
solver runs on (the pair-dictionary SELL-64 plan of the matrix), timed with HIP
events on the stream it is launched on; the plain SELL-64 and CSR SpMVs and the
solve on plain SELL-64 plans are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DIM, LF, LC = 3, 7, 1
SEED = 28054777172512
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the level-7 kernels, from two hardware-counter runs of
# `bench.py --spmv-only` (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch); FETCH_SIZE is scaled
# by 2.0 and WRITE_SIZE by 1.0, the factors that make the same counters read 1 GiB for a
# 1 GiB device-to-device copy (profiles/spmv_traffic_3d_q1_l7.md)
SPMV_TRAFFIC_BYTES = {"jacobi_step": 55483.9 * 2048 + 16771.0 * 1024,
                      "sell64_diag": 38712.2 * 2048 + 16771.0 * 1024,
                      "sell64_dict": 40575.7 * 2048 + 16771.0 * 1024,
                      "sell64": 295808.9 * 2048 + 16771.0 * 1024,
                      "csr_u64": 474623.3 * 2048 + 16771.0 * 1024}


def load_feat():
    spec = importlib.util.spec_from_file_location(
        "feat3amd", os.path.join(ROOT, "feat3.doxytest_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["feat3amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, idx_bytes=8, val_bytes=8):
    """Algorithmic bytes of r = A*x: val + col + row_ptr + x + r, each once."""
    return nnz * (val_bytes + idx_bytes) + (n + 1) * idx_bytes + n * val_bytes + n * val_bytes


def cpu_baseline(feat, b, iters_gpu):
    """The reference's own generic code (oracle/_ref) on one host core, on the
    same matrices; falls back to the C oracle where _ref is absent."""
    from oracle.pyoracle import Oracle, Reference, LIB_REF, Csr, csr_transpose, COARSE_JACOBI, COARSE_DIRECT
    levels = []
    for lvl in range(LF, LC - 1, -1):
        q = feat.q1_problem(DIM, lvl)
        lv = dict(A=Csr(q["n"], q["n"], q["val"], q["col"], q["rowptr"]), filt=q["filt"])
        if lvl > LC:
            P = Csr(q["n"], q["pcols"], q["pval"], q["pcol"], q["prowptr"])
            lv["P"], lv["R"] = P, csr_transpose(P)
        levels.append(lv)
    n = levels[0]["A"].rows
    if os.path.exists(LIB_REF):
        kind = "reference"
        t0 = time.perf_counter()
        r = Reference().mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_JACOBI)
        secs = r["seconds"] if r.get("seconds") else time.perf_counter() - t0
    else:
        kind = "port"
        t0 = time.perf_counter()
        r = Oracle().mg_pcg_solve(levels, np.zeros(n), b, coarse=COARSE_DIRECT)
        secs = time.perf_counter() - t0
    return dict(value=secs, unit="s", cores=1, kind=kind,
                sample=f"one full solve of the same workload ({r['iters']} PCG iterations, "
                       f"GPU took {iters_gpu}), single thread, {platform.processor() or 'x86_64'} "
                       f"host of {os.cpu_count()} logical CPUs",
                iters=r["iters"])


STOKES_LF, STOKES_LC = 6, 1


def stokes_cpu_reference(feat, x0, b):
    """The reference's AmaVanka multigrid (oracle/_ref) on one host core, same matrices."""
    from oracle.pyoracle import Reference, LIB_REF, Bcsr, Csr, csr_transpose
    if not os.path.exists(LIB_REF):
        return None
    levels = []
    for lvl in range(STOKES_LF, STOKES_LC - 1, -1):
        s = feat.stokes_problem(lvl)
        nv, np_ = s["nv"], s["np"]
        lv = dict(nv=nv, np=np_, filt=s["filt"], filt_val=s["filt_val"], prim=s["prim"], dual=s["dual"],
                  A=Bcsr(nv, nv, 2, 2, *s["A"]), B=Bcsr(nv, np_, 2, 1, *s["B"]), D=Bcsr(np_, nv, 1, 2, *s["D"]))
        if lvl > STOKES_LC:
            pv, pp = s["Pv"], s["Pp"]
            lv["Pv"], lv["Pp"] = Csr(nv, pv[3], pv[0], pv[1], pv[2]), Csr(np_, pp[3], pp[0], pp[1], pp[2])
            lv["Rv"], lv["Rp"] = csr_transpose(lv["Pv"]), csr_transpose(lv["Pp"])
        levels.append(lv)
    r = Reference().stokes_mg_solve(levels, x0, b, tol_rel=1e-8)
    return dict(value=r["seconds"], unit="s", cores=1, kind="reference", iters=r["iters"])


def stokes_leg(feat, torch, dev, args):
    t0 = time.perf_counter()
    st = feat.Stokes(STOKES_LF, STOKES_LC, steps=5, omega_rich=0.7, tol_rel=1e-8, max_iter=100)
    setup = time.perf_counter() - t0
    z = st.sizes(0)
    n = 2 * z["nv"] + z["np"]
    b = torch.zeros(n, dtype=torch.float64, device=dev)
    st.filter(0, b, "rhs")
    x0 = torch.zeros(n, dtype=torch.float64, device=dev)
    st.filter(0, x0, "sol")
    x = torch.empty_like(x0)
    res, times = None, []
    for k in range(args.warmup + args.steps):
        x.copy_(x0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = st.solve(x, b)
        torch.cuda.synchronize()
        if k >= args.warmup:
            times.append(time.perf_counter() - t)
    assert res["status"] == "success", res
    t_init = float(np.median([st.init_numeric() for _ in range(3)]))
    out = {"workload": f"2D Stokes lid-driven cavity Q2/P1dc level {STOKES_LF} (64x64 quads, {2 * z['nv']} velocity "
                       f"+ {z['np']} pressure dofs), Richardson(MG-V(Richardson 0.7 x5, Schwarz(AmaVanka))), "
                       f"levels {STOKES_LF}..{STOKES_LC}, tol 1e-8",
           "s_per_solve": float(np.mean(times)), "iters": res["iters"],
           "final_rel_defect": res["def_final"] / res["def_init"], "vanka_macros": z["macros"],
           "vanka_local_size": z["stride"], "vanka_init_numeric_s_all_levels": t_init,
           "setup_s": setup}
    log(f"stokes L{STOKES_LF}: {out['s_per_solve'] * 1e3:.2f} ms per solve, {res['iters']} iterations, "
        f"rel. defect {out['final_rel_defect']:.3e}, AmaVanka init {t_init * 1e3:.2f} ms")
    if not args.no_cpu_baseline:
        cb = stokes_cpu_reference(feat, x0.cpu().numpy(), b.cpu().numpy())
        if cb:
            out["cpu_baseline"] = cb
            log(f"stokes cpu reference: {cb['value']:.3f} s, {cb['iters']} iterations")
    st.close()
    return out


BATCH = dict(seed=28054777172512, ninst=64, lmin=3, lmax=8, sigma=1.0)


def batch_leg(feat, torch, dev, args):
    """64 independent 2D Q1 diffusion MG-PCG solves (levels 3..8, log-normal
    coefficients): the batched solver vs the same solves one after another."""
    t0 = time.perf_counter()
    bs = feat.BatchSolver(BATCH["seed"], BATCH["ninst"], BATCH["lmin"], BATCH["lmax"], BATCH["sigma"], 1e-8, 100)
    setup = time.perf_counter() - t0
    b_host = feat.batch_rhs(bs, BATCH["seed"])
    b = torch.from_numpy(b_host).to(dev)
    x = torch.zeros(bs.n, dtype=torch.float64, device=dev)
    times, res = [], None
    for k in range(args.warmup + args.steps):
        x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = bs.solve(x, b)
        torch.cuda.synchronize()
        if k >= args.warmup:
            times.append(time.perf_counter() - t)
    assert all(s == "success" for s in res["status"]), res["status"]
    t_batch = float(np.mean(times))
    tunes, us = bs.tuning()
    # the same 64 solves one after another with the single-instance solver
    cfg = feat.mgpcg_cfg(coarse="jacobi", coarse_omega=1.0, tol_rel=1e-8, max_iter=100)
    singles = []
    for i, L in enumerate(bs.levels):
        coef = feat.lognormal_cells(BATCH["seed"] + 1000 + i, 4 ** int(L), BATCH["sigma"])
        lo, hi = int(bs.off[i]), int(bs.off[i + 1])
        singles.append((feat.MgPcg(cfg, levels=feat.batch_instance_levels(int(L), coef)), lo, hi))
    xs = torch.zeros(bs.n, dtype=torch.float64, device=dev)
    seq = []
    for k in range(args.warmup + args.steps):
        xs.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for s, lo, hi in singles:
            s.solve(xs[lo:hi], b[lo:hi])
        torch.cuda.synchronize()
        if k >= args.warmup:
            seq.append(time.perf_counter() - t)
    for s, _, _ in singles:
        s.close()
    t_seq = float(np.mean(seq))
    err = float(torch.linalg.norm(x - xs) / torch.linalg.norm(xs))
    out = {"workload": "64 independent 2D Q1 diffusion problems, levels 3..8 from the seed "
                       f"({bs.n} dofs in total), log-normal(0,1) cell coefficients, PCG + MG V(4,4) "
                       "Richardson(Jacobi 0.5), coarse level 1 Jacobi, tol 1e-8 each",
           "s_per_batch": t_batch, "s_sequential_single_solves": t_seq, "speedup_vs_sequential": t_seq / t_batch,
           "iters_min": int(res["iters"].min()), "iters_max": int(res["iters"].max()),
           "rel_diff_vs_sequential": err, "levels": [int(v) for v in bs.levels], "setup_s": setup,
           "coarse_tail_levels": bs.tail_levels,
           "autotuned_shapes": [int(t) for t in tunes], "autotuned_spmv_us": [float(u) for u in us]}
    log(f"batch of 64: {t_batch * 1e3:.2f} ms (sequential single solves {t_seq * 1e3:.2f} ms, "
        f"{t_seq / t_batch:.2f}x), iterations {out['iters_min']}..{out['iters_max']}, |x-x_seq|/|x_seq| {err:.2e}")
    if not args.no_cpu_baseline:
        from oracle.pyoracle import Reference, LIB_REF, Csr, csr_transpose, COARSE_JACOBI
        if os.path.exists(LIB_REF):
            ref = Reference()
            tot = 0.0
            for i, L in enumerate(bs.levels):
                coef = feat.lognormal_cells(BATCH["seed"] + 1000 + i, 4 ** int(L), BATCH["sigma"])
                lvls = []
                for q in feat.batch_instance_levels(int(L), coef):
                    lv = dict(A=Csr(q["n"], q["n"], q["val"], q["col"], q["rowptr"]), filt=q["filt"])
                    if q.get("pcols"):
                        P = Csr(q["n"], q["pcols"], q["pval"], q["pcol"], q["prowptr"])
                        lv["P"], lv["R"] = P, csr_transpose(P)
                    lvls.append(lv)
                lo, hi = int(bs.off[i]), int(bs.off[i + 1])
                r = ref.mg_pcg_solve(lvls, np.zeros(hi - lo), b_host[lo:hi], coarse=COARSE_JACOBI)
                tot += r["seconds"]
            out["cpu_baseline"] = {"value": tot, "unit": "s", "cores": 1, "kind": "reference",
                                   "sample": "the 64 solves one after another, reference generic code"}
            log(f"batch cpu reference: {tot:.3f} s")
    bs.close()
    return out


def assembly_leg(feat, torch, dev, args):
    """Device assembly of the structured Q1 level (stiffness, boundary list,
    prolongation, restriction) against the host assembly of the same arrays, and
    the MG-PCG setup with device- vs host-assembled levels."""
    k, _ = feat.libs()
    stream = torch.cuda.current_stream()
    h = feat.Handle(0, stream=stream.cuda_stream)
    d = feat.q1_problem_device(torch, h, DIM, LF)
    names = ("val", "col", "rowptr", "filt", "pval", "pcol", "prowptr", "rval", "rcol", "rrowptr")
    ptrs = [d[nm].data_ptr() for nm in names]

    def launch():
        feat.check(k.feat_q1_assemble_device(h.h, DIM, LF, *ptrs), "feat_q1_assemble_device")

    for _ in range(3):
        launch()
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        launch()
    e1.record(stream)
    e1.synchronize()
    dev_ms = e0.elapsed_time(e1) / reps
    # write-only: 16 B per entry of A, P and R, the three row-pointer arrays, the boundary list
    wbytes = sum(d[nm].numel() * d[nm].element_size() for nm in names)
    t = time.perf_counter()
    q = feat.q1_problem(DIM, LF)
    host_s = time.perf_counter() - t
    same = (np.array_equal(d["val"].cpu().numpy().view(np.uint64), q["val"].view(np.uint64))
            and np.array_equal(d["col"].cpu().numpy().view(np.uint64), q["col"])
            and np.array_equal(d["pval"].cpu().numpy().view(np.uint64), q["pval"].view(np.uint64))
            and np.array_equal(d["filt"].cpu().numpy().view(np.uint64), q["filt"]))
    del d, q
    h.close()
    torch.cuda.synchronize()
    cfg = feat.mgpcg_cfg(nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0, coarse="direct",
                         tol_rel=1e-8, max_iter=100)
    setup = {}
    for host in (False, True):
        torch.cuda.synchronize()
        t = time.perf_counter()
        s = feat.MgPcg(cfg, q1=(DIM, LF, LC), host_assembly=host)
        torch.cuda.synchronize()
        setup["host" if host else "device"] = time.perf_counter() - t
        s.close()
    # ---- Q2/P1dc Stokes level (A, B, D, filters, both prolongations and restrictions)
    hs = feat.Handle(0, stream=stream.cuda_stream)
    stokes = {}
    for lvl in (STOKES_LF, 9):
        sd = feat.stokes_problem_device(torch, hs, lvl)
        sa = feat.StokesLevelDev()
        for key, pre in (("A", "a"), ("B", "b"), ("D", "d"), ("Pv", "pv"), ("Pp", "pp"), ("Rv", "rv"), ("Rp", "rp")):
            for nm, tt in zip(("val", "col", "ptr"), sd[key][:3]):
                setattr(sa, f"{pre}_{nm}", tt.data_ptr())
        for nm in ("filt", "filt_val", "prim", "dual"):
            setattr(sa, nm, sd[nm].data_ptr())
        for _ in range(3):
            feat.check(k.feat_stokes_assemble_device(hs.h, lvl, C.byref(sa)), "feat_stokes_assemble_device")
        e0.record(stream)
        for _ in range(reps):
            feat.check(k.feat_stokes_assemble_device(hs.h, lvl, C.byref(sa)), "feat_stokes_assemble_device")
        e1.record(stream)
        e1.synchronize()
        sms = e0.elapsed_time(e1) / reps
        sbytes = sum(tt.numel() * tt.element_size() for key in ("A", "B", "D", "Pv", "Pp", "Rv", "Rp")
                     for tt in sd[key][:3]) + sum(sd[nm].numel() * 8 for nm in ("filt", "filt_val", "prim", "dual"))
        t = time.perf_counter()
        feat.stokes_problem(lvl)
        shost = time.perf_counter() - t
        stokes[f"level_{lvl}"] = {"device_ms": sms, "bytes_written": sbytes, "gbps_written": sbytes / (sms * 1e-3) / 1e9,
                                  "host_s": shost}
        log(f"stokes assembly level {lvl}: device {sms * 1e3:.1f} us ({sbytes / (sms * 1e-3) / 1e9:.0f} GB/s written), "
            f"host {shost * 1e3:.1f} ms")
        del sd
    hs.close()
    log(f"assembly level {LF}: device {dev_ms:.2f} ms ({wbytes / (dev_ms * 1e-3) / 1e9:.0f} GB/s written), "
        f"host {host_s:.2f} s, identical {same}; MG-PCG setup device {setup['device']:.2f} s, host {setup['host']:.2f} s")
    return {"what": f"Q1 level {LF}: stiffness (CSR f64/u64), boundary list, prolongation from level {LF - 1} "
                    "and restriction (its transpose)",
            "device_ms": dev_ms, "bytes_written": wbytes, "gbps_written": wbytes / (dev_ms * 1e-3) / 1e9,
            "host_s": host_s, "host_what": "host assembly of stiffness, boundary and prolongation (no transpose), "
                                          "one core",
            "bitwise_identical": bool(same),
            "mgpcg_setup_s": {"device_assembly": setup["device"], "host_assembly_and_upload": setup["host"],
                              "includes": "levels 7..1, SELL-64 dictionary plans, autotuning, coarse factorisation"},
            "stokes_q2p1dc": dict(stokes, what="A (2x2 blocks), B, D, velocity filter, mean-filter vectors, Q2 and "
                                               "P1dc prolongations and restrictions; host: the same without the "
                                               "restrictions, one core")}


def smoothers_leg(feat, torch, dev, args):
    """SOR / SSOR (level-set and multicolour order) and ILU(0) on the level-7 Q1
    matrix: one application each (HIP events), the ILU(0) factorisation, and PCG
    around SSOR / ILU(0) to 1e-8 on level 5 (3D)."""
    k, _ = feat.libs()
    stream = torch.cuda.current_stream()
    h = feat.Handle(0, stream=stream.cuda_stream)
    d = feat.q1_problem_device(torch, h, DIM, LF, transfer=False)
    n, nnz = d["n"], int(d["val"].numel())
    desc = feat.CsrF64(n, n, nnz, d["val"].data_ptr(), d["col"].data_ptr(), d["rowptr"].data_ptr())
    x = torch.rand(n, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {}
    t = time.perf_counter()
    lv = feat.SorPlan(h, desc)
    setup_lv = time.perf_counter() - t
    t = time.perf_counter()
    mc = feat.SorPlan(h, desc, multicolor=True)
    setup_mc = time.perf_counter() - t
    # bytes of one sweep: the entries it reads (8 B value + 4 B column), diagonal, x and y
    lower = (nnz - n) // 2
    sweep_bytes = lower * 12 + n * 8 * 3
    for name, plan in (("level_sets", lv), ("multicolour", mc)):
        ms_sor = timed(lambda: plan.sor(y, x, 1.0), 5)
        ms_ssor = timed(lambda: plan.ssor(y, x, 1.0), 5)
        out[f"sor_{name}"] = {"colours": plan.ncolors, "ms_per_apply": ms_sor,
                              "gbps_sweep_bytes": sweep_bytes / (ms_sor * 1e-3) / 1e9}
        out[f"ssor_{name}"] = {"ms_per_apply": ms_ssor, "gbps_sweep_bytes": 2 * sweep_bytes / (ms_ssor * 1e-3) / 1e9}
        log(f"{name}: {plan.ncolors} colours, SOR {ms_sor:.3f} ms, SSOR {ms_ssor:.3f} ms per application")
    out["sor_level_sets"]["setup_s"] = setup_lv
    out["sor_multicolour"]["setup_s"] = setup_mc
    lv.close()
    mc.close()
    t = time.perf_counter()
    ilu = feat.IluPlan(h, desc)
    torch.cuda.synchronize()
    setup_ilu = time.perf_counter() - t
    ms_fact = timed(ilu.init_numeric, 3)
    ms_ilu = timed(lambda: ilu.apply(y, x), 5)
    out["ilu0"] = {"levels_l": ilu.levels_l, "levels_u": ilu.levels_u, "ms_factorise": ms_fact, "ms_per_apply": ms_ilu,
                   "setup_s": setup_ilu}
    log(f"ilu(0): {ilu.levels_l}+{ilu.levels_u} levels, factorise {ms_fact:.2f} ms, apply {ms_ilu:.2f} ms")
    ilu.close()
    del d
    h.close()
    # PCG to 1e-8 on 3D Q1 level 5 around the device preconditioners
    q = feat.q1_problem(3, 5)
    from types import SimpleNamespace
    A = SimpleNamespace(rows=q["n"], cols=q["n"], val=q["val"], col=q["col"], rowptr=q["rowptr"])
    b = feat.rhs_random(SEED, q["n"], -1.0, 1.0)
    b[q["filt"].astype(np.int64)] = 0.0
    bd = torch.from_numpy(b).to(dev)
    # (ILU: omega_pre is the fill-in level p)
    for pre, om, tag in (("jacobi", 1.0, "jacobi"), ("ssor", 1.0, "ssor"), ("ilu", 0.0, "ilu"), ("ilu", 1.0, "ilu1"),
                         ("ilu", 2.0, "ilu2")):
        sv = feat.KrylovSolver("pcg", A, q["filt"], pre, omega_pre=om, tol_rel=1e-8, max_iter=1000)
        xs = torch.zeros(q["n"], dtype=torch.float64, device=dev)
        sv.solve(xs, bd)
        xs.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = sv.solve(xs, bd)
        torch.cuda.synchronize()
        out[f"pcg_{tag}_3d_l5"] = {"s_per_solve": time.perf_counter() - t, "iters": r["iters"],
                                   "final_rel_defect": r["def_final"] / r["def_init"]}
        log(f"pcg-{tag} 3D L5: {r['iters']} iterations, {out[f'pcg_{tag}_3d_l5']['s_per_solve'] * 1e3:.1f} ms")
        sv.close()
    out["what"] = (f"3D Q1 level {LF} matrix ({n} rows): one SOR / SSOR / ILU(0) application, level-set order = the "
                   "generic sequential sweeps bit for bit, multicolour = the CUDA backend's colour order; PCG on level 5")
    return out


def mixed_varcoef_leg(feat, torch, dev, args):
    """Mixed precision where the matrices are the traffic: the 3D Q1 anisotropic diffusion
    problem of the learned-smoother leg (log-normal sigma 1.5, K = diag(1, 1, 10)) at level 7
    (2,146,689 dofs).  Variable coefficients leave no pair dictionary: every level runs on
    banded plans (the grid-stencil rows: values only, columns from the node position; FP64
    8 B, FP32 4 B, FP16 2 B per entry; the plain SELL-64 plan adds 2 B column offsets).
    PCG + MG V(4,4) Richardson(Jacobi 0.5), coarse level 1, tol 1e-8 on the FP64 defect."""
    import scipy.sparse as sp
    t0 = time.perf_counter()
    lv = feat.diffusion3d_levels(LF, LC, SEED + 1000, 1.5, 10.0)
    n = lv[0]["n"]
    bh = feat.rhs_random(SEED + 2000, n, -1.0, 1.0)
    bh[lv[0]["filt"].astype(np.int64)] = 0.0
    b = torch.from_numpy(bh).to(dev)
    S = sp.csr_matrix((lv[0]["val"], lv[0]["col"].astype(np.int64), lv[0]["rowptr"].astype(np.int64)), shape=(n, n))
    setup = time.perf_counter() - t0
    out = {"workload": "3D Q1 -div(c K grad u) = f, K = diag(1,1,10), log-normal c (sigma 1.5, seed + 1000) per "
                       "cell, level 7 (2,146,689 dofs, 57,066,625 nnz), rhs uniform[-1,1] (seed + 2000); PCG + MG "
                       "V(4,4) Richardson(Jacobi 0.5), levels 7..1, exact coarse solve, tol 1e-8 on the FP64 defect",
           "host_assembly_s": setup}
    x64 = None
    for prec in ("f64", "f32", "f16"):
        cfg = feat.mgpcg_cfg(coarse="direct", max_iter=300, precision=prec, graphs=False)
        s_ = feat.MgPcg(cfg, levels=lv)
        x = torch.zeros(n, dtype=torch.float64, device=dev)
        ts = []
        for k in range(1 + max(1, args.steps // 3)):
            x.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            r = s_.solve(x, b)
            torch.cuda.synchronize()
            if k:
                ts.append(time.perf_counter() - t)
        s_.close()
        assert r["status"] == "success", (prec, r)
        xh = x.cpu().numpy()
        d = bh - S @ xh
        d[lv[0]["filt"].astype(np.int64)] = 0.0
        e = {"s_per_solve": float(np.mean(ts)), "iters": r["iters"],
             "true_rel_residual_fp64": float(np.linalg.norm(d) / np.linalg.norm(bh))}
        if prec == "f64":
            x64 = xh
        else:
            e["speedup_vs_fp64"] = out["f64"]["s_per_solve"] / e["s_per_solve"]
            e["rel_error_vs_fp64_solution"] = float(np.linalg.norm(xh - x64) / np.linalg.norm(x64))
        out[prec] = e
        log(f"variable-coefficient L7 {prec}: {e['s_per_solve'] * 1e3:.2f} ms, {r['iters']} iterations, "
            f"true rel. residual {e['true_rel_residual_fp64']:.2e}")
    return out


def learned_leg(feat, torch, dev, args):
    """Learned smoother on 3D Q1 anisotropic log-normal diffusion, level 6 (274,625 dofs):
    flexible PCG(MG-V(4,4)) with the Jacobi smoother vs the trained BF16-MLP smoother,
    both to a 1e-8 FP64 relative defect; training on the GPU from the multigrid's own
    residual/error fields (fixed seeds, deterministic)."""
    t0 = time.perf_counter()
    # the network on the first of the four steps of each finest smoother call, damped Jacobi on
    # the other three (measured: 1 learned step 21.2 ms / 26 it, 2: 22.4 ms / 24 it, all four:
    # 27.2 ms / 23 it, Jacobi 24.9 ms / 37 it)
    learned_steps = 1
    s = feat.LearnedMG(level_fine=6, level_coarse=1, lr=1e-3, learned_steps=learned_steps)
    rho = s.generate()
    e_jac = s.energy()
    t1 = time.perf_counter()
    # trained on the fixed Jacobi-multigrid snapshots: networks trained after an on-policy
    # refresh (snapshots of the learned smoother's own cycles) failed their validation solves
    rounds, steps, on_policy = 4, 300, False
    fit = s.fit(rounds=rounds, steps=steps, on_policy=on_policy, seed=1)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    e_lrn = s.energy()
    setup = time.perf_counter() - t0 - t_train

    def timed(learned):
        for _ in range(args.warmup):
            r = s.solve(learned)
        ts = []
        for _ in range(args.steps):
            r = s.solve(learned)
            ts.append(r["seconds"])
        return r, float(np.mean(ts))

    rj, tj = timed(False)
    rl, tl = timed(True)
    s.close()
    assert rj["status"] == "success" and rl["status"] == "success", (rj, rl)
    out = {"workload": "3D Q1 -div(c K grad u) = f, K = diag(1,1,10), log-normal c (sigma 1.5), level 6, "
                       "274,625 dofs; flexible PCG + MG V(4,4), levels 6..1, tol 1e-8 (FP64 defect)",
           "jacobi": {"iters": rj["iters"], "s_per_solve": tj, "final_rel_defect": rj["def_final"] / rj["def_init"]},
           "learned": {"iters": rl["iters"], "s_per_solve": tl, "final_rel_defect": rl["def_final"] / rl["def_init"],
                       "fell_back": rl["fell_back"], "fallback_iter": rl["fallback_iter"]},
           "speedup_vs_jacobi": tj / tl, "iteration_ratio": rj["iters"] / max(rl["iters"], 1),
           "learned_steps_per_smoother_call": learned_steps,
           "train": {"adam_steps": rounds * steps, "rounds": rounds, "on_policy_refresh": on_policy, "s": t_train,
                     "setup_and_data_s": setup, "validation": fit,
                     "one_step_rel_energy_jacobi": e_jac, "one_step_rel_energy_learned": e_lrn},
           "rho_jacobi": rho}
    log(f"learned smoother L6: Jacobi {rj['iters']} it {tj * 1e3:.2f} ms, learned {rl['iters']} it {tl * 1e3:.2f} ms "
        f"(fallback {rl['fell_back']}), training {t_train:.2f} s (validation {fit}), one-step energy "
        f"{e_jac:.3f} -> {e_lrn:.3f}")
    return out


def read_ceilings(feat, torch, dev, stream, h, nbytes, reps):
    """GB/s of feat_stream_read_f64 over nbytes: one buffer read repeatedly, and eight
    nbytes windows of a larger buffer read in turn (each from HBM, not the last-level cache)."""
    k, _ = feat.libs()
    nd = (nbytes // 8 + 1) & ~1
    big = torch.ones(8 * nd, dtype=torch.float64, device=dev)
    sink = torch.zeros(1, dtype=torch.float64, device=dev)
    out = {"bytes": nd * 8}
    for tag, nwin in (("same_bytes_repeated_gbps", 1), ("same_bytes_from_hbm_gbps", 8)):
        for w in range(nwin):
            feat.check(k.feat_stream_read_f64(h.h, feat.ptr(big[w * nd:]), nd, feat.ptr(sink)), "stream read")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for i in range(reps):
            k.feat_stream_read_f64(h.h, feat.ptr(big[(i % nwin) * nd:]), nd, feat.ptr(sink))
        e1.record(stream)
        e1.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        out[tag] = nd * 8 / (us * 1e-6) / 1e9
        out[tag.replace("_gbps", "_us")] = us
    log(f"read ceiling over {nd * 8 / 1e6:.1f} MB: {out['same_bytes_repeated_gbps']:.0f} GB/s repeated, "
        f"{out['same_bytes_from_hbm_gbps']:.0f} GB/s from HBM")
    return out


def spmv_only(feat, torch, dev, stream, args):
    """SpMV r = A*x on the level-7 matrix in each format, args.spmv_reps launches each,
    then args.spmv_reps device-to-device copies of 1 GiB (the counter calibration)."""
    q = feat.q1_problem(DIM, LF)
    n = q["n"]
    val = torch.from_numpy(q["val"]).to(dev)
    col = torch.from_numpy(q["col"].view(np.int64)).to(dev)
    rp = torch.from_numpy(q["rowptr"].view(np.int64)).to(dev)
    k, _ = feat.libs()
    h = feat.Handle(0, stream=stream.cuda_stream)
    xs = torch.rand(n, dtype=torch.float64, device=dev)
    r = torch.empty(n, dtype=torch.float64, device=dev)
    plans = {"sell64_diag": feat.SellPlan(h, val, col, rp, n, n),
             "sell64_dict": feat.SellPlan(h, val, col, rp, n, n, diag=False),
             "sell64": feat.SellPlan(h, val, col, rp, n, n, dict=False)}
    mats = {"csr_u64": feat.csr_matrix(val, col, rp, n, n)}
    mats.update({kk: p.m for kk, p in plans.items()})
    out = {}
    for name, m in mats.items():
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.spmv_reps):
            feat.check(k.feat_matrix_apply_f64(h.h, C.byref(m), feat.ptr(r), 1.0, feat.ptr(xs), 0.0, None), name)
        e1.record(stream)
        e1.synchronize()
        out[name + "_us"] = e0.elapsed_time(e1) * 1e3 / args.spmv_reps
        log(f"{name}: {out[name + '_us']:.1f} us per launch")
    # one Jacobi step on the diagonal plan per call (the solve's dominant kernel)
    inv_t = torch.rand(n, dtype=torch.float64, device=dev)
    rhs_t = torch.rand(n, dtype=torch.float64, device=dev)
    xo, tmp_t = torch.empty_like(rhs_t), torch.rand(n, dtype=torch.float64, device=dev)
    keep_ws, ws = feat.reduce_ws(torch, dev)
    dm = plans["sell64_diag"].m
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.spmv_reps):
        feat.check(k.feat_rj_smooth_from1_f64(h.h, C.byref(dm), feat.ptr(inv_t), None, 2, 0.5, feat.ptr(rhs_t), None,
                                              feat.ptr(xo), feat.ptr(tmp_t), None, None, None, None, ws), "step")
    e1.record(stream)
    e1.synchronize()
    out["jacobi_step_us"] = e0.elapsed_time(e1) * 1e3 / args.spmv_reps
    log(f"jacobi step: {out['jacobi_step_us']:.1f} us per launch")
    out["read_ceiling"] = read_ceilings(feat, torch, dev, stream, h, int(dm.nnz) + 256 * 12 + 4 * n * 8, args.spmv_reps)
    src = torch.empty(1 << 27, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    src.fill_(1.0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.spmv_reps):
        feat.check(k.feat_copy(h.h, feat.ptr(dst), feat.ptr(src), src.numel() * 8), "copy")
    e1.record(stream)
    e1.synchronize()
    out["copy_1gib_us"] = e0.elapsed_time(e1) * 1e3 / args.spmv_reps
    log(f"copy 1 GiB: {out['copy_1gib_us']:.1f} us per launch")
    for p in plans.values():
        p.close()
    h.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the FP64 MG-PCG solves (for kernel-trace profiles of the solve)")
    ap.add_argument("--learned-only", action="store_true", help="only the learned-smoother leg")
    ap.add_argument("--stokes-only", action="store_true", help="only the Stokes leg")
    ap.add_argument("--batch-only", action="store_true", help="only the 64-instance batch leg")
    ap.add_argument("--mixed-only", action="store_true",
                    help="only the mixed-precision leg on the variable-coefficient level-7 problem")
    ap.add_argument("--assembly-only", action="store_true", help="only the device-assembly leg")
    ap.add_argument("--smoothers-only", action="store_true", help="only the SOR / SSOR / ILU(0) leg")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the SpMV loops (pair dictionary, plain SELL-64, CSR) and a 1 GiB device copy, "
                         "for hardware-counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "domain decomposition is not implemented here.", file=sys.stderr)
        sys.exit(2)

    import torch
    feat = load_feat()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    if args.spmv_only:
        spmv_only(feat, torch, dev, stream, args)
        return
    if args.learned_only:
        print(json.dumps({"learned_smoother": learned_leg(feat, torch, dev, args)}))
        return
    if args.stokes_only:
        print(json.dumps({"stokes": stokes_leg(feat, torch, dev, args)}))
        return
    if args.mixed_only:
        print(json.dumps({"mixed_precision_variable_coefficient": mixed_varcoef_leg(feat, torch, dev, args)}))
        return
    if args.batch_only:
        print(json.dumps({"batch": batch_leg(feat, torch, dev, args)}))
        return
    if args.smoothers_only:
        print(json.dumps({"smoothers": smoothers_leg(feat, torch, dev, args)}))
        return
    if args.assembly_only:
        print(json.dumps({"assembly": assembly_leg(feat, torch, dev, args)}))
        return

    # ---- setup (device assembly of levels 7..2, host assembly of level 1; not timed)
    t0 = time.perf_counter()
    # eager launches: on this path they beat replaying recorded HIP graphs by ~2 % (the
    # graph variant is timed below and reported beside)
    cfg = feat.mgpcg_cfg(nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0, coarse="direct",
                         tol_rel=1e-8, max_iter=100, graphs=False)
    solver = feat.MgPcg(cfg, q1=(DIM, LF, LC))
    A = solver.matrix(0)
    n, nnz = int(A.rows), int(A.nnz)
    b_host = feat.rhs_random(SEED, n, -1.0, 1.0)
    m = 2 ** LF + 1
    idx = np.arange(n)
    c = np.stack([idx % m, (idx // m) % m, idx // (m * m)])
    bnd = np.any((c == 0) | (c == m - 1), axis=0)
    b_host[bnd] = 0.0                        # filter_rhs of the homogeneous Dirichlet filter
    b = torch.from_numpy(b_host).to(dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t0:.1f} s: n={n} nnz={nnz}")

    # ---- MG-PCG time to solution
    res = None
    for _ in range(args.warmup):
        x.zero_()
        torch.cuda.synchronize()
        res = solver.solve(x, b)
    times = []
    for k in range(args.steps):
        x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = solver.solve(x, b)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        log(f"solve {k}: {times[-1] * 1e3:.2f} ms, {res['iters']} iterations, "
            f"rel. defect {res['def_final'] / res['def_init']:.3e}")
    assert res["status"] == "success", res
    t_mean = float(np.mean(times))

    if args.headline_only:
        solver.close()
        print(json.dumps({"metric": "MG-PCG time-to-1e-8 residual (s), 3D Q1 Poisson 2.1M dofs (headline solves only)",
                          "value": t_mean, "unit": "s", "iters": res["iters"], "steps": args.steps}))
        return

    # ---- the same solver replaying recorded HIP graphs (two per PCG iteration)
    se = feat.MgPcg(feat.mgpcg_cfg(nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0, coarse="direct",
                                   tol_rel=1e-8, max_iter=100, graphs=True), q1=(DIM, LF, LC))
    xe = torch.zeros(n, dtype=torch.float64, device=dev)
    te = []
    for k in range(args.warmup + args.steps):
        xe.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        re_ = se.solve(xe, b)
        torch.cuda.synchronize()
        if k >= args.warmup:
            te.append(time.perf_counter() - t)
    se.close()
    assert re_["iters"] == res["iters"] and bool(torch.equal(xe, x)), "graph and eager solves differ"
    t_eager = float(np.mean(te))
    log(f"recorded graphs: {t_eager * 1e3:.2f} ms per solve (eager {t_mean * 1e3:.2f} ms), identical result")

    # ---- the same solver on plain SELL-64 plans (full FP64 values, i16 column offsets) and
    # on entry-order dictionary plans
    plan_rel_diff = {}

    def solve_with_plans(flags, what):
        feat.runtime_sell_flags(flags)
        sp = feat.MgPcg(cfg, q1=(DIM, LF, LC))
        feat.runtime_sell_flags(0)
        xp = torch.zeros(n, dtype=torch.float64, device=dev)
        tp = []
        for k in range(args.warmup + args.steps):
            xp.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            rp_ = sp.solve(xp, b)
            torch.cuda.synchronize()
            if k >= args.warmup:
                tp.append(time.perf_counter() - t)
        sp.close()
        # every SpMV row is bit-identical across plans; the dots and norms are summed per
        # workgroup chunk, whose rows differ (248 on diagonal plans, 256 on the others)
        rd = float(torch.linalg.norm(xp - x) / torch.linalg.norm(x))
        assert rp_["iters"] == res["iters"] and rd <= 1e-12, f"{what} and diagonal plans differ: {rd}"
        plan_rel_diff[what] = rd
        tt = float(np.mean(tp))
        log(f"{what}: {tt * 1e3:.2f} ms per solve (diagonal dictionary {t_mean * 1e3:.2f} ms), "
            f"same iterations, |x - x_diag| / |x_diag| = {rd:.1e}")
        return tt

    t_plain = solve_with_plans(feat.SELL_NO_DICT, "plain SELL-64 plans")
    t_entry = solve_with_plans(feat.SELL_NO_DIAG, "entry-order dictionary plans")
    xp = torch.zeros(n, dtype=torch.float64, device=dev)

    # ---- the same solver with the default launch shapes (no autotuning)
    su = feat.MgPcg(feat.mgpcg_cfg(nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0, coarse="direct",
                                   tol_rel=1e-8, max_iter=100, autotune=False, graphs=False), q1=(DIM, LF, LC))
    tu = []
    for k in range(args.warmup + args.steps):
        xp.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        ru_ = su.solve(xp, b)
        torch.cuda.synchronize()
        if k >= args.warmup:
            tu.append(time.perf_counter() - t)
    su.close()
    assert ru_["iters"] == res["iters"] and bool(torch.equal(xp, x)), "tuned and untuned solves differ"
    t_untuned = float(np.mean(tu))
    log(f"default launch shapes: {t_untuned * 1e3:.2f} ms per solve (autotuned {t_mean * 1e3:.2f} ms), identical result")

    # ---- the mixed-precision forms: FP32 V-cycle (FP32 or FP16 matrix values)
    # inside the FP64 PCG; outer residual, update, convergence test and coarse solve FP64
    kk, _ = feat.libs()
    hh = feat.Handle(0, stream=stream.cuda_stream)
    bnd_t = torch.from_numpy(bnd).to(dev)
    mixed = {}
    for prec in ("f32", "f16"):
        cfgm = feat.mgpcg_cfg(nu_pre=4, nu_post=4, omega_jac=0.5, omega_rich=1.0, coarse="direct",
                              tol_rel=1e-8, max_iter=100, precision=prec, graphs=False)
        sm = feat.MgPcg(cfgm, q1=(DIM, LF, LC))
        xm = torch.zeros(n, dtype=torch.float64, device=dev)
        for _ in range(args.warmup):
            xm.zero_()
            torch.cuda.synchronize()
            rm = sm.solve(xm, b)
        tm = []
        for k in range(args.steps):
            xm.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            rm = sm.solve(xm, b)
            torch.cuda.synchronize()
            tm.append(time.perf_counter() - t)
        assert rm["status"] == "success", rm
        # FP64 verification: true residual with the FP64 CSR, error vs the FP64 solve
        rr = torch.empty(n, dtype=torch.float64, device=dev)
        feat.check(kk.feat_csr_apply_f64_u64(hh.h, feat.ptr(rr), -1.0, feat.ptr(xm), 1.0, feat.ptr(b),
                                             C.c_void_p(A.val), C.c_void_p(A.col), C.c_void_p(A.rowptr), n, n,
                                             nnz, 0), "residual")
        rr[bnd_t] = 0.0
        tmean = float(np.mean(tm))
        mixed[prec] = {"s_per_solve": tmean, "iters": rm["iters"], "speedup_vs_fp64": t_mean / tmean,
                       "speedup_vs_fp64_plain_plans": t_plain / tmean,
                       "true_rel_residual_fp64": float(torch.linalg.norm(rr) / torch.linalg.norm(b)),
                       "rel_error_vs_fp64_solution": float(torch.linalg.norm(xm - x) / torch.linalg.norm(x))}
        sm.close()
        log(f"mixed {prec}: {tmean * 1e3:.2f} ms, {rm['iters']} iterations, true rel. residual "
            f"{mixed[prec]['true_rel_residual_fp64']:.3e}, |x-x64|/|x64| {mixed[prec]['rel_error_vs_fp64_solution']:.3e}, "
            f"speedup {t_mean / tmean:.2f}x")
    hh.close()

    # ---- Stokes: Q2/P1dc lid-driven cavity level 6, Richardson(MG-V(Richardson(Schwarz(AmaVanka))))
    mixed_vc = mixed_varcoef_leg(feat, torch, dev, args)
    stokes = stokes_leg(feat, torch, dev, args)
    batch = batch_leg(feat, torch, dev, args)
    learned = learned_leg(feat, torch, dev, args)
    assembly = assembly_leg(feat, torch, dev, args)
    smoothers = smoothers_leg(feat, torch, dev, args)

    # ---- SpMV r = A*x on the same matrix, HIP events on the launch stream:
    # the caller's CSR (u64) and the SELL-64 plan the solver runs on
    k, _ = feat.libs()
    h = feat.Handle(0, stream=stream.cuda_stream)
    xs = torch.rand(n, dtype=torch.float64, device=dev)
    r = torch.empty(n, dtype=torch.float64, device=dev)
    csr = feat.Matrix(feat.FMT_CSR, feat.IDX_U64, n, n, nnz, A.val, A.col, A.rowptr, None, 0, 1.0)
    plan = feat.Matrix()
    feat.check(k.feat_sell_create_ex_f64(h.h, C.byref(A), feat.SELL_NO_DICT, C.byref(plan)), "feat_sell_create_ex_f64")
    eplan = feat.Matrix()
    feat.check(k.feat_sell_create_ex_f64(h.h, C.byref(A), feat.SELL_NO_DIAG, C.byref(eplan)), "feat_sell_create_ex_f64")
    assert eplan.index == feat.IDX_DICT
    dplan = feat.Matrix()
    feat.check(k.feat_sell_create_f64(h.h, C.byref(A), C.byref(dplan)), "feat_sell_create_f64")
    assert dplan.index == feat.IDX_DIAG

    def time_apply(m):
        for _ in range(5):
            feat.check(k.feat_matrix_apply_f64(h.h, C.byref(m), feat.ptr(r), 1.0, feat.ptr(xs), 0.0, None), "spmv")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.spmv_reps):
            k.feat_matrix_apply_f64(h.h, C.byref(m), feat.ptr(r), 1.0, feat.ptr(xs), 0.0, None)
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / args.spmv_reps

    csr_ms = time_apply(csr)
    sell_ms = time_apply(plan)
    edict_ms = time_apply(eplan)
    dict_ms = time_apply(dplan)

    # one fused Richardson(Jacobi) step on the diagonal plan (the solve's most frequent
    # launch): smoother calls of 9 and of 1 steps, the difference over 8 steps
    inv_t = torch.rand(n, dtype=torch.float64, device=dev)
    rhs_t = torch.rand(n, dtype=torch.float64, device=dev)
    xo, tmp_t = torch.empty_like(rhs_t), torch.empty_like(rhs_t)
    keep_ws, ws = feat.reduce_ws(torch, dev)

    # a two-step smoother whose first iterate is already in tmp is exactly one launch of the
    # Jacobi step kernel (x = tmp + omega * filter(inv * (rhs - A tmp)))
    def call_step():
        return k.feat_rj_smooth_from1_f64(h.h, C.byref(dplan), feat.ptr(inv_t), None, 2, 0.5, feat.ptr(rhs_t),
                                          None, feat.ptr(xo), feat.ptr(tmp_t), None, None, None, None, ws)

    def time_loop(call, reps):
        for _ in range(3):
            feat.check(call(), "warm-up")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            call()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    step_ms = time_loop(call_step, args.spmv_reps)
    # measured read ceilings (feat_stream_read_f64): a buffer of the Jacobi step's bytes read
    # again and again (held in the last-level cache like the step's own arrays), and the same
    # bytes from HBM (eight such windows of a larger buffer in turn)
    ceil = read_ceilings(feat, torch, dev, stream, h, int(dplan.nnz) + 256 * 12 + 4 * n * 8, args.spmv_reps)
    idx_bytes = {feat.IDX_D16: 2, feat.IDX_U32: 4}[plan.index]
    nslices = (n + 63) // 64
    # bytes each kernel must move: its own arrays once, x once, r once
    csr_bytes = spmv_bytes(n, nnz)
    sell_bytes = int(plan.nnz) * (8 + idx_bytes) + n * 2 + (nslices + 1) * 8 + 2 * n * 8
    # pair dictionary in entry order: one byte per stored entry, row lengths, slice offsets,
    # the 3 KB pair table, x, r; by diagonal: one byte per slot (27 per row), the table, x, r
    edict_bytes = int(eplan.nnz) + n * 2 + (nslices + 1) * 8 + 256 * 12 + 2 * n * 8
    dict_bytes = int(dplan.nnz) + 256 * 12 + 2 * n * 8
    edict_gbs = edict_bytes / (edict_ms * 1e-3) / 1e9
    csr_gbs = csr_bytes / (csr_ms * 1e-3) / 1e9
    sell_gbs = sell_bytes / (sell_ms * 1e-3) / 1e9
    dict_gbs = dict_bytes / (dict_ms * 1e-3) / 1e9
    # Jacobi step: the plan's bytes, x_k (gathered and read by the epilogue: once), rhs,
    # inverse diagonal, x_{k+1}
    step_bytes = int(dplan.nnz) + 256 * 12 + 4 * n * 8
    step_gbs = step_bytes / (step_ms * 1e-3) / 1e9
    log(f"jacobi step on the diagonal plan: {step_ms * 1e3:.1f} us, {step_gbs:.0f} GB/s")
    ref_model = nnz * 16 + n * 8             # product_matvec-bench.cpp:41-44 model (GiB = 1024^3)
    log(f"spmv csr-u64: {csr_ms * 1e3:.1f} us, {csr_gbs:.0f} GB/s | sell-64 ({'i16 delta' if idx_bytes == 2 else 'u32'}): "
        f"{sell_ms * 1e3:.1f} us, {sell_gbs:.0f} GB/s own bytes, {csr_bytes / (sell_ms * 1e-3) / 1e9:.0f} GB/s CSR-equivalent"
        f" | pair dictionary (entry order): {edict_ms * 1e3:.1f} us, {edict_gbs:.0f} GB/s"
        f" | pair dictionary by diagonal: {dict_ms * 1e3:.1f} us, {dict_gbs:.0f} GB/s own bytes")
    padded = int(plan.nnz)
    dict_padded = int(dplan.nnz)
    for m_ in (plan, eplan, dplan):
        feat.check(k.feat_sell_destroy_f64(C.byref(m_)), "feat_sell_destroy_f64")
    h.close()
    achieved = step_gbs

    out = {
        "metric": "MG-PCG time-to-1e-8 residual (s) + SpMV HBM GB/s, 3D Q1 Poisson 2.1M dofs",
        "value": t_mean,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mean * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "3D Q1 Poisson unit cube level 7 (128^3 cells, 2,146,689 dofs, 57,066,625 nnz "
                        "27-point CSR f64/u64), rhs uniform[-1,1] seed 28054777172512, PCG + MG V(4,4) "
                        "Richardson(Jacobi 0.5), levels 7..1, exact coarse solve, tol 1e-8",
            "dofs": n, "nnz": nnz, "levels": LF - LC + 1, "iters": res["iters"],
            "final_rel_defect": res["def_final"] / res["def_init"],
            "graph_replay_s_per_solve": t_eager,
            "launches": "eager (the graph replay of the same kernels is graph_replay_s_per_solve)",
            "plain_sell_plans_s_per_solve": t_plain,
            "entry_order_dictionary_plans_s_per_solve": t_entry,
            "other_plans_rel_diff_x": plan_rel_diff,
            "default_launch_shapes_s_per_solve": t_untuned,
            "matrix_plan": "SELL-64 pair dictionary by diagonal (1 byte per offset slot, 9 triples of offsets, "
                           "42 distinct (col-row, value) pairs)",
            "parallelism": "single GPU",
        },
        "mixed_precision": {"form": "FP32 V-cycle on FP32 / FP16-valued SELL-64 plans (FP64 coarse solve) "
                                    "inside FP64 PCG",
                            "fp32": mixed["f32"],
                            "fp16_values_where_smaller": mixed["f16"],
                            "note": "on this constant-coefficient hierarchy every operator has a 1-byte pair "
                                    "dictionary plan, smaller than FP16 values, so the FP16 form stores only "
                                    "the transfer operators in FP16 and uses the FP32 dictionary elsewhere",
                            "variable_coefficient_level7": mixed_vc},
        "spmv": {
            "sell64_diag": {"kernel": "k_sell_diag_rows<double,EpiApply<double>>", "us_per_launch": dict_ms * 1e3,
                            "bytes_per_launch": dict_bytes, "gbps": dict_gbs,
                            "csr_equivalent_gbps": csr_bytes / (dict_ms * 1e-3) / 1e9,
                     "jacobi_step": {"kernel": "k_sell_diag_rows<double,EpiSmooth<double>> (one fused "
                                               "Richardson(Jacobi) step, the solve's most frequent launch)",
                                     "us_per_launch": step_ms * 1e3, "achieved": step_gbs,
                                     "frac": step_gbs / PEAK_HBM_GBS, "algorithmic_bytes": step_bytes}, "stored_slots": dict_padded},
            "sell64_dict": {"kernel": "k_sell_dict_rows<double,EpiApply<double>,8>", "us_per_launch": edict_ms * 1e3,
                            "bytes_per_launch": edict_bytes, "gbps": edict_gbs,
                            "csr_equivalent_gbps": csr_bytes / (edict_ms * 1e-3) / 1e9, "stored_entries": int(eplan.nnz)},
            "sell64": {"kernel": "k_sell_rows<double,short,true,EpiApply<double>>", "us_per_launch": sell_ms * 1e3,
                       "bytes_per_launch": sell_bytes, "gbps": sell_gbs,
                       "csr_equivalent_gbps": csr_bytes / (sell_ms * 1e-3) / 1e9,
                       "index": "i16 delta" if idx_bytes == 2 else "u32", "padded_entries": padded},
            "csr_u64": {"kernel": "k_csr_rows<double,unsigned long,EpiApply<double>>", "us_per_launch": csr_ms * 1e3,
                        "bytes_per_launch": csr_bytes, "gbps": csr_gbs,
                        "gibps_reference_model": ref_model / (csr_ms * 1e-3) / 1024 ** 3},
        },
        "stokes": stokes,
        "batch": batch,
        "learned_smoother": learned,
        "assembly": assembly,
        "smoothers": smoothers,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": SPMV_TRAFFIC_BYTES["jacobi_step"],
                     "algorithmic_bytes": step_bytes, "us_per_launch": step_ms * 1e3,
                     "kernel": "k_sell_diag_rows<double,EpiSmooth<double>,9,7>: one fused Richardson(Jacobi) "
                               "step x_{k+1} = x_k + 0.5 filter(inv_diag (rhs - A x_k)) on the level-7 diagonal "
                               "pair-dictionary plan (the solve's dominant kernel: 35 of its launches per solve), "
                               "bytes = codes (1 B per offset slot, 27 slots in 7 words per row) + pair table + "
                               "x_k + rhs + inv_diag + x_{k+1}",
                     "measured_read_ceiling": ceil,
                     "frac_of_read_ceiling_same_bytes": achieved / ceil["same_bytes_repeated_gbps"],
                     "frac_of_read_ceiling_hbm": achieved / ceil["same_bytes_from_hbm_gbps"],
                     "spmv_diag_apply": {"achieved": dict_gbs, "frac": dict_gbs / PEAK_HBM_GBS,
                                         "traffic": SPMV_TRAFFIC_BYTES["sell64_diag"], "algorithmic_bytes": dict_bytes},
                     "entry_order_dictionary": {"achieved": edict_gbs, "frac": edict_gbs / PEAK_HBM_GBS,
                                                "traffic": SPMV_TRAFFIC_BYTES["sell64_dict"],
                                                "algorithmic_bytes": edict_bytes},
                     "csr_equivalent_gbps": csr_bytes / (dict_ms * 1e-3) / 1e9,
                     "plain_sell64": {"achieved": sell_gbs, "frac": sell_gbs / PEAK_HBM_GBS,
                                      "traffic": SPMV_TRAFFIC_BYTES["sell64"], "algorithmic_bytes": sell_bytes},
                     "csr_u64": {"achieved": csr_gbs, "frac": csr_gbs / PEAK_HBM_GBS,
                                 "traffic": SPMV_TRAFFIC_BYTES["csr_u64"], "algorithmic_bytes": csr_bytes}},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline: reference MG-PCG on one host core ...")
        cb = cpu_baseline(feat, b_host, res["iters"])
        out["cpu_baseline"] = {kk: cb[kk] for kk in ("value", "unit", "cores", "kind", "sample")}
        log(f"cpu baseline: {cb['value']:.2f} s, {cb['iters']} iterations")
    solver.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
