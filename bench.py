"""Benchmark: DG transport operator apply and time-to-1e-8 solve on one GPU.

Workload (uniform absorber, "C2"): [0,1]^2 x [0, 2 pi), 32 x 32 spatial
columns x 16 angular cells, 5 nodes per direction (p = 4): 2,048,000 DOF;
kappa = 1 + 0.5 sin(2 pi x) sin(2 pi y), sigma = 0.7 kappa, Henyey-Greenstein
g = 0.8, unit isotropic inflow on the left wall, no forcing.  Data are
synthetic (seeded random operand vectors) on the deterministic mesh.

A "step" is one application of the interior transport operator A_II
(mass + interior convection + upwind boundary convection + scattering,
FP64) to a vector already resident in HBM.  ``value`` is DOF/s per apply.
After the timed steps the script also solves A_II x = rhs to a true FP64
relative residual of 1e-8 with (a) plain FP64 GMRES + block-Jacobi and
(b) FP64 iterative refinement around reduced-precision applies; trains the
learned coarse preconditioner on solver-generated pairs of the random-field
workload (64 x 64 x 16 cells, p = 4, 8,192,000 DOF) and times plain FP64,
mixed and learned+mixed solves of an unseen field there; times the error
indicator kernels on the same mesh; times the same three solves on the
batched uneven-order hp mesh with two strong scatterers (8,263,762 DOF),
the coarse model trained on the two-scatterer family; times the oracle's assembled CSR
SpMV on a bounded sample on the host CPU; and solves a smaller mesh with
C2's coefficients (16 x 16 x 8 cells, 256,000 DOF) to 1e-8 on the host CPU
(scipy GMRES with the oracle's block-Jacobi) and on the GPU.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solve]
                       [--no-cpu] [--no-learned] [--no-indicators] [--no-hp] [--hp-fields N]
                       [--learned-fields N] [--learned-kappa-scales S,...] [--calib]
                       [--profile-apply {c2-f64,c2-mixed,c3-f64,c3-mixed}]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "radiative-transfer-amr_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-indicators", action="store_true")
    ap.add_argument("--no-hp", action="store_true",
                    help="skip the hp-mesh (C3) time-to-1e-8 section")
    ap.add_argument("--hp-fields", type=int, default=17,
                    help="training fields of the learned correction on the hp mesh")
    ap.add_argument("--learned-fields", type=int, default=17)
    ap.add_argument("--learned-kappa-scales", default="1,30",
                    help="comma-separated kappa scales of the random-field workload "
                         "(1: as specified, median kappa 1; 30: optically thick)")
    ap.add_argument("--calib", action="store_true",
                    help="run one plain 1 GiB device copy after the timed region "
                         "(counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--path", type=int, default=-1,
                    help="apply path for --profile-apply (-1 automatic, 0 generic, "
                         "1 fused parts, 3 hp column kernel)")
    ap.add_argument("--profile-apply", choices=sorted(PROFILE_WORKLOADS),
                    help="only time the interior operator apply of one workload and "
                         "precision (for rocprofv3 counter runs; prints its own JSON line)")
    return ap.parse_args()


# --profile-apply workloads: (mesh, coefficients, precision)
PROFILE_WORKLOADS = {
    "c2-f64": ("c2", "f64"), "c2-mixed": ("c2", "mixed"),
    "c3-f64": ("c3", "f64"), "c3-mixed": ("c3", "mixed"),
}


def profile_apply(args):
    """Interior operator applies only (``--profile-apply``): C2 (uniform
    absorber, 32 x 32 x 16 cells, p = 4) or C3 (batched uneven-order hp
    mesh of 8,263,762 DOF with two strong scatterers), FP64 or mixed (FP32
    convection, BF16 scattering); ``warmup`` untimed and ``steps`` timed
    launches on random input, HIP events on the launch stream; with
    ``--calib`` the two 1 GiB calibration copies follow."""
    import torch
    from rtx import problems
    from rtx.operator import TransportOperator, OP_SYSTEM, OP_RESTRICT, PREC_F64, PREC_MIXED
    mesh_name, prec_name = PROFILE_WORKLOADS[args.profile_apply]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if mesh_name == "c2":
        mesh, wl = workload()
    else:
        mesh, wl = problems.random_hp_mesh(0), problems.two_scatterers_c3()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi, device=dev)
    op.set_path(args.path)
    prec = PREC_F64 if prec_name == "f64" else PREC_MIXED
    dt = torch.float64 if prec == PREC_F64 else torch.float32
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    x = (torch.rand(op.ndof, dtype=torch.float64, device=dev, generator=gen) * 2 - 1).to(dt)
    op.mask(x)
    y = torch.empty_like(x)
    ops = OP_SYSTEM | OP_RESTRICT
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        op.apply(x, y, ops=ops, precision=prec)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        op.apply(x, y, ops=ops, precision=prec)
    ev1.record(stream)
    torch.cuda.synchronize()
    us = ev0.elapsed_time(ev1) / max(args.steps, 1) * 1e3
    if args.calib:
        _calibration_copies(op, dev)
    bytes_, flops = algorithmic(op)
    print(json.dumps({"profile_apply": args.profile_apply, "path": args.path, "ndof": op.ndof,
                      "us_per_apply": us, "gdof_per_s": op.ndof / us * 1e-3,
                      "algorithmic_bytes_f64": bytes_, "algorithmic_flops": flops,
                      "steps": args.steps, "warmup": args.warmup}), flush=True)


def _calibration_copies(op, dev):
    """Two copies of 1 GiB for the traffic counters: the runtime's blit, then
    rtx_copy_f64 (plain per-lane FP64 loads and stores, the access type of
    the operator kernels)."""
    import ctypes as C
    import torch
    from rtx import _lib
    from rtx.operator import stream_handle
    big = torch.empty(1 << 27, dtype=torch.float64, device=dev)   # 1 GiB
    big2 = torch.empty_like(big)
    big2.copy_(big)
    _lib.check(op.lib.rtx_copy_f64(C.c_void_p(big.data_ptr()), C.c_void_p(big2.data_ptr()),
                                   big.numel(), stream_handle()), "rtx_copy_f64")
    torch.cuda.synchronize()


def workload():
    from rtx import problems
    mesh = problems.uniform_mesh([1.0, 1.0], [5, 5, 5], nref_ang=4, nref_spt=5)
    return mesh, problems.absorber_c2()


def algorithmic(op):
    """Bytes and flops of one FP64 apply (read x, write y, read kappa and
    sigma at the spatial nodes; scattering 2 N_sp Nth^2 per column, volume
    and face terms)."""
    t = op.t
    import numpy as np
    nsp = int(t.nsp)
    bytes_ = 16 * t.ndof + 16 * nsp
    nang = t.col_nang.astype(np.int64)
    nsp_c = (t.col_nx * t.col_ny).astype(np.int64)
    scat = int((2 * nsp_c * nang * nang).sum())
    cell_n = (t.col_nx[t.cell_col] * t.col_ny[t.cell_col] * t.cell_nth).astype(np.int64)
    vol = int((cell_n * (2 * (t.col_nx[t.cell_col] + t.col_ny[t.cell_col]) + 8)).sum())
    face = 0
    for l in range(0, t.nlink):
        k, s, F = t.link_target[l], t.link_src[l], t.link_face[l]
        n0 = t.col_nx[t.cell_col[k]] if F % 2 else t.col_ny[t.cell_col[k]]
        n1 = t.col_nx[t.cell_col[s]] if F % 2 else t.col_ny[t.cell_col[s]]
        face += 2 * (int(n1) * int(t.cell_nth[s]) * int(t.cell_nth[k]) + int(n1) * int(n0) * int(t.cell_nth[k]))
    return bytes_, scat + vol + face


def measured_traffic():
    """HBM bytes per apply launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this same workload (counter values scaled by the
    factors a 1 GiB copy kernel gives), or None when it is absent."""
    import re
    path = os.path.join(ROOT, "profiles", "k_uniform_e_c2_f64.md")
    try:
        m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", open(path).read(), re.M)
    except OSError:
        return None
    return int(m.group(1)) if m else None


def _cpu_threads():
    """Host threads for the multi-core baseline: the cores this process may
    run on, capped by OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def _time_reps(fn, budget_s, max_reps=2000):
    fn()
    reps, t = 0, 0.0
    t0 = time.perf_counter()
    while t < budget_s and reps < max_reps:
        fn()
        reps += 1
        t = time.perf_counter() - t0
    return t / reps, reps


def cpu_baseline(mesh, wl, ncols=128, budget_s=4.0):
    """The oracle's assembled FP64 CSR A = (BC - IC) + M + S on the rows of a
    sample of columns, multiplied by the oracle's C CSR product on every host
    thread (the value) and by scipy on one thread; DOF/s = rows / time per
    SpMV."""
    import numpy as np
    from oracle import assemble as asm, build_c
    keys = sorted(mesh.cols)
    sample = keys[:: max(1, len(keys) // ncols)][:ncols]
    t0 = time.perf_counter()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi, cols=sample)
    A = asm.system_matrix(V)
    rows = np.unique(A.nonzero()[0])
    A = A[rows].tocsr()
    A.indices = A.indices.astype(np.int32)
    t_asm = time.perf_counter() - t0
    x = np.random.default_rng(0).uniform(-1, 1, A.shape[1])
    per1, reps1 = _time_reps(lambda: A @ x, budget_s)
    if not os.path.exists(build_c.OUT):
        build_c.build()
    lib = build_c.load()
    nt = _cpu_threads()
    if not np.array_equal(build_c.csr_spmv(lib, A, x, nt), A @ x):
        raise RuntimeError("the C CSR product differs from scipy's")
    indptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
    y = np.empty(A.shape[0])
    per, reps = _time_reps(lambda: lib.oracle_csr_spmv(
        A.shape[0], indptr.ctypes.data, A.indices.ctypes.data, A.data.ctypes.data,
        x.ctypes.data, y.ctypes.data, nt), budget_s)
    return {"value": rows.size / per, "unit": "DOF/s", "cores": nt, "kind": "port",
            "sample": "rows of {} of {} columns ({} rows, {} nnz, {:.0f} MB CSR) of the "
                      "oracle's assembled FP64 CSR A = (BC - IC) + M + S; the oracle's C CSR "
                      "product (OpenMP, {} threads, {} reps, bit-identical to scipy); "
                      "assembly of the sample took {:.1f} s".format(
                          len(sample), len(keys), rows.size, A.nnz,
                          (A.data.nbytes + A.indices.nbytes + indptr.nbytes) / 1e6, nt, reps,
                          t_asm),
            "scipy_1thread": {"value": rows.size / per1, "unit": "DOF/s", "cores": 1,
                              "reps": reps1}}


def cpu_solve_baseline(nref_ang=3, nref_spt=4):
    """Time-to-1e-8 on the host CPU beside the GPU, on a workload small
    enough for the CPU: C2's coefficients on 16 x 16 x 8 cells, p = 4
    (256,000 DOF).  CPU: the oracle's assembled interior system A_II and
    right-hand side f_I - A_IB b_B (``src/rt/solve.py:49-100``), scipy's
    GMRES(50) (left-preconditioned, as PETSc's default) with the oracle's
    per-cell block-Jacobi inverses, the CSR product and the block apply on
    every host thread through the oracle's C library; initial guess f_I as
    the reference's, then restarted until the true relative residual is
    1e-8.  GPU: the same system by ``_timed_solve`` (FP64 and mixed).
    ``seconds`` = block-Jacobi setup + solve on both sides; the CPU's
    assembly time is reported apart (the GPU's plan build likewise)."""
    import numpy as np
    import scipy.sparse.linalg as spla
    import torch
    from oracle import assemble as asm, system as osys, build_c
    from rtx import problems
    from rtx.operator import TransportOperator
    mesh = problems.uniform_mesh([1.0, 1.0], [5, 5, 5], nref_ang=nref_ang, nref_spt=nref_spt)
    wl = problems.absorber_c2()
    t0 = time.perf_counter()
    V = asm.assemble(mesh, wl.kappa, wl.sigma, wl.Phi)
    A = asm.system_matrix(V)
    f = asm.forcing(mesh, wl.f)
    b = asm.bcs(mesh, wl.bcs)
    mask = asm.intr_mask(mesh)
    A_II, rhs, I, _ = osys.interior_system(A, f, b, mask)
    A_II.indices = A_II.indices.astype(np.int32)
    t_asm = time.perf_counter() - t0
    if not os.path.exists(build_c.OUT):
        build_c.build()
    lib = build_c.load()
    nt = _cpu_threads()
    indptr = np.ascontiguousarray(A_II.indptr, dtype=np.int64)

    def matvec(v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty(A_II.shape[0])
        lib.oracle_csr_spmv(A_II.shape[0], indptr.ctypes.data, A_II.indices.ctypes.data,
                            A_II.data.ctypes.data, v.ctypes.data, out.ctypes.data, nt)
        return out

    t0 = time.perf_counter()
    invs = osys.block_jacobi_inverses(A, mesh, mask)
    pos = np.full(mask.size, -1, dtype=np.int64)
    pos[I] = np.arange(I.size)
    idx = np.ascontiguousarray(pos[np.stack(osys.bj_block_dofs(mesh))])
    inv = np.ascontiguousarray(np.stack(invs))
    del invs
    t_bj = time.perf_counter() - t0
    Aop = spla.LinearOperator(A_II.shape, matvec=matvec, dtype=np.float64)
    M = spla.LinearOperator(A_II.shape, matvec=lambda r: build_c.bj_apply(lib, idx, inv, r, nt),
                            dtype=np.float64)
    its = [0]

    def count(_):
        its[0] += 1

    nb = np.linalg.norm(rhs)
    x = np.ascontiguousarray(f[I])
    rtol, res, rounds = 1e-8, np.inf, 0
    t0 = time.perf_counter()
    while rounds < 5:
        x, _ = spla.gmres(Aop, rhs, x0=x, rtol=rtol, atol=0.0, restart=50, maxiter=5000,
                          M=M, callback=count, callback_type="pr_norm")
        rounds += 1
        res = float(np.linalg.norm(rhs - matvec(x)) / nb)
        if res <= 1e-8:
            break
        rtol *= 0.1
    t_solve = time.perf_counter() - t0
    out = {"workload": "uniform absorber 16x16x8 cells, p=4 ({} DOF), C2's coefficients"
                       .format(mask.size),
           "cpu": {"seconds": t_bj + t_solve, "setup_seconds": t_bj, "solve_seconds": t_solve,
                   "assembly_seconds": t_asm, "iterations": its[0], "rel_residual": res,
                   "converged": res <= 1e-8, "cores": nt,
                   "method": "scipy GMRES(50) + the oracle's block-Jacobi, CSR product and "
                             "block apply in C (OpenMP, {} threads)".format(nt)}}
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    grhs = _rhs(op, wl)
    for prec in ("fp64", "mixed"):
        xg, out[prec] = _timed_solve(op, grhs, prec)
        xg = xg.cpu().numpy()[I]
        out[prec]["rel_l2_vs_cpu"] = float(np.linalg.norm(xg - x) / np.linalg.norm(x))
        out["speedup_{}_vs_cpu".format(prec)] = out["cpu"]["seconds"] / out[prec]["seconds"]
        out["speedup_{}_vs_cpu_warm".format(prec)] = t_solve / out[prec]["solve_seconds_warm"]
    del op
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the operator and solver have no multi-GPU "
              "path (--gpus {} requested)".format(args.gpus), file=sys.stderr)
        sys.exit(2)
    if args.profile_apply:
        profile_apply(args)
        return
    import numpy as np
    import torch
    from rtx.operator import TransportOperator, OP_SYSTEM, OP_RESTRICT, PREC_F64, PREC_F32, PREC_MIXED
    from rtx import krylov

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    mesh, wl = workload()
    t0 = time.perf_counter()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi, device=dev)
    t_plan = time.perf_counter() - t0
    n = op.ndof
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    x = torch.rand(n, dtype=torch.float64, device=dev, generator=gen) * 2 - 1
    op.mask(x)
    y = torch.empty_like(x)
    ops = OP_SYSTEM | OP_RESTRICT
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        op.apply(x, y, ops=ops)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        op.apply(x, y, ops=ops)
    ev1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall0
    ms = ev0.elapsed_time(ev1) / args.steps
    bytes_, flops = algorithmic(op)
    gbs = bytes_ / (ms * 1e-3) / 1e9
    tflops = flops / (ms * 1e-3) / 1e12
    # the binding resource: the larger of bytes / HBM peak and flops / FP64 peak
    fp64_bound = flops / (FP64_PEAK_TFLOPS * 1e12) > bytes_ / (HBM_PEAK_GBS * 1e9)
    if args.calib:
        _calibration_copies(op, dev)

    out = {
        "metric": "time-to-1e-8 residual & DOF/s per apply: learned+mixed vs plain FP64, 1 GPU",
        "value": n / (ms * 1e-3), "unit": "DOF/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "uniform absorber 32x32x16 cells, p=4 (2,048,000 DOF), HG g=0.8, "
                               "sigma/kappa=0.7", "ndof": n, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma" if fp64_bound else "hbm",
                     "achieved": tflops if fp64_bound else gbs,
                     "peak": FP64_PEAK_TFLOPS if fp64_bound else HBM_PEAK_GBS,
                     "unit": "TFLOP/s" if fp64_bound else "GB/s",
                     "frac": tflops / FP64_PEAK_TFLOPS if fp64_bound else gbs / HBM_PEAK_GBS,
                     "traffic": measured_traffic(),
                     "kernel": "k_uniform_e (operator apply A_II, FP64, output layout, order bin 5x5x16)",
                     "algorithmic_bytes": bytes_, "algorithmic_flops": flops,
                     "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
                     "fp64_tflops": tflops, "fp64_frac": tflops / FP64_PEAK_TFLOPS},
        "plan_seconds": t_plan, "timed_wall_seconds": t_wall,
    }
    if not args.no_solve:
        out["solve"] = solve_section(op, mesh, wl)
    if not args.no_indicators:
        out["indicators"] = indicators_section(op, mesh, wl, cpu=not args.no_cpu)
    if not args.no_learned or not args.no_hp:
        del op
        torch.cuda.empty_cache()
    if not args.no_hp:
        out["hp"] = hp_section(fields=args.hp_fields,
                               log=lambda m: print(m, file=sys.stderr, flush=True))
        torch.cuda.empty_cache()
    if not args.no_learned:
        out["learned"] = [learned_section(fields=args.learned_fields, kappa_scale=float(ks),
                                          log=lambda m: print(m, file=sys.stderr, flush=True))
                          for ks in args.learned_kappa_scales.split(",") if ks.strip()]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, wl)
        if not args.no_solve:
            out["cpu_baseline"]["solve"] = cpu_solve_baseline()
    print(json.dumps(out), flush=True)


def indicators_section(op, mesh, wl, reps=20, cpu=True):
    """The error-indicator kernels on the same mesh, on a random nodal field:
    average device time per call (HIP events on the launch stream) and the
    algorithmic bytes each reads; beside them, the oracle's loops on a
    4 x 4-column mesh of the same order (per DOF)."""
    import ctypes as C
    import numpy as np
    import torch
    from rtx.error import ErrorEvaluator
    from rtx.operator import stream_handle
    t0 = time.perf_counter()
    ev = ErrorEvaluator(mesh, device=op.device, tables=op.t)
    t = ev.t
    u = torch.rand(t.ndof, dtype=torch.float64, device=op.device)
    start, pair, scale, pool = ev.jump_pairs()
    ev.spatial_jump(u)                          # uploads the pair tables
    ev.angular_jump(u)
    t_setup = time.perf_counter() - t0
    vp = lambda x: C.c_void_p(x.data_ptr())
    f64 = lambda n: torch.empty(max(int(n), 1), dtype=torch.float64, device=op.device)
    xy, th, n2, ce, cc = f64(t.nang), f64(t.nsp), f64(t.ncell), f64(t.ncell), f64(t.ncol)
    hr = torch.rand(int(ev.hr_off[-1]), dtype=torch.float64, device=op.device)
    ds, dp, dsc, dpl = ev._d_jump
    st = stream_handle()
    calls = {
        "moments": lambda: ev.lib.rtx_indicator_moments(ev.plan, vp(u), vp(xy), vp(th), vp(n2), st),
        "angular_jump": lambda: ev.lib.rtx_error_angular_jump(ev.plan, vp(u), vp(ev._ang_nbr),
                                                              vp(ev._d_lend), vp(ce), st),
        "spatial_jump": lambda: ev.lib.rtx_error_spatial_jump(ev.plan, vp(u), vp(ds), vp(dp),
                                                              vp(dsc), vp(dpl), vp(cc), st),
        "analytic_l2": lambda: ev.lib.rtx_error_l2(ev.plan, vp(u), vp(hr), vp(ev.d_hr_off),
                                                   vp(ev.d_tab), vp(ce), st),
    }
    P = pair.reshape(-1, 9)
    c0, c1 = t.cell_col[P[:, 0]], t.cell_col[P[:, 1]]
    n0 = np.where(P[:, 2] % 2 == 0, t.col_ny[c0], t.col_nx[c0])
    n1 = np.where(P[:, 2] % 2 == 0, t.col_ny[c1], t.col_nx[c1])
    trace_vals = int(np.sum(n0 * t.cell_nth[P[:, 0]] + n1 * t.cell_nth[P[:, 1]]))
    alg = {"moments": 16 * t.ndof, "angular_jump": 8 * t.ndof,
           "spatial_jump": 8 * trace_vals, "analytic_l2": 8 * (t.ndof + int(ev.hr_off[-1]))}
    stream = torch.cuda.current_stream()
    res = {"workload": "C2 mesh, random nodal field", "setup_seconds": t_setup,
           "spatial_jump_pairs": int(P.shape[0])}
    for name, fn in calls.items():
        for _ in range(3):
            _lib_check(fn(), name)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / reps * 1e3
        res[name] = {"us": us, "algorithmic_bytes": alg[name],
                     "gbs": alg[name] / (us * 1e-6) / 1e9}
    if cpu:
        res["cpu_oracle"] = _indicator_cpu_sample()
    return res


def _lib_check(rc, what):
    if rc != 0:
        raise RuntimeError("{} failed ({})".format(what, rc))


def _indicator_cpu_sample():
    """The oracle's loops (angular and spatial jumps) on a 4 x 4-column,
    16-cell, p = 4 mesh: seconds per DOF, one thread."""
    import numpy as np
    from oracle import indicator as oind
    from rtx import problems
    m = problems.uniform_mesh([1.0, 1.0], [5, 5, 5], nref_ang=4, nref_spt=2)
    n = m.get_ndof()
    u = np.random.default_rng(0).uniform(-1, 1, n)
    out = {"sample": "4 x 4 columns x 16 cells, p = 4 ({} DOF), oracle/indicator.py, one thread".format(n),
           "cores": 1}
    for name, fn in (("angular_jump", oind.angular_jump), ("spatial_jump", oind.spatial_jump)):
        t0 = time.perf_counter()
        fn(m, u)
        dt = time.perf_counter() - t0
        out[name] = {"seconds": dt, "dof_per_s": n / dt}
    return out


def _rhs(op, wl):
    """rhs = f_I - A_IB b_B (full-length, zero on the inflow boundary), as in
    rt.Problem.solve."""
    import torch
    from rtx.operator import OP_SYSTEM, OP_QUADW
    f = op.apply(torch.from_numpy(op.nodal(wl.f)).to(op.device), ops=OP_QUADW)
    b = torch.from_numpy(op.nodal(wl.bcs)).to(op.device)
    bb = op.mask(b.clone(), keep_interior=False)
    return op.mask(f - op.apply(bb, ops=OP_SYSTEM), keep_interior=True)


def _timed_solve(op, rhs, precision, make_pc=None, inner_rtol=1e-2):
    """Set up a solver session (block-Jacobi factorisation, and the coarse
    preconditioner when ``make_pc(bj)`` is given), then solve twice: the
    first ("cold") solve also runs each chunk of iterations eagerly once and
    captures its graph, the second ("warm") replays the graphs, as a repeated
    solve (an adaptive loop) would.  ``seconds`` = setup + cold solve."""
    import torch
    from rtx import learned
    from rtx.krylov import BlockJacobi
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bj = BlockJacobi(op, keep_f64=precision == "fp64", keep_fp4=precision != "fp64")
    pc = make_pc(bj) if make_pc is not None else None
    # fp64: GMRES(50) as the reference's KSP; mixed: inner GMRES(15) on the FP16 basis
    sess = learned.Session(op, pc=pc, bj=bj, precision=precision)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    out = {}
    for tag in ("cold", "warm"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, info = sess.solve(rhs, rtol=1e-8, inner_rtol=inner_rtol)
        torch.cuda.synchronize()
        out[tag] = (time.perf_counter() - t0, x, info)
    t_cold, _, info_c = out["cold"]
    t_warm, x, info = out["warm"]
    return x, {"seconds": t_setup + t_cold, "setup_seconds": t_setup,
               "solve_seconds_cold": t_cold, "solve_seconds_warm": t_warm,
               "iterations": int(info["iteration_number"]),
               "rel_residual": float(info["rel_residual"]),
               "converged": bool(info["converged_reason"] > 0),
               "fallback": bool(info.get("fallback", False)),
               "gated": bool(info.get("gated", False))}


def _speedups(res, a, b, key):
    res["speedup_{}_vs_{}".format(a, b)] = res[b]["seconds"] / res[a]["seconds"]
    res["speedup_{}_vs_{}_warm".format(a, b)] = \
        res[b]["solve_seconds_warm"] / res[a]["solve_seconds_warm"]


def solve_section(op, mesh, wl):
    """Time-to-1e-8 true FP64 relative residual on the apply workload's
    interior system: plain FP64 GMRES(50) + block-Jacobi vs FP64 iterative
    refinement around FP32 / BF16 applies with BF16 block-Jacobi inverses.
    Times include the block-Jacobi factorisation."""
    rhs = _rhs(op, wl)
    res = {}
    x64, res["fp64"] = _timed_solve(op, rhs, "fp64")
    xm, res["mixed"] = _timed_solve(op, rhs, "mixed")
    res["mixed"]["rel_l2_vs_fp64"] = float((xm - x64).norm() / x64.norm())
    _speedups(res, "mixed", "fp64", "seconds")
    return res


def _learned_solves(op, mesh, field, eval_wl, fields, log=None, ncoarse=64, epochs=100):
    """Train the coarse model on solver pairs of ``fields`` seeded members of
    ``field`` on ``op``'s mesh, then time FP64, mixed and learned+mixed
    solves of ``eval_wl`` (coefficients set on ``op``) to an FP64 residual of
    1e-8.  Returns the timing record."""
    import torch
    from rtx import learned
    from rtx._lib import PREC_MIXED
    grid = learned.CoarseGrid(op, ncoarse)
    t0 = time.perf_counter()
    X, Y = learned.generate_pairs(op, grid, range(fields), cycles=20, restart=10, skip=1,
                                  field=field, direction_stride=2, log=log)
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t0
    nc = grid.nc
    mlp = learned.MLP([5 * nc, 128, 128, 128, nc], op.device, seed=0)
    mlp.n_moments = 3 * nc
    t0 = time.perf_counter()
    xs, ys, hist = learned.train(mlp, X, Y, epochs=epochs, batch=64, lr=1e-3, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    npairs = int(X.shape[0])
    del X, Y
    op.set_coeffs(eval_wl.kappa, eval_wl.sigma)
    rhs = _rhs(op, eval_wl)
    res = {"pairs": npairs, "fields": fields, "pair_generation_seconds": t_pairs,
           "training_seconds": t_train, "validation_rel_error": hist[-1][1]}
    x64, res["fp64"] = _timed_solve(op, rhs, "fp64")
    xm, res["mixed"] = _timed_solve(op, rhs, "mixed")

    def make_pc(bj):
        model = learned.CoarseModel(mlp, 3 * nc, xs, ys)
        model.validation_error = hist[-1][1]
        model.set_coefficients(grid.coefficient_channels(eval_wl.kappa, eval_wl.sigma))
        return learned.TwoLevelPC(op, bj, grid, model, apply_precision=PREC_MIXED)

    xl, res["learned_mixed"] = _timed_solve(op, rhs, "mixed", make_pc=make_pc)
    for k, x in (("mixed", xm), ("learned_mixed", xl)):
        res[k]["rel_l2_vs_fp64"] = float((x - x64).norm() / x64.norm())
    _speedups(res, "learned_mixed", "fp64", "seconds")
    _speedups(res, "mixed", "fp64", "seconds")
    return res


def hp_section(fields=17, log=None):
    """The north star's target workload: the batched uneven-order hp mesh
    (16 x 16 x 8 base, seeded random h-refinement, node counts 3..9:
    8,263,762 DOF) with the two strong scatterers (kappa = 1 + 9 max(g1, g2),
    sigma / kappa = 0.1 + 0.89 max(g1, g2), HG g = 0.9, unit inflow on the
    left wall).  Time to an FP64 relative residual of 1e-8: plain FP64
    GMRES(50) + block-Jacobi, mixed refinement, and learned+mixed with the
    coarse model trained on ``fields`` seeded members of the two-scatterer
    family (blob centres, radius, strength and ratios at random).  Also the
    default (hp column kernel) and generic FP64 / mixed apply times."""
    import torch
    from rtx import problems
    from rtx.operator import TransportOperator, OP_SYSTEM, OP_RESTRICT, PREC_F64, PREC_MIXED
    t0 = time.perf_counter()
    mesh = problems.random_hp_mesh(0)
    wl = problems.two_scatterers_c3()
    op = TransportOperator(mesh, wl.kappa, wl.sigma, wl.Phi)
    t_setup = time.perf_counter() - t0
    res = {"workload": "two strong scatterers, hp mesh 16x16x8 base + random h-refinement, "
                       "node counts 3..9 ({} DOF)".format(op.ndof),
           "ndof": op.ndof, "setup_seconds": t_setup, "apply_us": {}}
    stream = torch.cuda.current_stream()
    x = torch.rand(op.ndof, dtype=torch.float64, device=op.device) * 2 - 1
    op.mask(x)
    for path, pname in ((-1, "default"), (0, "generic")):
        op.set_path(path)
        for prec, xx in ((PREC_F64, x), (PREC_MIXED, x.float())):
            y = torch.empty_like(xx)
            for _ in range(3):
                op.apply(xx, y, ops=OP_SYSTEM | OP_RESTRICT, precision=prec)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(20):
                op.apply(xx, y, ops=OP_SYSTEM | OP_RESTRICT, precision=prec)
            e1.record(stream)
            torch.cuda.synchronize()
            res["apply_us"]["{}_{}".format(pname, "f64" if prec == PREC_F64 else "mixed")] = \
                e0.elapsed_time(e1) / 20 * 1e3
    op.set_path(-1)
    res.update(_learned_solves(op, mesh, problems.two_scatterer_field, wl, fields, log=log))
    del op
    return res


def learned_section(fields=17, eval_seed=10_000, kappa_scale=1.0, log=None):
    """The learned coarse preconditioner on the random-field workload: 64 x 64
    columns x 16 angular cells, p = 4 (8,192,000 DOF), log kappa ~ GRF
    (length 0.15, mean log kappa_scale, variance 1), sigma / kappa ~
    U(0.9, 0.999), HG g = 0.5, unit source, vacuum inflow.  Pairs from
    ``fields`` seeded fields (solver-generated), a 4-layer width-128 MLP
    trained on the device, then fp64 / mixed / learned+mixed solves of an
    unseen field (seed ``eval_seed``) to an FP64 residual of 1e-8."""
    from rtx import problems
    from rtx.operator import TransportOperator
    t0 = time.perf_counter()
    mesh = problems.uniform_mesh([1.0, 1.0], [5, 5, 5], nref_ang=4, nref_spt=6)

    def field(s):
        return problems.grf_scatterer(s, n=64, kappa_scale=kappa_scale)

    wl0 = field(0)
    op = TransportOperator(mesh, wl0.kappa, wl0.sigma, wl0.Phi)
    t_setup = time.perf_counter() - t0
    wl = field(eval_seed)
    res = {"workload": "grf scatterer 64x64x16 cells, p=4 (8,192,000 DOF), kappa_scale {}, "
                       "seed {}, sigma/kappa {:.4f}".format(kappa_scale, eval_seed, wl.ratio),
           "setup_seconds": t_setup}
    res.update(_learned_solves(op, mesh, field, wl, fields, log=log))
    return res


if __name__ == "__main__":
    main()
