"""Benchmark of the ADER-DG elastic tet4 element update on one MI355X.

Workload (the headline): 32^3 hexes x 6 = 196,608 tets, interior vertices
jittered by +-10% (seeded), per-element rho ~ U[2,3], vp ~ U[4,6],
vs = vp / sqrt(3), degree 4 (order 5, 35 modes), 16 fused runs in FP32,
global time stepping at the CFL time step.  Beside it the plain FP64
single-run update on the same mesh is timed (the drop-in), the FP32
single-run update, and the FP64 oracle (the vanilla-semantics CPU
restatement) on a bounded sample of the same workload on the host cores.

A step of the fused-16 run is the reference's (src/impl/seismic/steps.inc):
the receivers' time prediction at the step start (64 receivers), one local
pass over all elements, the point sources over the step (one Ricker
moment-tensor source per run), one neighbour pass.  Inputs are resident in
HBM before the timed region.  Prints one JSON line.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DOF-updates/s, tet4 elastic O5 FP32 fused-16 vs FP64 plain; % roofline"
PEAK = {"fp32_mfma": 157.3e12, "fp64": 78.6e12, "hbm": 8.0e12}

# Dense algorithmic FLOPs per element, step and run (vanilla GEMM shapes,
# src/impl/seismic/kernels/*Vanilla.hpp): local = CK (+axpy) + volume + local
# surface, neigh = neighbouring surface, proj = the neighbouring surface's
# first product, tInt times the neighbour flux matrix (SurfIntVanilla.hpp
# neighFluxInt, 4 faces x 2 x 9 x N x NF).  When the local kernel writes the
# face projections, it performs proj (as the reference's local step does for
# faces whose neighbour is on another rank, AderDg.hpp:419-473), so proj counts
# towards the local kernel and not the neighbour kernel.
FLOPS = {4: {"local": 20466 + 630 + 20520 + 35280, "neigh": 35280, "proj": 4 * 2 * 9 * 20 * 10},
         5: {"local": 67986 + 1260 + 54810 + 85320, "neigh": 85320, "proj": 4 * 2 * 9 * 35 * 15}}


def alg_bytes(order, runs, s, s_c):
    """Compulsory bytes per element-step: DOFs read+write, tInt write+read,
    dense star + two flux-solver sets, connectivity."""
    n = order * (order + 1) * (order + 2) // 6
    return s * 9 * n * runs * 4 + s_c * (3 * 81 + 2 * 4 * 81) + 80


def local_bytes(order, runs, prec=4, proj=False):
    """Compulsory bytes per element of one local launch: DOFs read and
    written, tInt written, dense star matrices and local flux solvers read,
    and with proj the face projections written (4 faces x 3 x 16 x 16)."""
    n = order * (order + 1) * (order + 2) // 6
    return 3 * 9 * n * runs * prec + (3 * 81 + 4 * 81) * prec + (4 * 3 * 256 * prec if proj else 0)


def launch_bytes(order, runs, variant, prec=4):
    """Compulsory bytes per element of the dominant FP32 launch: the local
    kernel (local_bytes), or the fused one, which also reads the face
    projections its neighbours wrote and the neighbour-side flux solvers."""
    if variant != "fused":
        return local_bytes(order, runs, prec, proj=variant == "proj")
    return local_bytes(order, runs, prec, proj=True) + 4 * 3 * 256 * prec + 4 * 81 * prec + 80


# FP32 variants: (projections exchanged, XCD-aware schedule, fused multi-step)
FP32_VARIANTS = {"fused": (True, True, True), "proj": (True, True, False), "tiles": (False, True, False)}

# HBM traffic of the dominant FP32 kernel per variant
TRAFFIC_PROFILE = {"fused": os.path.join(ROOT, "profiles", "traffic_k_local_fused_o5_fp32_r16_196608tets.md"),
                   "proj": os.path.join(ROOT, "profiles", "traffic_k_local_o5_fp32_r16_196608tets.md"),
                   "tiles": os.path.join(ROOT, "profiles", "traffic_k_local_tiles_o5_fp32_r16_196608tets.md")}
# MFMA utilisation (SQ_VALU_MFMA_BUSY_CYCLES over GRBM_GUI_ACTIVE x SIMDs) per kernel
MFMA_PROFILE = os.path.join(ROOT, "profiles", "mfma_util_o5_196608tets.md")


def dom_kernel_name(variant, order):
    return {"fused": f"k_local<{order}, true, true, true, true>", "proj": f"k_local<{order}, true, true, true, false>",
            "tiles": f"k_local<{order}, true, true, false, false>"}[variant]


def read_mfma_util(path, n_el, order, runs, kernel):
    """The MFMA busy fraction of `kernel` from a committed rocprofv3 counter
    summary (a table row `| kernel | ... | util |`, whose last column is the
    busy fraction), or None when absent or for another workload."""
    try:
        with open(path) as f:
            text = f.read()
        m = re.search(r"elements\s*=\s*(\d+)\s*\n\s*order\s*=\s*(\d+)", text)
        if not m or (int(m.group(1)), int(m.group(2))) != (n_el, order):
            return None
        for line in text.splitlines():
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 3 and cells[0].strip("`") == kernel and cells[1] == str(runs):
                return float(cells[-1])
    except (OSError, ValueError):
        return None
    return None


def read_traffic(path, n_el, order, runs):
    """HBM bytes per launch of the local kernel from a committed rocprofv3
    summary: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by
    bytes / counter for a copy kernel of known size.  None when the summary
    is absent or describes another workload."""
    try:
        kv = {}
        with open(path) as f:
            for line in f:
                m = re.match(r"^\s*([a-z_0-9]+)\s*=\s*([-+0-9.eE]+)\s*$", line)
                if m:
                    kv[m.group(1)] = float(m.group(2))
        if (int(kv["elements"]), int(kv["order"]), int(kv["runs"])) != (n_el, order, runs):
            return None
        return (kv["local_fetch_kib"] * kv["fetch_scale"] + kv["local_write_kib"] * kv["write_scale"]) * 1024.0
    except (OSError, KeyError, ValueError):
        return None


def copy_calibration(mib):
    """A plain copy kernel of known size (an elementwise negation: reads and
    writes `mib` MiB once each), for scaling the FETCH_SIZE / WRITE_SIZE
    counters to bytes."""
    import torch
    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 5,
                                           "ok": bool((dst == -1).all())}}))


def build_problem(n_hex, order, seed=3343):
    from edge_amd import mesh as M
    from edge_amd import setups as S
    m = M.box(n_hex, jitter=0.1, seed=seed, block=4)
    rho, lam, mu = S.random_materials(m.n_el, seed=seed)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    # a P plane wave of each element's medium: wave vector 2 pi (1, 1, 1)
    q0 = S.project(m, order, S.PlaneWave(k=(2 * np.pi,) * 3), chunk=4096,
                   materials=(rho, lam, mu))[..., 0]  # [el][9][N]
    return m, rho, lam, mu, dt, q0


def point_setup(solver, m, runs, seed=3343, n_recv=64):
    """One Ricker point source per run (peak frequency U[1, 5], a random
    symmetric moment tensor with N(0, 1) entries, at a seeded point) and
    n_recv seeded receivers recording every run at every step start."""
    from edge_amd.points import PointSources, Receivers, moment_tensor_scalings, ricker
    rng = np.random.default_rng(seed + 1)
    lo, hi = m.ve_crds.min(axis=0), m.ve_crds.max(axis=0)
    pts = lo + (hi - lo) * rng.uniform(0.05, 0.95, size=(runs, 3))
    f0 = rng.uniform(1.0, 5.0, runs)
    dts = 1e-3 / f0  # 1,000 samples per period
    t0 = -3.0 / f0
    series = [ricker(f0[r], 0.0 - t0[r], dts[r], int(6.0 / f0[r] / dts[r]) + 1) for r in range(runs)]
    scas = []
    for r in range(runs):
        a = rng.normal(size=(3, 3))
        scas.append(moment_tensor_scalings(np.triu(a) + np.triu(a, 1).T))
    src = PointSources(solver, pts, np.arange(runs), np.array(scas), t0, dts, series)
    recv = Receivers(solver, lo + (hi - lo) * rng.uniform(0.05, 0.95, size=(n_recv, 3)), freq=1.0)
    return src, recv


def time_solver(solver, dofs, tint, dt, steps, warmup, points=None, fused=False):
    """Times exactly `steps` steps between synchronisations, after `warmup`
    untimed ones; per-kernel durations from HIP events recorded around each
    launch on the stream the kernels run on.  With points = (sources,
    receivers), every step is the reference's (src/impl/seismic/steps.inc):
    receivers at the step start, local, sources over the step, neigh.
    fused: the steps run as local, neigh_local x (steps - 1), neigh (one
    launch per step, the neighbour step of step k - 1 with the local step of
    step k; the sources between the launches, the receivers on the
    step-start DOFs the fused launches write to taps).
    Returns (wall seconds, {"local" | "fused" | "neigh": mean seconds})."""
    import torch
    stream = torch.cuda.current_stream()
    src, recv = points if points is not None else (None, None)
    out = taps = tau = None
    if recv is not None:
        out = torch.empty((recv.n, 1, 9, solver.n_runs), dtype=dofs.dtype, device=dofs.device)
        tau = np.zeros((recv.n, 1))
        if fused:
            taps = torch.zeros_like(dofs)
            solver.set_taps(recv.el, taps)

    def run(n, i0, rec):
        def timed(name, fn):
            if rec is None:
                fn()
                return
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            rec.setdefault(name, []).append((a, b))

        for k in range(n):
            t = (i0 + k) * dt
            if not fused or k == 0:
                if recv is not None:
                    recv.eval_into(dofs, tau, out, stream)
                timed("local", lambda: solver.local(dofs, tint, dt, stream=stream))
            else:
                timed("fused", lambda: solver.neigh_local(dofs, tint, dt, stream=stream))
                if recv is not None:
                    recv.eval_into(taps, tau, out, stream)
            if src is not None:
                src.apply(t, t + dt, dofs, stream)
            if not fused:
                timed("neigh", lambda: solver.neigh(dofs, tint, stream=stream))
        if fused:
            timed("neigh", lambda: solver.neigh(dofs, tint, stream=stream))

    if warmup > 0:
        run(warmup, 0, None)
    torch.cuda.synchronize()
    rec = {}
    t0 = time.perf_counter()
    run(steps, warmup, rec)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if taps is not None:
        solver.set_taps(None)
    return wall, {k: float(np.mean([a.elapsed_time(b) for a, b in v])) * 1e-3 for k, v in rec.items()}


def host_cpu():
    """(CPU model, logical CPUs this process may run on, physical cores among them)."""
    model, cores = "?", set()
    try:
        allowed = os.sched_getaffinity(0)
    except AttributeError:
        allowed = set(range(os.cpu_count() or 1))
    try:
        cpu, phys, core = None, None, None
        with open("/proc/cpuinfo") as f:
            for line in f.read().splitlines() + [""]:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    cpu = int(v)
                elif k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and cpu is not None:
                    if cpu in allowed:
                        cores.add((phys, core))
                    cpu, phys, core = None, None, None
    except OSError:
        pass
    return model, len(allowed), len(cores) or len(allowed)


def omp_threads(n):
    """Sets the OpenMP thread count of the oracle's parallel loops."""
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
    except OSError:
        pass


def config1_single_thread(budget_s=8.0):
    """The oracle on the reference's CPU-runnable case: unit cube of 8^3 hexes
    x 6 = 3,072 tets, periodic, degree 3 (order 4, 20 modes), FP64, one run,
    rho = 1, lambda = 2, mu = 1, P plane wave, global time stepping at the
    CFL step; one thread, as many of its 100 steps as fit the budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from edge_amd import mesh as M
    from edge_amd import setups as S
    m = M.box(8)
    one = np.ones(m.n_el)
    order = 4
    dt = S.dt_cfl(m, one, 2 * one, one, order)
    d = S.project(m, order, S.PlaneWave(k=(2 * np.pi,) * 3))
    t = np.zeros_like(d)
    mo = orc.MeshOracle(m, order, 1, rho=one, lam=2 * one, mu=one)
    omp_threads(1)
    t0 = time.perf_counter()
    steps = 0
    while steps < 100 and time.perf_counter() - t0 < budget_s:
        mo.step(d, t, dt, 1)
        steps += 1
    wall = time.perf_counter() - t0
    omp_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    n = S.n_modes(order)
    return {"value": m.n_el * 9 * n * steps / wall, "unit": "DOF-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} of the 100 steps of the 3,072-tet O4 FP64 single-run plane-wave case, {wall:.1f} s, "
                      f"one thread"}


def cpu_baseline(m, order, runs, rho, lam, mu, dt, q0, budget_s=15.0):
    """FP64 oracle (vanilla-semantics CPU restatement; LIBXSMM is not
    available offline) on a bounded element sample, all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    star, fs, sample = orc.setup_mesh_sample(m, rho, lam, mu, n_sample=4096)
    n_s = star.shape[0]
    mo = orc.SampleOracle(order, runs, star, fs, sample)
    dofs = np.ascontiguousarray(np.repeat(q0[:n_s, ..., None], runs, axis=-1))
    tint = np.zeros_like(dofs)
    t0 = time.perf_counter()
    n_el = 0
    reps = 0
    while time.perf_counter() - t0 < budget_s and reps < 64:
        chunk = 256
        for e0 in range(0, n_s, chunk):
            mo.local(dofs, tint, dt, e0, min(chunk, n_s - e0))
            mo.neigh(dofs, tint, e0, min(chunk, n_s - e0))
            n_el += min(chunk, n_s - e0)
            if time.perf_counter() - t0 > budget_s:
                break
        reps += 1
    wall = time.perf_counter() - t0
    n = order * (order + 1) * (order + 2) // 6
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model, logical, physical = host_cpu()
    return {"value": n_el * 9 * n * runs / wall, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "threads": threads, "cpu_model": model, "logical_cpus": logical, "physical_cores": physical,
            "sample": f"{n_el} element-steps (local+neigh) of the {m.n_el}-tet O{order} R={runs} FP64 workload, "
                      f"{wall:.1f} s, OpenMP over elements with {threads} threads on {physical} physical cores "
                      f"({model}); vanilla-semantics restatement, LIBXSMM unavailable",
            "config1_single_thread": config1_single_thread()}


def visco_flops(order, n_rms):
    """Dense vanilla FLOPs per element-step and run of the viscoelastic
    update (TimePredVanilla / VolIntVanilla / SurfIntVanilla with N_RMS > 0,
    no CK mode truncation): (local, neigh)."""
    n = order * (order + 1) * (order + 2) // 6
    nf = order * (order + 1) // 2
    nv = (order - 1) * order * (order + 1) // 6
    ck_stage = 3 * (2 * 9 * n * n + 2 * 81 * n + 2 * 18 * n) + n_rms * (2 * 36 * n + 2 * 6 * n) + 2 * (9 + 6 * n_rms) * n
    ck = (order - 1) * ck_stage
    vol = 3 * (2 * 9 * nv * n + 2 * 81 * n + 2 * 18 * n) + n_rms * (2 * 36 * n + 3 * 6 * n)
    face = 2 * 9 * n * nf + 2 * 81 * nf + 2 * 9 * nf * n + 2 * 54 * nf + 2 * 6 * nf * n
    return ck + vol + 4 * face + n_rms * 2 * 6 * n, 4 * face + n_rms * 2 * 6 * n


def ensemble_row(args, timed, flops_el_run):
    """24 independent viscoelastic meshes (3 mechanisms, Q_p in U[50, 200],
    Q_s = Q_p / 2, f_c = 1 Hz, ratio 10) advanced as one mesh: FP32 with 8
    runs on the MFMA kernels, beside FP64 with one run."""
    import torch
    from edge_amd import mesh as M
    from edge_amd import setups as S
    from edge_amd.solver import ViscoSolver
    rng = np.random.default_rng(3343)
    sizes = [int(x) for x in rng.integers(8, 33, 24)]
    meshes, mats = [], []
    for i, n in enumerate(sizes):
        m = M.box(n, jitter=0.1, seed=3343 + i)
        r = np.random.default_rng(10_000 + i)
        rho = r.uniform(2.0, 3.0, m.n_el)
        vp = r.uniform(4.0, 6.0, m.n_el)
        mu = rho * (vp / np.sqrt(3.0)) ** 2
        qp = np.full(m.n_el, r.uniform(50, 200))
        meshes.append(m)
        mats.append((rho, rho * vp ** 2 - 2 * mu, mu, qp))
    u, ranges = M.union(meshes)
    rho, lam, mu, qp = [np.concatenate([x[k] for x in mats]) for k in range(4)]
    for runs, prec in ((8, 32), (1, 64)):
        t0 = time.time()
        s = ViscoSolver(u, args.order, runs, prec, rho, lam, mu, qp, qp / 2, 3, 1.0, 10.0)
        t_setup = time.time() - t0
        dt = S.dt_cfl(u, rho, s.setup_a()["lam_e"], s.setup_a()["mu_e"], args.order)
        tune = [] if args.no_autotune else s.autotune(dt, 2)
        d = torch.randn(s.dof_shape(), dtype=s.dtype, device="cuda")
        a = 0.01 * torch.randn(s.dof_a_shape(), dtype=s.dtype, device="cuda")
        t = torch.zeros_like(d)
        # one graph-captured multi-step call (fused launches where the solver has them)
        n_steps = 6
        ms = timed(lambda st: s.step(d, t, dt, n_steps, a, graph=True, stream=st), 1) / n_steps
        peak = PEAK["fp32_mfma"] if prec == 32 else PEAK["fp64"]
        n = s.n_modes
        print(json.dumps({"row": "batched viscoelastic ensemble", "meshes": len(sizes), "hexes_per_axis": sizes,
                          "elements": u.n_el, "order": args.order, "runs": runs, "precision": prec, "n_rms": 3,
                          "kernel": s.kernel_info()[0], "hi_occ": bool(s.kernel_info()[1] & 32),
                          "fused": bool(s.kernel_info()[1] & 16),
                          "autotune": [{"kernel": k, "projections": p, "xcd_schedule": x, "fused": f, "occupancy": o,
                                        "ms_per_step": c} for k, p, x, c, f, o in tune],
                          "graph": True, "steps_per_call": n_steps, "setup_s": round(t_setup, 1),
                          "ms_per_step": ms * 1e3, "dof_updates_per_s": u.n_el * 27 * n * runs / ms,
                          "roofline_frac": flops_el_run * runs * u.n_el / ms / peak}), flush=True)
        s.close()
        del d, a, t
        torch.cuda.empty_cache()


def fault_problem(order, runs, seed=3343):
    """48 x 48 x 24 hexes x 6 = 331,776 tets on the unit cube (periodic in x
    and y, outflow in z) with a planar fault at z = 0.5 (4,608 faces):
    rho = 2.67, vp = 6, vs = 3.464; linear slip weakening with mu_s = 0.677,
    mu_d = 0.525, d_c = 0.4, normal stress -120; per run an initial shear
    stress U[70, 81.24] and a seeded square nucleation patch (half-width 0.15)
    at 81.6, above the static strength 81.24."""
    from edge_amd import fault as FL
    from edge_amd import mesh as M
    from edge_amd import setups as S
    from edge_amd.mesh import FA_OUTFLOW, FA_PERIODIC
    m = M.box((48, 48, 24), bc=(FA_PERIODIC, FA_PERIODIC, FA_OUTFLOW), seed=seed)
    faces = FL.tag_plane(m, 2, 0.5)
    rho = np.full(m.n_el, 2.67)
    mu = rho * 3.464 ** 2
    lam = rho * 6.0 ** 2 - 2 * mu
    rng = np.random.default_rng(seed + 4)
    nsf = FL.n_sub_faces(order)
    ss0 = np.broadcast_to(rng.uniform(70.0, 81.24, runs)[None, None, :], (len(faces), nsf, runs)).copy()
    pts = FL.sub_face_points(m, order)
    ctr = rng.uniform(0.3, 0.7, (runs, 2))
    for r in range(runs):
        inside = (np.abs(pts[..., 0] - ctr[r, 0]) < 0.15) & (np.abs(pts[..., 1] - ctr[r, 1]) < 0.15)
        ss0[..., r][inside] = 81.6
    glob = np.array([[0.677] * runs, [0.525] * runs, [1 / 0.4] * runs])
    st0 = FL.lsw_state(len(faces), nsf, runs, sn0=-120.0, ss0_0=ss0, ss0_1=0.0, mus=glob[0])
    frame = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    dt = S.dt_cfl(m, rho, lam, mu, order)
    return m, rho, lam, mu, frame, glob, st0, dt


def fault_row(args, timed):
    """Elastic + dynamic rupture: 331,776 tets, 4,608 fault faces, degree 3,
    16 fused runs in FP32 (friction, net updates and the fault's sub-face
    projections in FP64), one graph-captured coupled step (local, fault,
    neighbour) per time step; beside it the fault phase alone."""
    import torch
    from edge_amd import fault as FL
    from edge_amd.solver import ElasticSolver
    order, runs = 4, 16
    m, rho, lam, mu, frame, glob, st0, dt = fault_problem(order, runs)
    s = ElasticSolver(m, order, runs, 32, rho, lam, mu)
    f = FL.Fault(s, m, rho, lam, mu, frame, glob, st0)
    d = torch.zeros(s.dof_shape(), dtype=torch.float32, device="cuda")
    t = torch.zeros_like(d)
    n_steps = 100
    # one multi-step call: fused neighbour + local launches with the fault
    # phase between them, captured as one graph
    ms = timed(lambda st: s.step(d, t, dt, n_steps, graph=True, stream=st), 1) / n_steps
    st, slip = f.get()
    fault_ms = timed(lambda st: f.apply(d, t, dt, stream=st), 5)
    fl = FLOPS[order]
    n = s.n_modes
    step_flop = (fl["local"] + fl["neigh"]) * runs * m.n_el
    print(json.dumps({"row": "config 4: elastic + dynamic rupture (coupled step)", "elements": m.n_el,
                      "fault_faces": f.n_faces, "sub_faces": f.n_sf, "order": order, "runs": runs, "precision": 32,
                      "graph": True, "fused": bool(s.kernel_info()[1] & 16), "steps": 2 * n_steps,
                      "t_end": 2 * n_steps * dt, "ms_per_step": ms * 1e3,
                      "dof_updates_per_s": m.n_el * 9 * n * runs / ms, "fault_phase_ms": fault_ms * 1e3,
                      "step_tflops": step_flop / ms / 1e12, "step_frac": step_flop / ms / PEAK["fp32_mfma"],
                      "slipping_sub_faces": float(slip.mean()), "max_slip": float(st[..., 11].max()),
                      "finite": bool(torch.isfinite(d).all())}), flush=True)
    f.close()
    s.close()


def rows(args):
    """Per-row measurements beside the headline (one JSON line per row):
    viscoelastic updates, local time stepping, receivers and sources, and the
    rupture net updates, each with HIP events around its launches."""
    import torch
    from edge_amd import mesh as M
    from edge_amd import setups as S
    from edge_amd.lts import Rate2Lts
    from edge_amd.rupture import Rupture
    from edge_amd.solver import ElasticSolver, ViscoSolver

    def timed(fn, reps):
        st = torch.cuda.current_stream()
        fn(st)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn(st)
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e-3 / reps

    want = set(args.only.split(",")) if args.only else {"visco", "ensemble", "lts", "points", "rupture", "fault"}
    if args.no_ensemble:
        want.discard("ensemble")
    if "fault" in want:
        fault_row(args, timed)
    if not want - {"fault"}:
        return
    m, rho, lam, mu, dt, q0 = build_problem(args.hex, args.order)
    n = q0.shape[2]
    # ---- viscoelastic (3 mechanisms, f_c = 1 Hz, ratio 10): FP32 8 runs and FP64 one run
    rng = np.random.default_rng(5)
    qp = rng.uniform(50, 200, m.n_el)
    fl, fn_ = visco_flops(args.order, 3)
    for runs, prec in (((8, 32), (1, 64)) if "visco" in want else ()):
        s = ViscoSolver(m, args.order, runs, prec, rho, lam, mu, qp, qp / 2, n_rms=3, freq_cen=1.0, freq_rat=10.0)
        tune = [] if args.no_autotune else s.autotune(dt, 2)
        d = torch.from_numpy(q0)[..., None].repeat(1, 1, 1, runs).to(s.dtype).cuda().contiguous()
        t, a = torch.zeros_like(d), s.alloc_a()
        loc = timed(lambda st: s.local(d, t, dt, dofs_a=a, stream=st), 2)
        nei = timed(lambda st: s.neigh(d, t, dofs_a=a, stream=st), 2)
        fused = bool(s.kernel_info()[1] & 16)
        # whole steps: one multi-step call (fused launches where the solver has them)
        step = timed(lambda st: s.step(d, t, dt, 6, a, stream=st), 1) / 6
        peak = PEAK["fp32_mfma"] if prec == 32 else PEAK["fp64"]
        kern = {1: "generic (LDS tiles, VALU)", 2: "MFMA (paired products, elementwise anelastic terms)",
                4: "FP64 MFMA (anelastic rows in the quantity tiles)"}
        print(json.dumps({"row": "a12 viscoelastic update", "kernel": kern.get(s.kernel_info()[0], "?"),
                          "hi_occ": bool(s.kernel_info()[1] & 32),
                          "autotune": [{"kernel": k, "projections": p, "xcd_schedule": x, "fused": f, "occupancy": o,
                                        "ms_per_step": c} for k, p, x, c, f, o in tune],
                          "order": args.order,
                          "runs": runs, "precision": prec, "n_rms": 3, "elements": m.n_el,
                          "local_ms": loc * 1e3, "neigh_ms": nei * 1e3, "fused": fused,
                          "step_ms": step * 1e3, "dof_updates_per_s": m.n_el * 27 * n * runs / step,
                          "tflops": (fl + fn_) * runs * m.n_el / step / 1e12,
                          "roofline_frac": (fl + fn_) * runs * m.n_el / step / peak}), flush=True)
        s.close()
        del d, t, a
    # ---- batched viscoelastic ensemble: 24 meshes of seeded sizes (6 n^3 tets,
    # n in [8, 32]) as one disjoint union, one graph-captured step per time step
    if "ensemble" in want:
        ensemble_row(args, timed, fl + fn_)
    if not want & {"lts", "points", "rupture"}:
        return
    # ---- local time stepping: two groups (halves of the mesh), FP32 16 runs
    s = ElasticSolver(m, args.order, 16, 32, rho, lam, mu)
    d = (torch.from_numpy(q0)[..., None].repeat(1, 1, 1, 16)).to(torch.float32).cuda().contiguous()
    t = torch.zeros_like(d)
    half = m.n_el // 2
    gts = timed(lambda st: s.step(d, t, dt / 2, 2, stream=st), 2)
    lts = Rate2Lts(s, [(0, half), (half, m.n_el - half)], dt / 2)
    cyc = timed(lambda st: lts.cycle(d, 1, stream=st), 2)
    print(json.dumps({"row": "a7 local time stepping (rate 2)", "order": args.order, "runs": 16, "precision": 32,
                      "elements": m.n_el, "interface_elements": int((lts.el_lts != 0).sum()),
                      "gts_two_fine_steps_ms": gts * 1e3, "lts_cycle_ms": cyc * 1e3,
                      "element_updates_per_cycle": int(half * 2 + (m.n_el - half)),
                      "ms_per_element_update_ratio_lts_over_gts": (cyc / (half * 2 + m.n_el - half)) /
                                                                   (gts / (2 * m.n_el))}), flush=True)
    # ---- receivers (64) and sources (16) per step
    s.set_lts(None, None)
    src, recv = point_setup(s, m, 16)
    out = torch.empty((recv.n, 1, 9, 16), dtype=torch.float32, device="cuda")
    tau = np.zeros((recv.n, 1))
    r_ms = timed(lambda st: recv.eval_into(d, tau, out, st), 5)
    s_ms = timed(lambda st: src.apply(0.0, dt, d, st), 5)
    print(json.dumps({"row": "a3 / (f) receivers and point sources", "receivers": recv.n, "sources": src.n,
                      "receiver_eval_ms": r_ms * 1e3, "source_apply_ms": s_ms * 1e3}), flush=True)
    src.close()
    recv.close()
    s.close()
    del d, t
    # ---- rupture net updates: 4,608 fault faces, 49 sub-faces (degree 3), 16 runs, FP64
    nf, nsf, R = 4608, 49, 16
    rng = np.random.default_rng(7)
    nrm = np.tile([0.0, 0.0, 1.0], (nf, 1))
    tan = np.tile([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], (nf, 1, 1))
    mat = np.tile([2.67, 32.0, 32.0], (nf, 1))
    st0 = np.zeros((nf, nsf, R, 14))
    st0[..., 0] = -120.0
    st0[..., 2] = rng.uniform(70, 81.24, (nf, nsf, R))
    st0[..., 5] = 0.677
    glob = np.stack([np.full(R, 0.677), np.full(R, 0.525), np.full(R, 1 / 0.4)])
    rup = Rupture(nsf, 343, R, nrm, tan, np.full(nf, 0.01), np.eye(3)[[2, 0, 1]], mat, mat, np.full(nf, 1e-3),
                  np.full(nf, 1e-3), glob, st0)
    dl = torch.tensor(rng.normal(size=rup.dof_shape()), device="cuda")
    dr = torch.tensor(rng.normal(size=rup.dof_shape()), device="cuda")
    nl, nr_ = rup.alloc(), rup.alloc()
    r_s = timed(lambda st: rup.net_updates(dl, dr, nl, nr_, 1e-3, stream=st), 5)
    items = nf * nsf * R
    byts = items * (4 * 9 * 8 + 2 * 13 * 8 + 1) + nf * 324 * 8
    print(json.dumps({"row": "a13/a14 rupture net updates", "faces": nf, "sub_faces": nsf, "runs": R,
                      "precision": 64, "ms": r_s * 1e3, "bytes_per_launch": byts, "gb_per_s": byts / r_s / 1e9,
                      "roofline_frac": byts / r_s / PEAK["hbm"]}), flush=True)
    rup.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--hex", type=int, default=32, help="hexes per axis (32 -> 196,608 tets)")
    ap.add_argument("--order", type=int, default=5)
    ap.add_argument("--kernel", type=int, default=0, help="0 auto, 1 generic, 2 MFMA-tiled")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 plain comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-ensemble", action="store_true", help="--rows without the 24-mesh ensemble")
    ap.add_argument("--only", default="", help="--rows: comma list of visco, ensemble, lts, points, rupture, fault")
    ap.add_argument("--no-autotune", action="store_true", help="measure the kernels --kernel selects, untuned")
    ap.add_argument("--fp32-variant", default="auto", choices=["auto"] + sorted(FP32_VARIANTS),
                    help="fix the FP32 kernel variant (no autotuning): fused, proj or tiles")
    ap.add_argument("--rows", action="store_true",
                    help="measure the other rows (viscoelastic, LTS, receivers/sources, rupture) instead")
    ap.add_argument("--no-points", action="store_true",
                    help="time the element update alone (no point sources and receivers in the step)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the element update's multi-GPU sharding is a design, not a measured path.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.rows:
        rows(args)
        return
    import torch
    from edge_amd.solver import KERNEL_EC64, KERNEL_MFMA, KERNEL_MFMA64, PATH_FUSED, PATH_PROJECTIONS, ElasticSolver

    order, runs = args.order, 16
    print(f"[bench] building the {6 * args.hex ** 3}-tet mesh (order {order})", file=sys.stderr, flush=True)
    m, rho, lam, mu, dt, q0 = build_problem(args.hex, order)
    n = q0.shape[2]
    scale = torch.linspace(0.25, 1.0, runs, dtype=torch.float64)  # distinct runs
    res = {}

    # ---- FP32 fused-16 (headline)
    s32 = ElasticSolver(m, order, runs, 32, rho, lam, mu, kernel=args.kernel)
    tune = {}
    if args.fp32_variant != "auto":  # a fixed variant (profiling): no autotuning
        s32.set_variant(KERNEL_MFMA, *FP32_VARIANTS[args.fp32_variant])
    elif not args.no_autotune:  # untimed set-up: the fastest kernel variant on this mesh
        tune["fp32"] = [{"kernel": k, "projections": p, "xcd_schedule": x, "fused": f, "ms_per_step": ms}
                        for k, p, x, ms, f, _ in s32.autotune(dt, 3)]
    flags = s32.kernel_info()[1]
    proj, fused = bool(flags & PATH_PROJECTIONS), bool(flags & PATH_FUSED)
    variant = "fused" if fused else ("proj" if proj else "tiles")
    d32 = (torch.from_numpy(q0)[..., None] * scale).to(torch.float32).cuda().contiguous()
    t32 = torch.zeros_like(d32)
    points = None if args.no_points else point_setup(s32, m, runs)
    print(f"[bench] fp32 fused-16 timing ({variant})", file=sys.stderr, flush=True)
    wall, ks = time_solver(s32, d32, t32, dt, args.steps, args.warmup, points, fused=fused)
    if points is not None:
        for p in points:
            p.close()
    ok = bool(torch.isfinite(d32).all())
    res["fp32"] = dict(wall=wall, k=ks, finite=ok)
    del d32, t32
    s32.close()

    # ---- FP64 plain single run (the drop-in)
    if not args.no_fp64:
        s64 = ElasticSolver(m, order, 1, 64, rho, lam, mu, kernel=args.kernel)
        if not args.no_autotune:
            tune["fp64"] = [{"kernel": k, "projections": p, "xcd_schedule": x, "ms_per_step": ms}
                            for k, p, x, ms, _, _ in s64.autotune(dt, 2)]
        d64 = torch.from_numpy(q0[..., None].copy()).cuda().contiguous()
        t64 = torch.zeros_like(d64)
        print("[bench] fp64 plain timing", file=sys.stderr, flush=True)
        w64, k64 = time_solver(s64, d64, t64, dt, max(1, args.steps // 2), args.warmup)
        res["fp64"] = dict(wall=w64, k=k64, steps=max(1, args.steps // 2))
        res["fp64"]["kernel"] = s64.kernel_info()[0]
        if res["fp64"]["kernel"] == KERNEL_EC64:
            # beside it the per-element FP64 MFMA kernels (one element per wave)
            s64.set_variant(KERNEL_MFMA64, False, True)
            d64.copy_(torch.from_numpy(q0[..., None].copy()))
            t64.zero_()
            print("[bench] fp64 per-element timing", file=sys.stderr, flush=True)
            wpe, kpe = time_solver(s64, d64, t64, dt, max(1, args.steps // 2), args.warmup)
            res["fp64_pe"] = dict(wall=wpe, k=kpe, steps=max(1, args.steps // 2))
        del d64, t64
        s64.close()
        # ---- FP32 single run (config 2's FP32 row), beside the FP64 one
        s1 = ElasticSolver(m, order, 1, 32, rho, lam, mu)
        if not args.no_autotune:
            tune["fp32_single_run"] = [{"kernel": k, "projections": p, "xcd_schedule": x, "ms_per_step": ms}
                                       for k, p, x, ms, _, _ in s1.autotune(dt, 2)]
        d1 = torch.from_numpy(q0[..., None].copy()).to(torch.float32).cuda().contiguous()
        t1 = torch.zeros_like(d1)
        print("[bench] fp32 single-run timing", file=sys.stderr, flush=True)
        w1, k1 = time_solver(s1, d1, t1, dt, max(1, args.steps // 2), args.warmup)
        res["fp32_r1"] = dict(wall=w1, k=k1, steps=max(1, args.steps // 2), kernel=s1.kernel_info()[0],
                              finite=bool(torch.isfinite(d1).all()))
        del d1, t1
        s1.close()

    dofs_per_step = m.n_el * 9 * n
    value = dofs_per_step * runs * args.steps / res["fp32"]["wall"]
    fl = FLOPS.get(order)
    roof = None
    if fl:
        k32 = res["fp32"]["k"]
        f_step = fl["local"] + fl["neigh"]
        if fused and "fused" in k32:
            # the fused launch: neighbour step k - 1 (projection lifts) and
            # local step k with the projections it writes: a whole step's FLOPs
            dom, f_dom, t_dom = "fused", f_step, k32["fused"]
            t_step = (k32["local"] + (args.steps - 1) * k32["fused"] + k32["neigh"]) / args.steps
            desc = ("fused neighbour step k - 1 + local step k (projection lifts; CK + volume + local surface + "
                    "neighbour-flux projections)")
        else:
            f_loc = fl["local"] + (fl["proj"] if proj else 0)
            dom, f_dom, t_dom = "local", f_loc, k32["local"]
            t_step = k32["local"] + k32["neigh"]
            desc = "local (CK + volume + local surface" + (" + neighbour-flux projections)" if proj else ")")
        achieved = f_dom * runs * m.n_el / t_dom
        step_ach = f_step * runs * m.n_el / t_step
        traffic = read_traffic(TRAFFIC_PROFILE[variant], m.n_el, order, runs)
        roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["fp32_mfma"] / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK["fp32_mfma"], "traffic": traffic,
                "traffic_alg": launch_bytes(order, runs, variant) * m.n_el,
                "kernel": desc, "kernel_ms": t_dom * 1e3, "flops_per_element_run": f_dom,
                "kernel_ms_all": {k: v * 1e3 for k, v in k32.items()},
                "step_tflops": step_ach / 1e12, "step_frac": step_ach / PEAK["fp32_mfma"]}
        util = read_mfma_util(MFMA_PROFILE, m.n_el, order, runs, dom_kernel_name(variant, order))
        if util is not None:
            roof["mfma_busy"] = util
    out = {
        "metric": METRIC, "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * res["fp32"]["wall"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"elastic tet4 O{order} FP32 fused-{runs}, {m.n_el} tets (32^3 hexes x 6, "
                               f"+-10% jitter, random rho/vp), plane-wave initial DOFs scaled per run"
                               + ("" if args.no_points else f", one Ricker moment-tensor point source per run, "
                                  f"64 receivers sampled every step"),
                   "elements": m.n_el, "order": order, "runs": runs, "dt": dt, "variant": variant,
                   "finite": res["fp32"]["finite"]},
        "roofline": roof,
    }
    if tune:
        out["autotune"] = tune
    if "fp64" in res:
        k64 = res["fp64"]["k"]
        v64 = dofs_per_step * res["fp64"]["steps"] / res["fp64"]["wall"]
        out["fp64_plain"] = {"value": v64, "unit": "DOF-updates/s", "ms_per_step": 1e3 * res["fp64"]["wall"] /
                             res["fp64"]["steps"], "local_ms": 1e3 * k64["local"], "neigh_ms": 1e3 * k64["neigh"]}
        if fl:
            a64 = (fl["local"] + fl["neigh"]) * m.n_el / (k64["local"] + k64["neigh"])
            out["fp64_plain"]["roofline_frac"] = a64 / PEAK["fp64"]
            name64 = (f"k_local_ec<double, {order}>" if res["fp64"]["kernel"] == KERNEL_EC64
                      else f"k_local_p64<{order}, true>")
            u64 = read_mfma_util(MFMA_PROFILE, m.n_el, order, 1, name64)
            out["fp64_plain"]["kernel"] = name64
            if u64 is not None:
                out["fp64_plain"]["mfma_busy_local"] = u64
        out["speedup_fp32_fused16_vs_fp64_plain"] = value / v64
    if "fp64_pe" in res:
        kpe = res["fp64_pe"]["k"]
        vpe = dofs_per_step * res["fp64_pe"]["steps"] / res["fp64_pe"]["wall"]
        out["fp64_per_element"] = {"value": vpe, "unit": "DOF-updates/s", "ms_per_step": 1e3 * res["fp64_pe"]["wall"]
                                   / res["fp64_pe"]["steps"], "local_ms": 1e3 * kpe["local"],
                                   "neigh_ms": 1e3 * kpe["neigh"], "kernel": f"k_local_p64<{order}, true>"}
        if fl:
            out["fp64_per_element"]["roofline_frac"] = ((fl["local"] + fl["neigh"]) * m.n_el /
                                                        (kpe["local"] + kpe["neigh"]) / PEAK["fp64"])
        out["speedup_fp32_fused16_vs_fp64_per_element"] = value / vpe
    if "fp32_r1" in res:
        k1 = res["fp32_r1"]["k"]
        out["fp32_single_run"] = {"value": dofs_per_step * res["fp32_r1"]["steps"] / res["fp32_r1"]["wall"],
                                  "unit": "DOF-updates/s", "ms_per_step": 1e3 * res["fp32_r1"]["wall"] /
                                  res["fp32_r1"]["steps"], "local_ms": 1e3 * k1["local"], "neigh_ms": 1e3 * k1["neigh"],
                                  "kernel": res["fp32_r1"]["kernel"], "finite": res["fp32_r1"]["finite"]}
        if fl:
            a1 = (fl["local"] + fl["neigh"]) * m.n_el / (k1["local"] + k1["neigh"])
            out["fp32_single_run"]["roofline_frac"] = a1 / PEAK["fp32_mfma"]
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(m, order, runs, rho, lam, mu, dt, q0, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
