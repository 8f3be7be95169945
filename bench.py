"""Benchmark of the sphere Helmholtz solve path on one MI355X.

Workload (one "step" = one solve): the level-7 icosahedral mesh (163,842
vertices, 163,841 inner unknowns, 1,146,869 CSR nonzeros), system
(M - tau L) u = M f with tau = 1e-2 assembled on the GPU, f a seeded real
spherical-harmonic field of degree <= 32, x0 = 0, GMRES(30) to a true FP64
relative residual of 1e-10.  `value` is plain-FP64 MGS solves/s (the drop-in
form, each restart cycle one cooperative kernel); the mixed-precision and
other forms are reported beside it with their error against the FP64 oracle
solution.  `roofline` describes that cycle kernel (GMRES(30) byte model over
its HIP-event duration), the whole solve, and the SpMV in cache (level 7)
and streamed from HBM (level 8).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  Inputs are resident in HBM before the timed region.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
LEVEL = 7
TAU = 1e-2


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def sph_field(xyz, lmax, seed):
    from tests._common import real_sph_harm_field
    return real_sph_harm_field(xyz, lmax, seed)


def spmv_bytes(n, nnz, sv, sx):
    """Algorithmic bytes of one CSR SpMV with int32 indices."""
    return nnz * (sv + 4) + 4 * (n + 1) + 2 * n * sx


def gmres_solve_bytes(n, nnz, iterations, restart=30, sv=8, sx=8):
    """Algorithmic bytes of a GMRES(restart) solve of `iterations` Arnoldi
    steps: step j reads the matrix and its source vector and writes w (one
    SpMV), then touches w twice and each of v_0..v_j once (j + 5 vectors of
    n); each cycle adds the residual (an SpMV and 2 vectors), v_0's
    normalisation (2) and the x update (k + 2 vectors for k steps).  A full
    30-step cycle is 31 SpMV + 621 vectors."""
    B = spmv_bytes(n, nnz, sv, sx)
    total, left = 0, iterations
    while left > 0:
        k = min(restart, left)
        left -= k
        total += k * B + sum(j + 5 for j in range(k)) * n * sx
        total += B + 2 * n * sx + 2 * n * sx + (k + 2) * n * sx
    return total


def time_solves(A, b, x, steps, warmup, stream, **kw):
    import torch
    infos = []
    for _ in range(warmup):
        A.solve(x, b, stream=stream, **kw)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        infos.append(A.solve(x, b, stream=stream, **kw))
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = sorted(s.elapsed_time(e) for s, e in ev)
    pct = lambda q: lat[min(len(lat) - 1, int(math.ceil(q * len(lat))) - 1)]
    return wall, lat, pct(0.5), pct(0.99), infos


def kernel_avg_ms(fn, launches, stream):
    """Average duration of one launch of fn: `launches` back-to-back launches
    on `stream` between two HIP events on that stream.  A spin kernel queued
    first holds the stream while the host enqueues every launch, so the events
    time the device running them back to back rather than the host's issue
    rate (a launch through ctypes costs the host a few microseconds, about
    the SpMV's own duration)."""
    import torch
    for _ in range(10):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        torch.cuda._sleep(20_000_000)
    s.record(stream)
    for _ in range(launches):
        fn()
    e.record(stream)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / launches


def traffic_from_profiles(kernel_file):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the factor a
    plain copy of known size gives), or None if it was not recorded."""
    path = os.path.join(ROOT, "profiles", kernel_file)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("traffic_bytes_per_launch")


def serve_leg(n_requests, rate, seed=20251202, verbose=False, load_fracs=(0.5, 0.8, 0.95),
              load_requests=128):
    """Ensemble serving: 256 members on levels 5/6/7/8 (40/30/20/10 %), each
    request one BarVortex implicit step (dt = 600 s, Earth-scaled units) of
    its member; Poisson arrivals at `rate` req/s.  The initial state of a
    member is the Rossby-Haurwitz wave-4 vorticity plus N(0, 1e-3 max|w|) per
    vertex, inverted to a stream function by the multigrid Poisson solve.
    Runs the FP64 and the mixed form; returns one dict per form."""
    import numpy as np
    import torch

    import mke_amd as P

    levels, weights, n_members = (5, 6, 7, 8), (0.4, 0.3, 0.2, 0.1), 256
    # at L8 the reference fixed point amplifies a mode at tau = 600 s (the
    # change grows x1.7 per two passes from ~1e-6): each request there runs
    # two BarVortex::calc steps of 300 s
    substeps = (1, 1, 1, 2)
    counts = [int(round(w * n_members)) for w in weights]
    counts[0] += n_members - sum(counts)
    omega = 7.292e-5
    wrh = 7.848e-6 / omega
    rng = np.random.default_rng(seed)
    init = {}
    for lv, cnt in zip(levels, counts):
        mesh = P.Mesh.sphere(lv)
        a = mesh.arrays()
        pr, inner = a["pr"], a["inner"]
        phi = np.arcsin(np.clip(pr[:, 2], -1, 1))
        lam = np.arctan2(pr[:, 1], pr[:, 0])
        w0 = 2 * wrh * np.sin(phi) - 30 * wrh * np.sin(phi) * np.cos(phi) ** 4 * np.cos(4 * lam)
        W = w0[None, :] + rng.standard_normal((cnt, len(w0))) * 1e-3 * np.abs(w0).max()
        mg = P.Multigrid(mesh, P.LAPLACE)
        M = P.Matrix.assemble(mesh, P.MASS)
        psis = []
        for c0 in range(0, cnt, 64):
            k = min(64, cnt - c0)
            Bm = torch.zeros((len(inner), k), dtype=torch.float64, device="cuda")
            for q in range(k):
                col = torch.zeros(len(inner), dtype=torch.float64, device="cuda")
                M.mult_vector(col, torch.as_tensor(W[c0 + q][inner], device="cuda"))
                Bm[:, q] = col
            Xm = torch.zeros_like(Bm)
            mg.solve_batch(Xm, Bm, rtol=1e-10, max_iters=500)
            for q in range(k):
                psi = np.zeros(len(w0))
                psi[inner] = Xm[:, q].cpu().numpy()
                psis.append(psi)
        init[lv] = psis
        log("serve: %d initial states at L%d" % (cnt, lv))
    members_level = [lv for lv, cnt in zip(levels, counts) for _ in range(cnt)]
    req_members = rng.integers(0, n_members, n_requests)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    out = {}
    for name, prec in (("serve_ensemble_f64", P.F64), ("serve_ensemble_mixed_f32", P.F32)):
        t_srv = time.perf_counter()
        srv = P.Server(levels=levels, precision=prec, verbose=verbose, substeps=substeps)
        log("serve: %s engines ready in %.1fs, slots %s" %
            (name, time.perf_counter() - t_srv, [srv.slots(lv) for lv in levels]))
        ids, k_at = [], {lv: 0 for lv in levels}
        for lv in members_level:
            ids.append(srv.add_member(lv, init[lv][k_at[lv]]))
            k_at[lv] += 1
        # warm-up: one request per slot of every level (graphs, workspaces)
        warm = [ids[members_level.index(lv) + s] for lv in levels for s in range(srv.slots(lv))]
        srv.run(warm, np.zeros(len(warm)))
        log("serve: warm-up done (%d steps)" % len(warm))
        start, finish, infos, st = srv.run(req_members, arrivals)
        lat = finish - arrivals
        its = np.array([i["iterations"] for i in infos])
        conv = np.array([i["converged"] for i in infos])
        div = np.array([i["diverged"] for i in infos])
        by_level = {}
        for lv in levels:
            sel = np.array([members_level[m] == lv for m in req_members])
            if sel.any():
                per_pass = lambda key: round(float(np.sum([infos[r][key] for r in np.flatnonzero(sel)])  # noqa: E731
                                                   / max(1, its[sel].sum())), 2)
                by_level["L%d" % lv] = dict(requests=int(sel.sum()),
                                            p50_ms=round(float(np.median(lat[sel])) * 1e3, 2),
                                            fp_passes_mean=round(float(its[sel].mean()), 2),
                                            gmres_steps_per_pass=dict(
                                                jacobian=per_pass("it_jacobian"),
                                                helmholtz=per_pass("it_helmholtz"),
                                                poisson=per_pass("it_poisson"),
                                                mass=per_pass("it_mass")),
                                            converged=int(conv[sel].sum()),
                                            diverged=int(div[sel].sum()),
                                            slots=srv.slots(lv))
        out[name] = dict(steps_per_s=round(st["throughput"], 3),
                         p50_ms=round(st["p50_latency_s"] * 1e3, 2),
                         p99_ms=round(st["p99_latency_s"] * 1e3, 2),
                         mean_ms=round(st["mean_latency_s"] * 1e3, 2),
                         p50_service_ms=round(st["p50_service_s"] * 1e3, 2),
                         mean_members_per_pass=round(st["mean_occupancy"], 3),
                         requests=int(n_requests), offered_req_per_s=rate,
                         members=n_members, converged=int(conv.sum()), diverged=int(div.sum()),
                         device_GB=round(st["device_bytes"] / 1e9, 2),
                         substeps=dict(zip(["L%d" % lv for lv in levels], substeps)),
                         levels=by_level)
        log("%s: %.2f steps/s, p50 %.0f ms, p99 %.0f ms, %.2f members/pass" %
            (name, st["throughput"], st["p50_latency_s"] * 1e3, st["p99_latency_s"] * 1e3,
             st["mean_occupancy"]))
        # latency below saturation: Poisson arrivals at fractions of the
        # throughput the burst measured (the same members and levels)
        loads = []
        for frac in load_fracs:
            r_l = frac * st["throughput"]
            mem_l = rng.integers(0, n_members, load_requests)
            arr_l = np.cumsum(rng.exponential(1.0 / r_l, load_requests))
            _, fin_l, inf_l, st_l = srv.run(mem_l, arr_l)
            loads.append(dict(offered_frac=frac, offered_req_per_s=round(r_l, 3),
                              requests=load_requests,
                              steps_per_s=round(st_l["throughput"], 3),
                              p50_ms=round(st_l["p50_latency_s"] * 1e3, 2),
                              p99_ms=round(st_l["p99_latency_s"] * 1e3, 2),
                              p50_service_ms=round(st_l["p50_service_s"] * 1e3, 2),
                              mean_members_per_pass=round(st_l["mean_occupancy"], 3),
                              converged=int(sum(i["converged"] for i in inf_l))))
            log("%s at %.0f%% of capacity (%.1f req/s): p50 %.0f ms, p99 %.0f ms" %
                (name, frac * 100, r_l, loads[-1]["p50_ms"], loads[-1]["p99_ms"]))
        out[name]["offered_loads"] = loads
        srv.close()
    return out


def serve_leg_child(n_requests, rate):
    """serve_leg in a child process of its own, as a server runs.  Each
    engine has one HIP stream, and HIP maps a process's streams onto a few
    hardware queues (GPU_MAX_HW_QUEUES, 4 by default): after this process's
    other legs, whose solver workspaces keep streams of their own, two
    engines can land on one queue and run one after the other (measured:
    24.6 instead of 29.5 steps/s for the FP64 form, 29.5 with 8 queues)."""
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--serve-only",
           "--serve-requests", str(n_requests), "--serve-rate", str(rate)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, cwd=ROOT)
    if r.returncode != 0:
        raise RuntimeError("serving leg failed (exit %d)" % r.returncode)
    return json.loads(r.stdout.strip().splitlines()[-1])["serve_only"]


def l8_leg(steps, oracle_check, seed=20251203):
    """Mixed precision on the level-8 mesh (655,362 vertices, 4,587,509
    nonzeros): (M - 1e-2 L) u = M f, f of degree <= 64 with N(0,1) max(l,1)^-2
    coefficients; FP64 forms and FP32/BF16 inner solves under FP64
    refinement, each to a true FP64 relative residual of 1e-10, with the
    forward error against a tight FP64 oracle solve of the same CSR system
    (rtol 1e-12, about the FP64 floor at this level; host cores) when
    oracle_check is set."""
    import numpy as np
    import torch

    import mke_amd as P
    from tests._common import real_sph_harm_field_lean

    stream = torch.cuda.current_stream()
    t = time.perf_counter()
    mesh = P.Mesh.sphere(8)
    cm, cl = P.helmholtz_coeffs(tau=1.0, sigma=0.0, theta=1.0, mu=TAU)
    A = P.Matrix.assemble(mesh, P.HELMHOLTZ, cm, cl, stream=stream)
    Mm = P.Matrix.assemble(mesh, P.MASS, stream=stream)
    a = mesh.arrays()
    f = real_sph_harm_field_lean(a["pr"], 64, seed, decay=2.0, device="cuda")[a["inner"]]
    b = torch.zeros(A.n_rows, dtype=torch.float64, device="cuda")
    Mm.mult_vector(b, torch.as_tensor(f, device="cuda"), stream=stream)
    torch.cuda.synchronize()
    log("L8: mesh, matrices and rhs in %.1fs (n=%d nnz=%d)" %
        (time.perf_counter() - t, A.n_rows, A.nnz))
    out, xs = {}, {}
    for name, kw in (("L8_fp64_mgs", dict(ortho=P.MGS)),
                     ("L8_fp64_cgs2", dict(ortho=P.CGS2)),
                     ("L8_mixed_f32_cgs2", dict(ortho=P.CGS2, precision=P.F32)),
                     ("L8_mixed_bf16_cgs2", dict(ortho=P.CGS2, precision=P.BF16))):
        x = torch.zeros_like(b)
        w, lat, q50, q99, inf = time_solves(A, b, x, steps, 1, stream, restart=30,
                                            rtol=1e-10, **kw)
        out[name] = dict(solves_per_s=round(len(inf) / w, 3), p50_ms=round(q50, 3),
                         p99_ms=round(q99, 3), iterations=inf[-1]["iterations"],
                         rel_residual=inf[-1]["rel_residual"], converged=inf[-1]["converged"],
                         level=8, rhs="real Y_lm, l<=64, c ~ N(0,1) max(l,1)^-2")
        xs[name] = x.cpu().numpy()
        log("%s: %.2f solves/s, p50 %.1f ms, %d iters, rel %.2e" %
            (name, out[name]["solves_per_s"], q50, inf[-1]["iterations"],
             inf[-1]["rel_residual"]))
    if oracle_check:
        from oracle import oracle as O
        t = time.perf_counter()
        rp, ci, v = A.csr()
        xt, info = O.gmres(rp, ci, v, b.cpu().numpy(), restart=30, rtol=1e-12,
                           max_iters=20000)
        nt = np.linalg.norm(xt)
        for name in out:
            out[name]["fwd_err_vs_oracle"] = float(np.linalg.norm(xs[name] - xt) / nt)
        log("L8 oracle solve: %d iterations in %.1fs" % (info["iterations"], time.perf_counter() - t))
    out["L8_spmv_fp64"] = l8_spmv_roofline(A, stream)
    return out


def calib_copy():
    """Five device copies of 256 MiB: what rocprof's byte counters read for
    them scales the counters of the other kernels."""
    import torch
    src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("calibration copy: 5 x 256 MiB")


def l8_spmv_roofline(A, stream, copies=8):
    """FP64 SpMV of the level-8 matrix: the same matrix repeated (its 65 MB
    stay in the 256 MB Infinity Cache) and round-robin over `copies` copies
    (about 0.5 GB, which must stream from HBM)."""
    import torch

    import mke_amd as P
    n, nnz = A.n_rows, A.nnz
    nbytes = nnz * 12 + 4 * (n + 1) + 2 * n * 8
    rp, ci, v = A.csr()
    mats = [A] + [P.Matrix.from_csr(n, n, rp, ci, v) for _ in range(copies - 1)]
    xs = [torch.randn(n, dtype=torch.float64, device="cuda") for _ in range(copies)]
    ys = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(copies)]
    for m, xx, yy in zip(mats, xs, ys):   # builds every copy's ELL before timing
        m.mult_vector(yy, xx, stream=stream)
    torch.cuda.synchronize()
    warm = sorted(kernel_avg_ms(lambda: A.mult_vector(ys[0], xs[0], stream=stream), 200, stream)
                  for _ in range(3))[1]
    it = [0]

    def cold():
        k = it[0] % copies
        it[0] += 1
        mats[k].mult_vector(ys[k], xs[k], stream=stream)

    cold_ms = sorted(kernel_avg_ms(cold, 200, stream) for _ in range(3))[1]
    for m in mats[1:]:
        m.close()
    res = dict(bytes_per_launch=nbytes, copies=copies,
               warm_us=round(warm * 1e3, 3), warm_gbs=round(nbytes / (warm * 1e-3) / 1e9, 1),
               cold_us=round(cold_ms * 1e3, 3),
               cold_gbs=round(nbytes / (cold_ms * 1e-3) / 1e9, 1))
    res["warm_frac"] = round(res["warm_gbs"] / HBM_PEAK_GBS, 4)
    res["cold_frac"] = round(res["cold_gbs"] / HBM_PEAK_GBS, 4)
    log("L8 spmv fp64: %.2f us repeated (%.0f GB/s), %.2f us over %d copies (%.0f GB/s)" %
        (res["warm_us"], res["warm_gbs"], res["cold_us"], copies, res["cold_gbs"]))
    return res


def barvortex_leg(level, members=8, steps=2):
    """One standalone BarVortex engine (no serving threads): `members` RH
    wave-4 members with seeded noise stepped together `steps` times (tau =
    300 s, the serving workload's level-8 sub-step).  For profiling a pass."""
    import numpy as np
    import torch

    import mke_amd as P
    from tests._common import rh_state

    mesh = P.Mesh.sphere(level)
    a = mesh.arrays()
    pr = a["pr"]
    lh = torch.as_tensor(2.0 * np.clip(pr[:, 2], -1, 1), device="cuda")
    U = torch.as_tensor(np.stack([rh_state(pr, seed=100 + k) for k in range(members)]),
                        device="cuda")
    bv = P.BarVortex(mesh, lh, tau=300.0 * 7.292e-5, sigma=1.6e-2, mu=8e-5, k1=1.0, k2=1.0,
                     theta=0.5, slots=members)
    U1 = torch.zeros_like(U)
    bv.calc(U1, U)   # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        infos = bv.calc(U1, U)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    passes = [i["iterations"] for i in infos]
    log("barvortex L%d: %.1f ms per step of %d members, passes %s, gmres per pass %s" %
        (level, dt * 1e3, members, passes,
         {k: round(float(np.mean([i[k] / max(1, i["iterations"]) for i in infos])), 1)
          for k in ("it_jacobian", "it_helmholtz", "it_poisson")}))
    return dict(ms_per_step=round(dt * 1e3, 2), level=level, members=members, passes=passes)


def baroclin_leg(level=5, members=8, steps=3):
    """Baroclin::calc on the device for `members` members at once: the
    reference driver's parameters (tau 1e-3, sigma 1.6e-2, mu 8e-2, sigma1 =
    sigma, mu1 = mu, alpha 1, no Coriolis) and its manufactured pair with its
    forcing, each member perturbed by a different smooth bump; block-
    preconditioned 4n x 4n solve.  Returns the variant dict."""
    import numpy as np
    import torch

    import mke_amd as P
    from oracle.baroclin import manufactured

    prm = dict(sigma=1.6e-2, mu=8e-2, sigma1=1.6e-2, mu1=8e-2, alpha=1.0)
    t = time.perf_counter()
    mesh = P.Mesh.sphere(level)
    a = mesh.arrays()
    xy = a["coords"][:, 0, :]
    u1, u2, f, g = manufactured(xy, 0.0, **prm)
    U1 = np.stack([u1 * (1.0 + 0.01 * k) for k in range(members)])
    U2 = np.stack([u2 * (1.0 - 0.01 * k) for k in range(members)])
    dev = lambda x: torch.as_tensor(np.ascontiguousarray(x), device="cuda")  # noqa: E731
    bc = P.Baroclin(mesh, dev(np.zeros(len(xy))), tau=1e-3, slots=members, **prm)
    log("baroclin: L%d engine for %d members ready in %.1fs" % (level, members,
                                                                time.perf_counter() - t))
    D1, D2 = torch.zeros_like(dev(U1)), torch.zeros_like(dev(U2))
    bc.calc(D1, D2, dev(U1), dev(U2), dev(f), dev(g))   # warm-up (graph capture)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        infos = bc.calc(D1, D2, dev(U1), dev(U2), dev(f), dev(g))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    out = dict(member_steps_per_s=round(members / dt, 3), ms_per_call=round(dt * 1e3, 2),
               level=level, members=members,
               passes=[i["iterations"] for i in infos],
               system_gmres_steps_per_pass=round(float(np.mean(
                   [i["it_system"] / max(1, i["iterations"]) for i in infos])), 2),
               jacobian_gmres_steps_per_pass=round(float(np.mean(
                   [i["it_jacobian"] / max(1, i["iterations"]) for i in infos])), 2),
               converged=all(i["converged"] for i in infos))
    log("baroclin L%d: %.1f member-steps/s (%.1f ms per call of %d), passes %s" %
        (level, out["member_steps_per_s"], out["ms_per_call"], members, out["passes"]))
    bc.close()
    return out


def l7_system(stream):
    """The metric's system: level-7 Helmholtz matrix and b = M f (f of degree
    <= 32, seed 20251202), assembled on the GPU."""
    import torch

    import mke_amd as P
    mesh = P.Mesh.sphere(LEVEL)
    cm, cl = P.helmholtz_coeffs(tau=1.0, sigma=0.0, theta=1.0, mu=TAU)
    A = P.Matrix.assemble(mesh, P.HELMHOLTZ, cm, cl, stream=stream)
    Mm = P.Matrix.assemble(mesh, P.MASS, stream=stream)
    arrays = mesh.arrays()
    f = sph_field(arrays["pr"], 32, 20251202)[arrays["inner"]]
    b = torch.zeros(A.n_rows, dtype=torch.float64, device="cuda")
    Mm.mult_vector(b, torch.tensor(f, dtype=torch.float64, device="cuda"), stream=stream)
    torch.cuda.synchronize()
    return mesh, A, Mm, b


def solve_only_leg(steps, warmup, stream, only=None):
    """The level-7 single right-hand-side forms alone: FP64 MGS as a
    persistent cooperative cycle and as a kernel per MGS pass, FP64 CGS2,
    mixed FP32 MGS (persistent) and CGS2."""
    import torch

    import mke_amd as P
    mesh, A, Mm, b = l7_system(stream)
    out = {}
    for name, kw in (("fp64_mgs_persistent", dict(ortho=P.MGS, persistent=True)),
                     ("fp64_mgs_per_pass", dict(ortho=P.MGS, persistent=False)),
                     ("fp64_cgs2", dict(ortho=P.CGS2)),
                     ("mixed_f32_mgs_persistent", dict(ortho=P.MGS, precision=P.F32)),
                     ("mixed_f32_cgs2", dict(ortho=P.CGS2, precision=P.F32))):
        if only and name != only:
            continue
        x = torch.zeros_like(b)
        w, lat, q50, q99, inf = time_solves(A, b, x, steps, warmup, stream, restart=30,
                                            rtol=1e-10, **kw)
        out[name] = dict(solves_per_s=round(steps / w, 3), p50_ms=round(q50, 3),
                         p99_ms=round(q99, 3), iterations=inf[-1]["iterations"],
                         rel_residual=inf[-1]["rel_residual"], converged=inf[-1]["converged"])
        launches = sum(i["cycle_launches"] for i in inf)
        if launches:
            out[name]["cycle_avg_launch_us"] = round(sum(i["cycle_ms"] for i in inf)
                                                     / launches * 1e3, 3)
            out[name]["cycle_launches_per_solve"] = launches / len(inf)
        log("%s: %.2f solves/s, p50 %.2f ms, p99 %.2f ms, %d iters" %
            (name, out[name]["solves_per_s"], q50, q99, inf[-1]["iterations"]))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(rp, ci, v, b, budget_s):
    """Oracle CSR GMRES(30)-MGS (OpenMP) on the host cores, same A and b, for
    about budget_s seconds on all threads and a bounded sample on one."""
    from oracle import oracle as O
    L = O.lib()
    threads = L.or_num_threads()
    t0 = time.perf_counter()
    n_solves, info = 0, None
    while True:
        x, info = O.gmres(rp, ci, v, b, restart=30, rtol=1e-10)
        n_solves += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    L.or_set_num_threads(1)
    try:
        t1 = time.perf_counter()
        n1 = 0
        while True:
            O.gmres(rp, ci, v, b, restart=30, rtol=1e-10)
            n1 += 1
            if time.perf_counter() - t1 >= budget_s / 2:
                break
        dt1 = time.perf_counter() - t1
    finally:
        L.or_set_num_threads(threads)
    return n_solves / dt, threads, n_solves, dt, info, n1 / dt1, n1, dt1


def cpu_assembly_and_config1(threads):
    """The oracle's row-parallel assembly (generate_matrix's OpenMP row loop)
    of the level-7 Helmholtz matrix, with the C library's cos (the
    reference's arithmetic) and with the correctly rounded one; and the
    config-1 solve (level 4, f of degree <= 8, GMRES(30)-MGS to 1e-10)."""
    from oracle import mesh_ref
    from oracle import oracle as O
    from tests._common import real_sph_harm_field
    L = O.lib()
    out = {}
    om = O.Mesh(mesh_ref.load_text(mesh_ref.sphere_mesh_text(LEVEL)))
    cm, cl = 1.0, -TAU
    for key, libm in (("libm_cos", 1), ("cr_cos", 0)):
        L.or_use_libm_cos(libm)
        try:
            t = time.perf_counter()
            om.assemble(O.KIND_HELMHOLTZ, cm, cl)
            out["assembly_L7_helmholtz_ms_" + key] = round((time.perf_counter() - t) * 1e3, 1)
        finally:
            L.or_use_libm_cos(0)
    a4 = mesh_ref.load_text(mesh_ref.sphere_mesh_text(4))
    m4 = O.Mesh(a4)
    rp4, ci4, v4 = m4.assemble(O.KIND_HELMHOLTZ, cm, cl)
    _, _, vm4 = m4.assemble(O.KIND_MASS)
    import numpy as np
    pr4 = np.asarray(a4["pr"]).reshape(-1, 3)
    b4 = O.spmv(rp4, ci4, vm4, real_sph_harm_field(pr4, 8, 20251202)[m4.inner])
    t = time.perf_counter()
    k = 0
    while time.perf_counter() - t < 1.0:
        _, inf4 = O.gmres(rp4, ci4, v4, b4, restart=30, rtol=1e-10)
        k += 1
    out["config1_L4_solves_per_s"] = round(k / (time.perf_counter() - t), 2)
    out["config1_L4_iterations"] = inf4["iterations"]
    out["threads"] = threads
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of oracle CPU solves for cpu_baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--cold-spmv", action="store_true",
                    help="also time the SpMV round-robin over 40 matrix copies (> 512 MB)")
    ap.add_argument("--serve-requests", type=int, default=256,
                    help="requests of the ensemble serving leg (0: skip it)")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--no-l8", action="store_true",
                    help="skip the level-8 mixed-precision leg")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the ensemble serving leg and print its results (profiling)")
    ap.add_argument("--barvortex-only", type=int, default=0, metavar="LEVEL",
                    help="run only a standalone 8-member BarVortex step at LEVEL (profiling)")
    ap.add_argument("--spmv-only", type=int, default=0, metavar="LEVEL",
                    help="run only the FP64 SpMV timing of the Helmholtz matrix at LEVEL, "
                         "repeated and over 8 copies (profiling)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 256 MiB device copy (calibrates rocprof byte counters)")
    ap.add_argument("--solve-only", action="store_true",
                    help="run only the level-7 single-RHS solve forms (timing, profiling)")
    ap.add_argument("--form", default=None,
                    help="with --solve-only: run this form only (e.g. fp64_mgs_persistent)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (single-GPU boxes); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import mke_amd as P

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    if args.serve_only:
        print(json.dumps({"serve_only": serve_leg(args.serve_requests, args.serve_rate)}))
        return
    if args.barvortex_only:
        print(json.dumps({"barvortex_only": barvortex_leg(args.barvortex_only)}))
        return
    if args.spmv_only:
        mesh = P.Mesh.sphere(args.spmv_only)
        cm, cl = P.helmholtz_coeffs(tau=1.0, sigma=0.0, theta=1.0, mu=TAU)
        A = P.Matrix.assemble(mesh, P.HELMHOLTZ, cm, cl)
        res = l8_spmv_roofline(A, torch.cuda.current_stream())
        if args.calib_copy:
            calib_copy()
        print(json.dumps({"spmv_only": dict(res, level=args.spmv_only)}))
        return
    stream = torch.cuda.current_stream()
    dev = "cuda:0"
    if args.solve_only:
        res = solve_only_leg(args.steps, args.warmup, stream, args.form)
        if args.calib_copy:
            calib_copy()
        print(json.dumps({"solve_only": res}))
        return

    t = time.perf_counter()
    mesh = P.Mesh.sphere(LEVEL)
    log("mesh L%d: %d points in %.2fs" % (LEVEL, mesh.n_points, time.perf_counter() - t))
    cm, cl = P.helmholtz_coeffs(tau=1.0, sigma=0.0, theta=1.0, mu=TAU)
    torch.cuda.synchronize()
    t = time.perf_counter()
    A = P.Matrix.assemble(mesh, P.HELMHOLTZ, cm, cl, stream=stream)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t
    # the same operator again: the first call also pays the process's first
    # use of the assembly kernels and the mesh's device upload
    t = time.perf_counter()
    A_again = P.Matrix.assemble(mesh, P.HELMHOLTZ, cm, cl, stream=stream)
    torch.cuda.synchronize()
    t_asm2 = time.perf_counter() - t
    del A_again
    Mm = P.Matrix.assemble(mesh, P.MASS, stream=stream)
    n, nnz = A.n_rows, A.nnz
    arrays = mesh.arrays()
    f = sph_field(arrays["pr"], 32, 20251202)[arrays["inner"]]
    b = torch.zeros(n, dtype=torch.float64, device=dev)
    Mm.mult_vector(b, torch.tensor(f, dtype=torch.float64, device=dev), stream=stream)
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    log("assembly (Helmholtz, GPU): %.1f ms incl. host structure (first call), %.1f ms again; "
        "n=%d nnz=%d" % (t_asm * 1e3, t_asm2 * 1e3, n, nnz))

    # ---- timed region: plain FP64 GMRES(30), MGS (the reference form), each
    #      restart cycle one cooperative kernel (HIP events around each) ----
    wall, lat, p50, p99, infos = time_solves(A, b, x, args.steps, args.warmup, stream,
                                             restart=30, rtol=1e-10, ortho=P.MGS)
    ok = all(i["converged"] and i["rel_residual"] <= 1e-10 for i in infos)
    iters = infos[-1]["iterations"]
    solves_s = args.steps / wall
    launches = sum(i["cycle_launches"] for i in infos)
    cyc_ms = sum(i["cycle_ms"] for i in infos)
    solve_bytes = gmres_solve_bytes(n, nnz, iters)
    cyc = None
    if launches:
        per_solve = launches / len(infos)
        cyc = dict(avg_launch_us=round(cyc_ms / launches * 1e3, 3),
                   launches_per_solve=per_solve,
                   bytes_per_launch=round(solve_bytes / per_solve),
                   cycle_kernel_ms_per_solve=round(cyc_ms / len(infos), 4))
        cyc["achieved"] = round(cyc["bytes_per_launch"] / (cyc_ms / launches * 1e-3) / 1e9, 1)
    log("fp64-mgs: %.2f solves/s, p50 %.2f ms, p99 %.2f ms, %d iters, rel %.2e, cycles %s" %
        (solves_s, p50, p99, iters, infos[-1]["rel_residual"], cyc))
    x64 = x.cpu().numpy()

    # ---- roofline of the FP64 SpMV (Solver::mult_vector), right after the
    #      timed region: median of 5 runs of 400 back-to-back launches ----
    y = torch.zeros_like(b)
    reps = sorted(kernel_avg_ms(lambda: A.mult_vector(y, b, stream=stream), 400, stream)
                  for _ in range(5))
    spmv_ms = reps[2]
    bspmv = spmv_bytes(n, nnz, 8, 8)
    achieved = bspmv / (spmv_ms * 1e-3) / 1e9
    log("spmv fp64: %.2f us per launch (%.0f GB/s; runs %s us)" %
        (spmv_ms * 1e3, achieved, [round(r * 1e3, 2) for r in reps]))

    rp, ci, v = A.csr()
    bh = b.cpu().numpy()

    # ---- variants ----
    variants = {}
    if not args.no_variants:
        for name, kw in (("fp64_mgs_per_pass", dict(ortho=P.MGS, persistent=False)),
                         ("fp64_cgs2", dict(ortho=P.CGS2)),
                         ("mixed_f32_cgs2", dict(ortho=P.CGS2, precision=P.F32)),
                         ("mixed_f32_mgs", dict(ortho=P.MGS, precision=P.F32)),
                         ("mixed_bf16_cgs2", dict(ortho=P.CGS2, precision=P.BF16))):
            xv = torch.zeros_like(b)
            w2, l2, q50, q99, inf2 = time_solves(A, b, xv, max(5, args.steps // 2), 2, stream,
                                                 restart=30, rtol=1e-10, **kw)
            variants[name] = dict(solves_per_s=round(len(inf2) / w2, 3), p50_ms=round(q50, 3),
                                  p99_ms=round(q99, 3), iterations=inf2[-1]["iterations"],
                                  rel_residual=inf2[-1]["rel_residual"],
                                  converged=inf2[-1]["converged"], x=xv)
            log("%s: %.2f solves/s, p50 %.2f ms, %d iters, rel %.2e, conv %s" %
                (name, variants[name]["solves_per_s"], q50, inf2[-1]["iterations"],
                 inf2[-1]["rel_residual"], inf2[-1]["converged"]))

    # ---- batched: 64 right-hand sides, block GMRES(30), FP32 basis + FP64 refinement ----
    batched = None
    if not args.no_variants:
        F = np.stack([sph_field(arrays["pr"], 32, 20251202 + q)[arrays["inner"]]
                      for q in range(64)], 1)
        Fd = torch.tensor(F, dtype=torch.float64, device=dev)
        B64 = torch.zeros_like(Fd)
        for q in range(64):  # b_q = M f_q through the library's SpMV
            col = torch.zeros(n, dtype=torch.float64, device=dev)
            Mm.mult_vector(col, Fd[:, q].contiguous(), stream=stream)
            B64[:, q] = col
        X64 = torch.zeros_like(B64)
        A.solve_block(X64, B64, rtol=1e-10, ortho=P.CGS1, stream=stream)  # warmup
        torch.cuda.synchronize()
        nb = max(2, args.steps // 5)
        s_ev, e_ev = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        lat_b = []
        t0 = time.perf_counter()
        for _ in range(nb):
            s_ev.record(stream)
            binfo = A.solve_block(X64, B64, rtol=1e-10, ortho=P.CGS1, stream=stream)
            e_ev.record(stream)
            torch.cuda.synchronize()
            lat_b.append(s_ev.elapsed_time(e_ev))
        wb = time.perf_counter() - t0
        batched = dict(solves_per_s=round(64 * nb / wb, 3), rhs=64, batch_ms=round(wb / nb * 1e3, 3),
                       block_iterations=binfo["iterations"], cycles=binfo["cycles"],
                       max_rel_residual=binfo["max_rel_residual"], converged=binfo["converged"],
                       x0=X64[:, 0].cpu().numpy())
        # block Gram-Schmidt on MFMA: step j of a 30-step cycle does V^T W and
        # W -= V H over j+1 basis blocks, 2 * 8192 n (j+1) FP32 flops
        its, full = binfo["iterations"], binfo["iterations"] // 30
        rem = its - 30 * full
        gs_flops = 16384.0 * n * (full * 465 + rem * (rem + 1) // 2)
        batched["gs_flops_per_solve"] = gs_flops
        batched["gs_tflops_over_solve"] = round(gs_flops / (wb / nb) / 1e12, 2)
        batched["gs_mfma_frac_over_solve"] = round(gs_flops / (wb / nb) / 157.3e12, 4)
        log("batched64: %.1f solves/s (%.1f ms per 64), %d block iters, max rel %.2e" %
            (batched["solves_per_s"], batched["batch_ms"], binfo["iterations"],
             binfo["max_rel_residual"]))
        # the same 64 systems as independent columns sharing each SpMM
        for name, prec, orth in (("batched64_f32_gmres_cgs2", P.F32, P.CGS2),
                                 ("batched64_f32_gmres_cgs1", P.F32, P.CGS1),
                                 ("batched64_f64_gmres_cgs2", P.F64, P.CGS2)):
            rpt, tune_ms = A.tune_batch(64, restart=30, ortho=orth, precision=prec,
                                        stream=stream)
            log("%s: autotuned %d rows per thread (cycle ms %s)" %
                (name, rpt, {k: round(v, 2) for k, v in tune_ms.items()}))
            Xc = torch.zeros_like(B64)
            A.solve_batch(Xc, B64, rtol=1e-10, precision=prec, ortho=orth, stream=stream)  # warmup
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(nb):
                cinfo = A.solve_batch(Xc, B64, rtol=1e-10, precision=prec, ortho=orth,
                                      stream=stream)
            torch.cuda.synchronize()
            wc = time.perf_counter() - t0
            variants[name] = dict(
                solves_per_s=round(64 * nb / wc, 3), rhs=64, batch_ms=round(wc / nb * 1e3, 3),
                max_iterations=max(i["iterations"] for i in cinfo),
                max_rel_residual=max(i["rel_residual"] for i in cinfo),
                converged=all(i["converged"] for i in cinfo), x=Xc[:, 0].clone(),
                rows_per_thread=rpt,
                tune_cycle_ms={str(k): round(v, 3) for k, v in tune_ms.items()})
            log("%s: %.1f solves/s (%.1f ms per 64), max iters %d, max rel %.2e" %
                (name, variants[name]["solves_per_s"], variants[name]["batch_ms"],
                 variants[name]["max_iterations"], variants[name]["max_rel_residual"]))

    # ---- multigrid-preconditioned forms (V-cycle over levels 2..7 of the same
    #      Helmholtz operator, bit-identical fine matrix), 1 and 64 right-hand sides ----
    if not args.no_variants:
        mg = P.Multigrid(mesh, P.HELMHOLTZ, cm, cl, stream=stream)
        for name, k, prec in (("fp64_mg_gmres_cgs2", 1, P.F64),
                              ("mixed_f32_mg_gmres_cgs2", 1, P.F32),
                              ("batched64_f64_mg_gmres_cgs2", 64, P.F64),
                              ("batched64_f32_mg_gmres_cgs2", 64, P.F32)):
            Bk = b.reshape(-1, 1).clone() if k == 1 else B64
            Xk = torch.zeros_like(Bk)
            mg.solve_batch(Xk, Bk, rtol=1e-10, precision=prec, stream=stream)  # warmup
            torch.cuda.synchronize()
            reps = max(10, args.steps // (1 if k == 1 else 2))
            lat_k = []
            t0 = time.perf_counter()
            for _ in range(reps):
                t1 = time.perf_counter()
                minfo = mg.solve_batch(Xk, Bk, rtol=1e-10, precision=prec, stream=stream)
                torch.cuda.synchronize()
                lat_k.append((time.perf_counter() - t1) * 1e3)
            wk = time.perf_counter() - t0
            lat_k.sort()
            variants[name] = dict(
                solves_per_s=round(k * reps / wk, 3), rhs=k, batch_ms=round(wk / reps * 1e3, 3),
                p50_ms=round(lat_k[len(lat_k) // 2], 3),
                max_iterations=max(i["iterations"] for i in minfo),
                max_rel_residual=max(i["rel_residual"] for i in minfo),
                converged=all(i["converged"] for i in minfo), x=Xk[:, 0].clone(),
                preconditioner="multigrid V-cycle, levels 2..7")
            log("%s: %.1f solves/s (%.2f ms per call of %d), max iters %d, max rel %.2e" %
                (name, variants[name]["solves_per_s"], variants[name]["batch_ms"], k,
                 variants[name]["max_iterations"], variants[name]["max_rel_residual"]))
        mg.close()

    # ---- level 8: mixed precision vs FP64 ----
    if not args.no_variants and not args.no_l8:
        variants.update(l8_leg(max(2, args.steps // 10), not args.no_cpu))

    # ---- Baroclin: two-layer implicit step, 8 members at level 5 ----
    if not args.no_variants:
        variants["baroclin_L5_8members"] = baroclin_leg()

    # ---- serving: ensemble BarVortex steps on mixed levels ----
    if not args.no_variants and args.serve_requests > 0:
        variants.update(serve_leg_child(args.serve_requests, args.serve_rate))

    cold_ms = cold_gbs = None
    if args.cold_spmv:
        # cold-cache variant: round-robin over 40 copies (> 512 MB working set)
        copies = [P.Matrix.from_csr(n, n, rp, ci, v) for _ in range(40)]
        xs = [torch.randn(n, dtype=torch.float64, device=dev) for _ in range(40)]
        ys = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(40)]
        for c, xx, yy in zip(copies, xs, ys):  # builds every copy's ELL before timing
            c.mult_vector(yy, xx, stream=stream)
        torch.cuda.synchronize()
        it = [0]

        def cold():
            k = it[0] % 40
            it[0] += 1
            copies[k].mult_vector(ys[k], xs[k], stream=stream)

        cold_ms = kernel_avg_ms(cold, 400, stream)
        cold_gbs = bspmv / (cold_ms * 1e-3) / 1e9
        log("spmv fp64 cold (40 copies): %.2f us (%.0f GB/s)" % (cold_ms * 1e3, cold_gbs))
    if args.calib_copy:
        calib_copy()

    # ---- forward error vs the FP64 oracle + CPU baseline ----
    cpu = None
    fwd = {}
    if not args.no_cpu:
        from oracle import oracle as O
        xt, _ = O.gmres(rp, ci, v, bh, restart=30, rtol=1e-14, max_iters=400000)
        nt = np.linalg.norm(xt)
        fwd["fp64_mgs"] = float(np.linalg.norm(x64 - xt) / nt)
        for name, d in variants.items():
            if d.get("x") is not None:
                fwd[name] = float(np.linalg.norm(d["x"].cpu().numpy() - xt) / nt)
        if batched is not None:
            batched["fwd_err_vs_oracle_col0"] = float(np.linalg.norm(batched["x0"] - xt) / nt)
        rate, threads, ns, dt, cinfo, rate1, n1, dt1 = cpu_baseline(rp, ci, v, bh,
                                                                    args.cpu_budget)
        extra = cpu_assembly_and_config1(threads)
        cpu = {"value": round(rate, 4), "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": "%d oracle FP64 GMRES(30)-MGS solves of the same L7 system "
                         "(%d iterations each) in %.1f s, OpenMP on %d threads; %d in %.1f s "
                         "on 1 thread" % (ns, cinfo["iterations"], dt, threads, n1, dt1),
               "cpu_model": cpu_model(), "value_1_thread": round(rate1, 4)}
        cpu.update(extra)
        log("cpu baseline: %.3f solves/s on %d threads, %.3f on 1; %s" %
            (rate, threads, rate1, extra))
    for d in variants.values():
        d.pop("x", None)
    if batched is not None:
        batched.pop("x0", None)
        variants["batched64_f32_block_gmres_bcgs1"] = batched
    for name in variants:
        if name in fwd:
            variants[name]["fwd_err_vs_oracle"] = fwd[name]

    out = {
        "metric": "Helmholtz solves/s @1e-10 rel residual, mixed vs FP64, L7 mesh; p50/p99 latency",
        "value": round(solves_s, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "L7 icosahedral sphere Helmholtz (M - 1e-2 L) u = M f, "
                               "GPU-assembled, FP64 GMRES(30)-MGS, rtol 1e-10, x0=0",
                   "level": LEVEL, "rows": n, "nnz": nnz, "rhs": "real Y_lm, l<=32, seed 20251202",
                   "parallelism": "single GPU"},
        "p50_ms": round(p50, 4),
        "p99_ms": round(p99, 4),
        "iterations_per_solve": iters,
        "converged_all": ok,
        "fwd_err_vs_oracle": fwd.get("fp64_mgs"),
        "assembly_ms": round(t_asm * 1e3, 3),
        "assembly_again_ms": round(t_asm2 * 1e3, 3),
        "variants": variants,
        "roofline": None,
        "cpu_baseline": cpu,
    }
    # the timed region's dominant kernel: the persistent MGS restart cycle
    # (one launch per cycle), algorithmic bytes of the GMRES(30) model over
    # its average duration from the HIP events around each launch
    whole = dict(bytes_per_solve=solve_bytes,
                 achieved=round(solve_bytes / (wall / args.steps) / 1e9, 1))
    whole["frac"] = round(whole["achieved"] / HBM_PEAK_GBS, 4)
    spmv = {"kernel": "ell_spmv_f64 (Solver::mult_vector), level 7, in the Infinity Cache",
            "bytes_per_launch": bspmv, "avg_launch_us": round(spmv_ms * 1e3, 3),
            "achieved": round(achieved, 1), "frac_in_cache": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles("ell_spmv_f64_L7_traffic.json")}
    l8s = variants.get("L8_spmv_fp64")
    if l8s:
        # the HBM figure: level 8, the matrix streamed from HBM (8 copies)
        spmv["level8_stream"] = dict(
            kernel="ell_spmv_f64_stream", bytes_per_launch=l8s["bytes_per_launch"],
            streamed_us=l8s["cold_us"], streamed_achieved=l8s["cold_gbs"],
            streamed_frac=l8s["cold_frac"], streamed_copies=l8s["copies"],
            repeated_achieved=l8s["warm_gbs"], repeated_frac=l8s["warm_frac"],
            traffic=traffic_from_profiles("ell_spmv_f64_stream_L8_traffic.json"))
    if cold_ms is not None:
        spmv.update(cold_achieved=round(cold_gbs, 1), cold_frac=round(cold_gbs / HBM_PEAK_GBS, 4),
                    cold_avg_launch_us=round(cold_ms * 1e3, 3))
    if cyc is not None:
        out["roofline"] = {
            "bound": "hbm", "kernel": "k_mgs_cycle<double,double> (one GMRES(30) restart cycle, "
                                      "MGS; latency-bound chain of dependent global sums)",
            "achieved": cyc["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(cyc["achieved"] / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles("k_mgs_cycle_L7_traffic.json"),
            "bytes_per_launch": cyc["bytes_per_launch"], "avg_launch_us": cyc["avg_launch_us"],
            "launches_per_solve": cyc["launches_per_solve"],
            "whole_solve": whole, "spmv": spmv}
    else:
        out["roofline"] = {"bound": "hbm", "kernel": spmv["kernel"], "achieved": achieved,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": spmv["traffic"],
                           "whole_solve": whole, "spmv": spmv}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
