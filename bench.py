#!/usr/bin/env python3
"""Benchmark of the element-assembly / matrix-free HVP / Jacobi-PCG Newton path on one MI355X.

Headline workload (default): compressible Neo-Hookean on a structured 160^3 Hex8
unit cube (4,096,000 elements, 4,173,281 nodes, 12,442,080 free dofs), 2x2x2
Gauss, mu = 1, lambda = 10, x = 0 clamped, u_x prescribed on x = 1 in 5 load
steps to a stretch of 1.3 (u_x = 0.3).  Solver: mixed-precision Newton (FP64
residual, energy and outer update; FP32 matrix-free HVP and Jacobi-PCG to 1e-3
for each correction), Newton to 1e-8 relative residual per load step.

One step = one load step solved to 1e-8 (homogeneous-stretch predictor, new
u_x on x = 1, Newton); steps cycle through the 5 load steps and restart from
u = 0 after the last.  Setup (mesh upload, tiling, constraints) is not timed,
as in the reference's setup/compile/run split
(tests/test_colored_matrix_benchmark.py:97-113).

metric value: Newton DOF-iterations/s = n_free * (inner CG iterations) / wall
seconds of the timed steps.  Alongside: the plain FP64 GPU path (colored FP64
Hessian -> CSR, FP64 Jacobi-PCG to 1e-10, Newton to 1e-8) on the same five load
steps, the FP64 oracle's residual at both final states, device memory of both,
the roofline of the dominant kernel of the mixed path (the FP32 HVP tile
kernel), timed with HIP events on its own stream, and a CPU baseline.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue peak of the part, in G wave64 instructions/s: one wave64 instruction per
# SIMD every two cycles on 1,024 SIMDs (256 CUs x 4) at 2.4 GHz.  Two cycles is what
# FP32 v_add_f32 / v_mul_f32 / v_fmac_f32 chains measure on gfx950 (`--workload valu`:
# 2.29-2.32 cycles per instruction including the launch ramp); v_fma_f32 with two
# scalar operands takes four (4.24), so a kernel's own floor lies between the two.
VALU_PEAK_GINST = 1024 * 2.4 / 2

# Counters of the headline's dominant kernel from its rocprofv3 profile (per launch):
# VALU wave-instructions (SQ_INSTS_VALU) and HBM bytes (FETCH_SIZE x 2 + WRITE_SIZE x 1,
# the factors of a plain 512 MiB device copy under the same counters).  The profile
# holds them as "key: value" lines; the bench divides them by the live launch time.
HEADLINE_PROFILE = os.path.join(ROOT, "profiles", "k_tile_assemble_hvp_hex8_nh_160cubed.md")


def profile_counters(path):
    """{key: float} of the "- key: value" lines of a committed profile summary."""
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        line = line.strip()
        if line.startswith("- ") and ":" in line:
            k, v = line[2:].split(":", 1)
            try:
                out[k.strip()] = float(v.split()[0].replace(",", ""))
            except (ValueError, IndexError):
                pass
    return out

# The C2 line's dominant kernel, the node-block SpMV at 64^3: its profile holds
# the HBM bytes per launch (FETCH_SIZE x 2 + WRITE_SIZE x 1) as a "- key: value" line.
C2_PROFILE = os.path.join(ROOT, "profiles", "k_pcg_bspmv_hex8_le_64cubed.md")


def build_problem(n):
    import numpy as np

    import tatva_amd as ta
    from tatva_amd import energies, lifter, solvers, sparse
    from tatva_amd.serve import traction_loads

    mesh = ta.Mesh.unit_cube(n)
    op = ta.Operator(mesh, ta.Hexahedron8())
    dens = energies.LinearElasticity.from_youngs_poisson(1.0, 0.3)
    f = traction_loads(mesh.coords, n)
    E = op.energy(dens, external_force=f)
    x0 = np.where(np.isclose(mesh.coords[:, 0], 0.0))[0]
    L = lifter.Lifter(mesh.n_nodes * 3, lifter.Fixed((x0[:, None] * 3 + np.arange(3)).ravel(), 0.0))
    t0 = time.perf_counter()
    pat = L.adapt_sparsity(sparse.pattern_from_mesh(mesh, 3))
    t1 = time.perf_counter()
    cm = sparse.ColoredMatrix.from_csr(pat)
    t2 = time.perf_counter()
    solver = solvers.Newton(E, L, mode="csr", tol=1e-9, lin_tol=1e-10, chunk=64, colored_matrix=cm)
    return dict(mesh=mesh, op=op, E=E, L=L, cm=cm, solver=solver, f=f, dens=dens,
                pattern_s=t1 - t0, coloring_s=t2 - t1)


def stretch_problem(n, kind="hex8", mu=1.0, lmbda=10.0, jitter=0.0, seed=0):
    """Compressible Neo-Hookean cube [0,1]^3: x = 0 clamped, u_x prescribed on x = 1
    (the load), u_y, u_z free there."""
    import numpy as np

    import tatva_amd as ta
    from tatva_amd import energies, lifter

    if kind == "hex8":
        mesh = ta.Mesh.unit_cube(n)
        el = ta.Hexahedron8()
    else:
        mesh = ta.Mesh.unit_cube(n, type="tetrahedron", jitter=jitter, seed=seed)
        el = ta.Tetrahedron4()
    op = ta.Operator(mesh, el)
    E = op.energy(energies.NeoHookean(mu, lmbda))
    x0 = np.where(np.isclose(mesh.coords[:, 0], 0.0))[0]
    x1 = np.where(np.isclose(mesh.coords[:, 0], 1.0))[0]
    dofs = np.concatenate([(x0[:, None] * 3 + np.arange(3)).ravel(), x1 * 3])
    L = lifter.Lifter(mesh.n_nodes * 3, lifter.Fixed(dofs, lifter.RuntimeValue("load")))

    def values(d):
        return np.concatenate([np.zeros(x0.size * 3), np.full(x1.size, d)])

    return dict(mesh=mesh, op=op, E=E, L=L, values=values)


def solve_load_steps(P, solver, n_steps, stretch, upto=None):
    """Load-stepped Newton: each step adds a homogeneous stretch increment
    (u_x += dx * x) as predictor, sets u_x = stretch * s / n_steps on x = 1 and
    solves; with upto, only the first upto of the n_steps steps.  Returns
    (u, [NewtonStats])."""
    import torch

    mesh = P["mesh"]
    ramp = torch.zeros(mesh.n_nodes * 3, dtype=torch.float64, device="cuda")
    ramp[0::3] = torch.as_tensor(mesh.coords[:, 0], device="cuda")
    u = torch.zeros(mesh.n_nodes * 3, dtype=torch.float64, device="cuda")
    out = []
    for s in range(1, (upto or n_steps) + 1):
        u.add_(ramp, alpha=stretch / n_steps)
        L = P["L"].at["load"].set(P["values"](stretch * s / n_steps))
        out.append(solver.solve(u, L))
        if P.get("progress"):
            st = out[-1]
            print(f"[{solver.mode}] load step {s}/{n_steps}: {st.newton_iterations} Newton, "
                  f"{st.cg_iterations} CG, {st.seconds:.2f} s", file=sys.stderr, flush=True)
    return u, out


def mixed_vs_fp64(E, L0, run, tol=1e-8, colored_matrix=None, keep_solutions=False, warmup=True,
                  modes=("csr", "mixed")):
    """The same problem solved by FP64 colored-CSR Newton and by mixed Newton
    (FP32 matrix-free PCG inside FP64 Newton/iterative refinement), both to
    ||r|| <= tol ||r_0||.  ``run(solver)`` performs the solve and returns
    (u, [NewtonStats]); it is called once untimed first.  Device memory per
    solver: the peak torch reserved during its solve, and the device memory in
    use (total - free, library buffers included) with the solver alive, above
    the level before it was built."""
    import gc

    import torch

    from tatva_amd import solvers

    res = {}
    sols = {}
    for mode in modes:
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        torch.cuda.reset_peak_memory_stats()
        free0, _total = torch.cuda.mem_get_info()
        kw = dict(lin_tol=1e-10, colored_matrix=colored_matrix) if mode == "csr" \
            else dict(inner_tol=1e-3)
        t_build = time.perf_counter()
        solver = solvers.Newton(E, L0, mode=mode, tol=tol, chunk=64, **kw)
        print(f"[{mode}] solver set up in {time.perf_counter() - t_build:.1f} s", file=sys.stderr,
              flush=True)
        if warmup:
            run(solver)     # warm-up (workspaces)
        solvers.set_profiling(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u, stats = run(solver)
        torch.cuda.synchronize()
        solvers.set_profiling(False)
        dt = time.perf_counter() - t0
        n_free = solver.red.size
        cg = sum(s.cg_iterations for s in stats)
        op_ms = sum(s.op_ms for s in stats)
        op_main = sum(s.op_main_ms for s in stats)
        op_n = sum(s.op_samples for s in stats)
        res[mode] = {
            "seconds": dt, "newton_iterations": sum(s.newton_iterations for s in stats),
            "cg_iterations": cg, "dof_iters_per_s": n_free * cg / dt,
            "converged": all(s.converged for s in stats),
            "worst_rel_residual": max(s.residual_history[-1] / s.residual_history[0]
                                      for s in stats),
            "operator_us": (op_ms / op_n * 1e3) if op_n else None,
            # the operator's first kernel alone: the HVP tile kernel / the SpMV
            "operator_main_kernel_us": (op_main / op_n * 1e3) if op_n else None,
        }
        free1, _total = torch.cuda.mem_get_info()
        res[mode]["device_gb_in_use"] = (free0 - free1) / 1e9
        res[mode]["torch_peak_reserved_gb"] = torch.cuda.max_memory_reserved() / 1e9
        sols[mode] = u
        del solver
    if keep_solutions:
        for mode in modes:
            res[f"_u_{mode}"] = sols[mode]
    if "csr" in sols and "mixed" in sols:
        a, b = sols["mixed"], sols["csr"]
        res["mixed_vs_fp64_rel_error"] = float((a - b).abs().max() / b.abs().max())
        res["dof_iters_speedup"] = res["mixed"]["dof_iters_per_s"] / res["csr"]["dof_iters_per_s"]
        res["time_to_solution_speedup"] = res["csr"]["seconds"] / res["mixed"]["seconds"]
    return res


def c1_latency(reps=20):
    """C1: Tri3 Poisson on a 64x64 unit square, u = 0 on the boundary, FP64
    colored CSR + PCG to 1e-10; median wall time per Newton solve."""
    import numpy as np
    import torch

    import tatva_amd as ta
    from tatva_amd import energies, lifter, solvers

    mesh = ta.Mesh.unit_square(64, 64)
    op = ta.Operator(mesh, ta.Tri3())
    E = op.energy(energies.Poisson(1.0))
    bnd = np.where(np.isclose(mesh.coords, 0).any(1) | np.isclose(mesh.coords, 1).any(1))[0]
    L = lifter.Lifter(mesh.n_nodes, lifter.Fixed(bnd, 0.0))
    solver = solvers.Newton(E, L, mode="csr", tol=1e-9, lin_tol=1e-10, chunk=32)
    times, its = [], 0
    for r in range(reps + 2):
        u = torch.zeros(mesh.n_nodes, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = solver.solve(u)
        torch.cuda.synchronize()
        if r >= 2:
            times.append(time.perf_counter() - t0)
            its = st.cg_iterations
    return {"median_ms": float(np.median(times)) * 1e3, "p99_ms": float(np.percentile(times, 99)) * 1e3,
            "cg_iterations": its, "n_free": solver.red.size}


def serve_stream(count, qps, seed=0, max_slots=64, chunk=8, hbm_budget=200e9, prefetch=2):
    """Continuous-batching server on the mixed Hex8 LE / Neo-Hookean request
    stream (n log-uniform in [8, 48]); first an untimed warm-up request of
    each material at a size inside the stream's range, so that one-time
    costs of the process (kernel loading, allocator growth, numpy's first
    large-array temporary check, which walks torch's shared libraries) land
    in neither path's first request."""
    from tatva_amd import serve

    srv = serve.Server(max_slots=max_slots, chunk=chunk, hbm_budget=hbm_budget, prefetch=prefetch)
    srv.run([serve.SolveRequest(rid=-2, n=40, material="linear_elastic", E=1.0, nu=0.3),
             serve.SolveRequest(rid=-1, n=40, material="neo_hookean", E=1.0, nu=0.3,
                                load_steps=2)], realtime=False)
    reqs = serve.make_stream(count, qps=qps, seed=seed)
    rep = srv.run(reqs, realtime=True)
    srv.close()
    s = rep.summary()
    s.update({
        "qps": qps, "max_slots": max_slots, "chunk": chunk, "hbm_budget_gb": hbm_budget / 1e9,
        "prefetch_threads": prefetch,
        "workload": f"{count} requests, Poisson arrivals at {qps}/s, hex8 cubes n in [8, 48] "
                    "log-uniform, E ~ LogNormal(0, 0.5), nu ~ U(0.2, 0.45), linear elastic "
                    "(traction 1e-3) / Neo-Hookean (2 load steps of u_x += 0.03) 50/50, "
                    "Newton to 1e-8 relative, FP32 batched PCG to 1e-3 per Newton step",
        "total_dofs": int(sum(r.n_dofs for r in reqs)),
        "all_converged": all(r.converged for r in rep.results),
    })
    print(f"[serve] {count} requests at {qps}/s, budget {hbm_budget / 1e9:.2f} GB: "
          f"p50 {s['p50_ms']:.1f} ms, p99 {s['p99_ms']:.1f} ms, {s['throughput_req_per_s']:.1f} "
          f"req/s, peak {s['peak_admitted']} slots", file=sys.stderr, flush=True)
    return s


def serve_sweep(args):
    """The serving stream at each QPS of --serve-qps (--serve-requests requests
    each, 200 GB budget), then once more at the highest QPS under the binding
    budget --serve-budget-gb (admissions wait for memory)."""
    out = [serve_stream(args.serve_requests, q, chunk=args.serve_chunk,
                        prefetch=args.serve_prefetch) for q in args.serve_qps]
    if args.serve_budget_gb > 0:
        out.append(serve_stream(args.serve_requests, max(args.serve_qps), chunk=args.serve_chunk,
                                hbm_budget=args.serve_budget_gb * 1e9,
                                prefetch=args.serve_prefetch))
    return out


def cpu_baseline(K_host, b_host, seconds=12.0, max_iters=400):
    """Oracle Jacobi-PCG (SciPy CSR SpMV + NumPy, one thread) on the same matrix,
    run for a bounded number of iterations; returns (DOF-iters/s, iterations, s)."""
    import numpy as np

    A = K_host
    dinv = 1.0 / A.diagonal()
    x = np.zeros_like(b_host)
    r = b_host.copy()
    z = dinv * r
    p = z.copy()
    rz = r @ z
    it = 0
    t0 = time.perf_counter()
    while it < max_iters and time.perf_counter() - t0 < seconds:
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        z = dinv * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        it += 1
    dt = time.perf_counter() - t0
    return A.shape[0] * it / dt, it, dt


def large_mixed(workload, oracle_check=True, upto=None, modes=("csr", "mixed")):
    """C3 (Tet4 Neo-Hookean 96^3 x 6, jittered, 10 load steps to stretch 1.3) or
    C5 (Hex8 Neo-Hookean 160^3, 5 load steps to stretch 1.3): mixed Newton vs
    the FP64 colored-CSR Newton on the same problem, both to 1e-8 relative
    (upto: only the first upto load steps, for short profiled runs)."""
    import numpy as np

    if workload == "c3":
        n, kind, steps, jitter = 96, "tet4", 10, 0.2
    else:
        n, kind, steps, jitter = 160, "hex8", 5, 0.0
    stretch = 0.3
    t0 = time.perf_counter()
    NH = stretch_problem(n, kind=kind, jitter=jitter, seed=0)
    NH["progress"] = True
    L0 = NH["L"].at["load"].set(NH["values"](0.0))
    setup = time.perf_counter() - t0
    print(f"{workload}: mesh {NH['mesh'].n_elements} elements, {setup:.1f} s", file=sys.stderr,
          flush=True)
    # no warm-up solve at this size: the first Newton step's set-up (graph
    # capture, workspaces) is a negligible share of the timed solve
    done = min(upto or steps, steps)
    res = mixed_vs_fp64(NH["E"], L0, lambda sv: solve_load_steps(NH, sv, steps, stretch, done),
                        keep_solutions=True, warmup=False, modes=modes)
    mesh = NH["mesh"]
    res["workload"] = (f"{kind} Neo-Hookean {n}^3{' x 6 tets' if kind == 'tet4' else ''}, "
                       f"mu=1, lambda=10, jitter={jitter}h (seed 0), x=0 clamped, u_x on x=1 "
                       f"in {steps} load steps to {stretch}"
                       f"{f' (first {done} run)' if done < steps else ''}, "
                       f"Newton to 1e-8 relative per step")
    res["elements"] = mesh.n_elements
    res["nodes"] = mesh.n_nodes
    res["setup_s"] = setup
    if oracle_check:
        from oracle import fem  # the checker only

        el = fem.Tet4() if kind == "tet4" else fem.Hex8()
        L_end = NH["L"].at["load"].set(NH["values"](stretch * done / steps))
        free = L_end.free_dofs
        # the oracle's element algebra in C over the host's threads
        Q = fem.ProblemOMP(el, mesh.coords, mesh.elements, fem.NeoHookean(1.0, 10.0),
                           threads=cpu_threads())
        # scale: the oracle residual of the final load applied to the undeformed body
        # (free dofs zero, prescribed dofs at their values)
        u_load = L_end.lift_from_zeros(np.zeros(free.size)).cpu().numpy()
        r_load = np.linalg.norm(Q.residual(u_load)[free])
        res["oracle_load_residual_norm"] = float(r_load)
        for mode in modes:
            uh = res.pop(f"_u_{mode}").cpu().numpy()
            r = Q.residual(uh)[free]
            res[mode]["oracle_final_residual_norm"] = float(np.linalg.norm(r))
            res[mode]["oracle_rel_residual"] = float(np.linalg.norm(r) / r_load)
    else:
        res.pop("_u_mixed", None)
        res.pop("_u_csr", None)
    return res


class LoadStepper:
    """Steps of the headline: load step k (1..n_steps) = predictor u_x += dx * x,
    u_x = k * dx on x = 1, Newton to the solver's tolerance.  After the last load
    step the next step restarts from u = 0."""

    def __init__(self, P, solver, n_steps, stretch):
        import torch

        mesh = P["mesh"]
        self.P, self.solver, self.n_steps, self.stretch = P, solver, n_steps, stretch
        self.ramp = torch.zeros(mesh.n_nodes * 3, dtype=torch.float64, device="cuda")
        self.ramp[0::3] = torch.as_tensor(mesh.coords[:, 0], device="cuda")
        self.u = torch.zeros(mesh.n_nodes * 3, dtype=torch.float64, device="cuda")
        self.k = 0          # load steps done in the current pass

    def load(self, k):
        return self.stretch * k / self.n_steps

    def step(self):
        if self.k == self.n_steps:
            self.u.zero_()
            self.k = 0
        self.k += 1
        self.u.add_(self.ramp, alpha=self.stretch / self.n_steps)
        L = self.P["L"].at["load"].set(self.P["values"](self.load(self.k)))
        return self.solver.solve(self.u, L)


def device_gb_in_use():
    import torch

    free, total = torch.cuda.mem_get_info()
    return (total - free) / 1e9


def mesh_is_affine(op):
    return bool(op.device_layout().get("hex_affine"))


def hvp_tile_bytes(op, dens_ncomp=3, cached_lin=0, geo=False, defer_n=0, mapped=True):
    """Algorithmic bytes of one FP32 HVP tile-kernel launch (k_tile_assemble,
    matrix-free PCG operator), from the mesh's device layout: per element its
    nne uint16 local node numbers (+ cached linearisation floats per quadrature
    point, + its 44-byte geometry cache entry, if any), per (tile, local node)
    slot its global node (int32), the node's FP64 coordinates (not with the
    geometry cache), FP32 state and tangent, the tangent's dof-map entries
    (int32) and the slot's FP32 partial sums written and the int32 position it
    is written to, plus per tile its color offsets and slot range.  Coordinates
    are counted per slot because tiles sharing a node each load it.  With the
    CG step's deferred update (defer_n entries of the CG vectors) the launch
    also streams z, p and x in and p and x out: 20 bytes per entry.  Without a
    dof map (the PCG on full-length vectors: mapped False) the tangent is read
    by node and no dof-map entries are loaded.  Hex8 and Tet4 tiles sum by
    per-tile incidence lists (a uint16 entry per element node, a uint16 list
    end per slot) and read only the tile's element count of its color
    offsets; Tri3 tiles read every color offset."""
    lay = op.device_layout()
    E, S, T = lay["n_elements"], lay["n_slots"], lay["n_tiles"]
    nne = 8 if op.element.kind == 2 else (4 if op.element.kind == 1 else 3)
    d = 3 if nne != 3 else 2
    nc = dens_ncomp
    b = 2 * nne * E + 4 * cached_lin * lay["n_quad"] * E + (44 * E if geo else 0)
    b += S * (4 + (0 if geo else 8 * d) + 4 * nc * (0 if cached_lin else 1)
              + (4 if mapped else 0) * nc + 4 * nc + 4 * nc + 4)
    incidence = nne in (4, 8)
    if incidence:
        b += 2 * nne * E + 2 * S + T * (2 + 8)
    else:
        b += T * (2 * (lay["max_colors"] + 1) + 8)
    b += 20 * defer_n
    return int(b), lay


def headline(args):
    """C5: mixed Newton on the Hex8 Neo-Hookean 160^3 cube (the timed line),
    then the FP64 colored-CSR Newton on the same load steps, oracle residuals."""
    import gc

    import numpy as np
    import torch

    from tatva_amd import solvers

    n, steps_per_pass, stretch = args.n, 5, 0.3
    t0 = time.perf_counter()
    P = stretch_problem(n, kind="hex8")
    L0 = P["L"].at["load"].set(P["values"](0.0))
    mesh, op = P["mesh"], P["op"]
    gb0 = device_gb_in_use()
    mixed = solvers.Newton(P["E"], L0, mode="mixed", tol=1e-8, inner_tol=1e-3, chunk=64)
    runner = LoadStepper(P, mixed, steps_per_pass, stretch)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    n_free = mixed.red.size
    print(f"[headline] {mesh.n_elements} elements, n_free {n_free}, setup {setup_s:.1f} s",
          file=sys.stderr, flush=True)
    for w in range(args.warmup):
        st = runner.step()
        print(f"[headline] warm-up {w + 1}/{args.warmup}: load step {runner.k}, "
              f"{st.newton_iterations} Newton, {st.cg_iterations} CG, {st.seconds:.2f} s",
              file=sys.stderr, flush=True)
    solvers.set_profiling(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        stats.append(runner.step())
        print(f"[headline] step {i + 1}/{args.steps}: load step {runner.k}, "
              f"{stats[-1].newton_iterations} Newton, {stats[-1].cg_iterations} CG, "
              f"{stats[-1].seconds:.2f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    solvers.set_profiling(False)
    gb_mixed = device_gb_in_use() - gb0
    cg_total = sum(s.cg_iterations for s in stats)
    value = n_free * cg_total / elapsed
    op_n = sum(s.op_samples for s in stats)
    tile_us = sum(s.op_main_ms for s in stats) / op_n * 1e3 if op_n else None
    op_us = sum(s.op_ms for s in stats) / op_n * 1e3 if op_n else None
    u_mixed = runner.u.clone()
    k_mixed = runner.k
    mixed_full = None
    if args.steps < steps_per_pass or args.steps % steps_per_pass:
        # the comparison needs one whole pass over the load steps: run one untimed
        runner.k = steps_per_pass
        t2 = time.perf_counter()
        full = [runner.step() for _ in range(steps_per_pass)]
        torch.cuda.synchronize()
        mixed_full = (time.perf_counter() - t2, full)
        u_mixed, k_mixed = runner.u.clone(), runner.k
    else:
        passes = args.steps // steps_per_pass
        mixed_full = (elapsed / passes, stats[-steps_per_pass:])
    res = {"mixed": {
        "seconds_per_pass": mixed_full[0],
        "newton_iterations": sum(s.newton_iterations for s in mixed_full[1]),
        "cg_iterations": sum(s.cg_iterations for s in mixed_full[1]),
        "converged": all(s.converged for s in mixed_full[1]),
        "worst_rel_residual": max(s.residual_history[-1] / s.residual_history[0]
                                  for s in mixed_full[1]),
        "device_gb_in_use": gb_mixed,
    }}
    res["mixed"]["dof_iters_per_s"] = n_free * res["mixed"]["cg_iterations"] / mixed_full[0]
    del runner, mixed
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if not args.no_fp64:
        gb1 = device_gb_in_use()
        t3 = time.perf_counter()
        csr = solvers.Newton(P["E"], L0, mode="csr", tol=1e-8, lin_tol=1e-10, chunk=64)
        torch.cuda.synchronize()
        csr_setup = time.perf_counter() - t3
        print(f"[fp64] colored CSR set up in {csr_setup:.1f} s (nnz {csr.colored_matrix.nnz}, "
              f"{csr.colored_matrix.n_colors} colors)", file=sys.stderr, flush=True)
        P["progress"] = True
        solvers.set_profiling(True)
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        u_csr, cstats = solve_load_steps(P, csr, steps_per_pass, stretch, upto=k_mixed)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t4
        solvers.set_profiling(False)
        P["progress"] = False
        cg = sum(s.cg_iterations for s in cstats)
        on = sum(s.op_samples for s in cstats)
        blocked = csr.colored_matrix.block_view(3) is not None
        res["fp64_csr"] = {
            "seconds_per_pass": dt, "setup_s": csr_setup,
            "newton_iterations": sum(s.newton_iterations for s in cstats), "cg_iterations": cg,
            "dof_iters_per_s": n_free * cg / dt,
            "converged": all(s.converged for s in cstats),
            "worst_rel_residual": max(s.residual_history[-1] / s.residual_history[0]
                                      for s in cstats),
            "spmv_us": sum(s.op_main_ms for s in cstats) / on * 1e3 if on else None,
            "spmv_layout": "node-block CSR" if blocked else "sliced ELL copy of the CSR",
            "nnz": csr.colored_matrix.nnz, "n_colors": csr.colored_matrix.n_colors,
            "device_gb_in_use": device_gb_in_use() - gb1,
        }
        res["dof_iters_speedup"] = res["mixed"]["dof_iters_per_s"] / res["fp64_csr"]["dof_iters_per_s"]
        res["time_to_solution_speedup"] = dt / res["mixed"]["seconds_per_pass"]
        res["mixed_vs_fp64_rel_error"] = float((u_mixed - u_csr).abs().max() / u_csr.abs().max())
        del csr
        gc.collect()
        torch.cuda.empty_cache()
    else:
        u_csr = None
    if not args.no_oracle:
        from oracle import fem  # the checker only

        L_end = P["L"].at["load"].set(P["values"](stretch * k_mixed / steps_per_pass))
        free = L_end.free_dofs
        u_load = L_end.lift_from_zeros(np.zeros(free.size)).cpu().numpy()
        vecs = [u_load, u_mixed.cpu().numpy()] + ([u_csr.cpu().numpy()] if u_csr is not None else [])
        t5 = time.perf_counter()
        # the oracle's element algebra in C over the host's threads (pinned to the
        # NumPy oracle to 1e-12 in the oracle's tests)
        Q = fem.ProblemOMP(fem.Hex8(), mesh.coords, mesh.elements, fem.NeoHookean(1.0, 10.0),
                           threads=cpu_threads())
        rs = [Q.residual(v) for v in vecs]
        r_load = float(np.linalg.norm(rs[0][free]))
        res["oracle_load_residual_norm"] = r_load
        res["oracle_seconds"] = time.perf_counter() - t5
        for key, r in zip(("mixed", "fp64_csr"), rs[1:]):
            res[key]["oracle_rel_residual"] = float(np.linalg.norm(r[free]) / r_load)
    # with the cache the PCG reads the linearisation (F^-1 and mu - lambda ln J: 10
    # floats per quadrature point) instead of the state
    cached = args.lin_cache != "never"     # the headline mesh is Hex8
    geo = not args.no_geo_cache and mesh_is_affine(op)
    defer = not args.no_defer
    full_vec = not args.no_full_vectors
    n_vec = mesh.n_nodes * 3 if full_vec else n_free   # length of the CG vectors
    tile_bytes, lay = hvp_tile_bytes(op, cached_lin=10 if cached else 0, geo=geo,
                                     defer_n=n_vec if defer else 0, mapped=not full_vec)
    hbm_rate = tile_bytes / (tile_us * 1e-6) / 1e9 if tile_us else None
    # the profile's counters are those of the default kernel on this workload
    default_kernel = n == 160 and not cached and geo and defer and full_vec
    prof = profile_counters(HEADLINE_PROFILE) if default_kernel else {}
    valu = prof.get("valu_wave_instructions_per_launch")
    traffic = prof.get("hbm_bytes_per_launch")
    achieved = valu / (tile_us * 1e-6) / 1e9 if (valu and tile_us) else None
    out = {
        "metric": "Newton DOF-iters/s + HBM GB/s (mixed vs FP64 @1e-8 rel); serve p50/p99 ms",
        "value": value,
        "unit": "DOF-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / max(args.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 HVP/CG + f64 residual/Newton",
        "data": "synthetic (structured mesh, prescribed boundary displacements)",
        "config": {
            "workload": f"hex8 Neo-Hookean {n}^3, mu=1, lambda=10, x=0 clamped, u_x on x=1 in "
                        f"{steps_per_pass} load steps to {stretch}; one step = one load step, "
                        "mixed Newton (FP32 matrix-free Jacobi-PCG to 1e-3) to 1e-8 relative",
            "elements": mesh.n_elements, "nodes": mesh.n_nodes, "n_free": n_free,
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": (f"k_tile_assemble<HEX8, NeoHookean<3,float>, float, "
                       f"{'kModeHvpLin' if cached else 'kModeHvp'}, "
                       f"{'kGeomCache' if geo else 'kGeomAffine'}> (FP32 HVP tile kernel of the "
                       "matrix-free PCG operator)"),
            "bound": "valu",
            "achieved": achieved,
            "peak": VALU_PEAK_GINST,
            "unit": "G wave64 VALU instructions/s",
            "frac": (achieved / VALU_PEAK_GINST) if achieved else None,
            "traffic": traffic,
            "traffic_source": os.path.relpath(HEADLINE_PROFILE, ROOT) if traffic else None,
            "valu_wave_instructions_per_launch": valu,
            "algorithmic_bytes": tile_bytes,
            "hbm_achieved_gbs": hbm_rate,
            "hbm_frac": hbm_rate / HBM_PEAK_GBS if hbm_rate else None,
            "geometry_cache": geo,
            "deferred_cg_update": defer,
            "full_length_cg_vectors": full_vec,
            "avg_launch_us": tile_us,
            "operator_us": op_us,
            "samples": op_n,
            "layout": lay,
            "note": "VALU roofline (neither HBM nor MFMA bounds this kernel): achieved = the "
                    "kernel's VALU wave-instructions per launch (rocprofv3 SQ_INSTS_VALU of this "
                    "workload, in the profile named in traffic_source) / the launch time measured "
                    "here with HIP events; peak = one wave64 instruction per SIMD per 2 cycles x "
                    "1,024 SIMDs x 2.4 GHz (the rate FP32 chains with vector operands measure; an "
                    "instruction with a scalar operand issues at half of it).  "
                    "hbm_achieved_gbs = algorithmic bytes / the same time",
        },
        "cpu_baseline": None,
        "newton_iterations": sum(s.newton_iterations for s in stats),
        "cg_iterations": cg_total,
        "converged": all(s.converged for s in stats),
        "setup_s": setup_s,
        **res,
    }
    return out


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on,
    capped by OMP_NUM_THREADS when the host sets it (the GPU host's share)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.strip().isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def cpu_baseline_headline(args):
    """The oracle's FP64 element kernels in C over OpenMP threads
    (oracle/fem_omp.c), driving the same inexact-Newton step on the headline
    mesh: residual assembly, Jacobi diagonal, then matrix-free Jacobi-PCG
    iterations (one HVP each; NumPy vector updates) for a bounded time.  The
    state is the first load step's start (u_x = 0.06 x, constraints set)."""
    import numpy as np

    from oracle import fem  # the CPU baseline only

    n = args.n
    threads = cpu_threads()
    X, el = fem.cube_hex8(n)
    x0 = np.where(np.isclose(X[:, 0], 0.0))[0]
    x1 = np.where(np.isclose(X[:, 0], 1.0))[0]
    fixed = np.concatenate([(x0[:, None] * 3 + np.arange(3)).ravel(), x1 * 3])
    size = X.shape[0] * 3
    free = np.setdiff1d(np.arange(size), fixed)
    u = np.zeros(size)
    u[0::3] = 0.06 * X[:, 0]
    u[(x0[:, None] * 3 + np.arange(3)).ravel()] = 0.0
    Q = fem.ProblemOMP(fem.Hex8(), X, el, fem.NeoHookean(1.0, 10.0), threads=threads)
    try:  # BLAS threads of the NumPy dot products would compete with the OpenMP ones
        from threadpoolctl import threadpool_limits
        limit = threadpool_limits(1, user_api="blas")
    except ImportError:
        limit = None
    t0 = time.perf_counter()
    b = -Q.residual(u)[free]
    dinv = 1.0 / Q.hessian_diag(u)[free]
    t_asm = time.perf_counter() - t0
    x = np.zeros(free.size)
    r = b.copy()
    z = dinv * r
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    full = np.zeros(size)
    its = 0
    while time.perf_counter() - t0 < args.cpu_seconds and np.linalg.norm(r) > 1e-3 * bn:
        full[free] = p
        Ap = Q.hvp(u, full)[free]
        alpha = rz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        z = dinv * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        its += 1
    dt = time.perf_counter() - t0
    if limit is not None:
        limit.restore_original_limits()
    return {"value": free.size * its / dt, "unit": "DOF-iters/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle FP64 Hex8 Neo-Hookean {n}^3 kernels in C (oracle/fem_omp.c, "
                      f"OpenMP, {threads} threads): residual + Jacobi diagonal assembly "
                      f"({t_asm:.1f} s) then {its} matrix-free Jacobi-PCG iterations of the first "
                      f"Newton step on {free.size} free dofs, {dt:.1f} s in total (assembly "
                      "included)"}


def c2_extra():
    """C2 (Hex8 linear elasticity 64^3, traction 1e-3 on x = 1): FP64 colored-CSR
    Newton vs mixed Newton on the same problem, with the oracle's FP64 residual
    of both solutions relative to the load's."""
    import numpy as np
    import torch

    from oracle import fem  # the checker only

    P = build_problem(64)
    nd = P["mesh"].n_nodes * 3

    def run_c2(sv):
        uu = torch.zeros(nd, dtype=torch.float64, device="cuda")
        return uu, [sv.solve(uu)]

    r = mixed_vs_fp64(P["E"], P["L"], run_c2, colored_matrix=P["cm"], keep_solutions=True)
    mesh, dn = P["mesh"], P["dens"]
    prob = fem.Problem(fem.Hex8(), mesh.coords, mesh.elements,
                       fem.LinearElasticity(dn.mu, dn.lmbda), f_ext=P["f"])
    free = P["L"].free_dofs
    r0 = np.linalg.norm(prob.residual(np.zeros(nd))[free])
    for mode in ("mixed", "csr"):
        uh = r.pop(f"_u_{mode}").cpu().numpy()
        r[mode]["oracle_rel_residual"] = float(np.linalg.norm(prob.residual(uh)[free]) / r0)
    r["workload"] = ("hex8 linear elasticity 64^3, E=1, nu=0.3, x=0 clamped, traction 1e-3 "
                     "on x=1; FP64 CSR PCG to 1e-10 vs mixed Newton to 1e-8")
    return r


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=None,
                    help="cells per side of the Hex8 cube (headline 160, c2 64)")
    ap.add_argument("--no-fp64", action="store_true",
                    help="headline: skip the FP64 colored-CSR comparison")
    ap.add_argument("--no-oracle", action="store_true",
                    help="headline: skip the oracle residuals of the final states")
    ap.add_argument("--lin-cache", choices=("auto", "always", "never"), default="never",
                    help="matrix-free PCG: re-linearise the density in every HVP (never: the "
                         "default) or read it from a cache written once per solve (auto: on "
                         "Hex8 meshes; always)")
    ap.add_argument("--no-geo-cache", action="store_true",
                    help="matrix-free PCG on affine Hex8 meshes: compute each element's J^-1 and "
                         "det J from the coordinates in every product instead of reading them "
                         "from the cache written once per solve (the default)")
    ap.add_argument("--no-defer", action="store_true",
                    help="matrix-free PCG: apply each iteration's p and x updates in its CG step "
                         "instead of in the next product's tile kernel (the default)")
    ap.add_argument("--no-full-vectors", action="store_true",
                    help="matrix-free PCG: run the CG on the reduced vectors through the dof map "
                         "instead of on full-length ones (the default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the mixed-vs-FP64 and C1 latency sections")
    ap.add_argument("--nh-n", type=int, default=48, help="cells per side, Neo-Hookean comparison")
    ap.add_argument("--nh-steps", type=int, default=2, help="load steps, Neo-Hookean comparison")
    ap.add_argument("--nh-stretch", type=float, default=0.1, help="final u_x on x = 1")
    ap.add_argument("--workload", choices=("headline", "c2", "c2mixed", "c3", "c4", "c5", "nh48",
                                           "copy", "valu"),
                    default="headline",
                    help="headline: the default line (Hex8 Neo-Hookean 160^3 mixed Newton, "
                         "FP64 path alongside); c2: the FP64 Hex8 elasticity 64^3 line; "
                         "c2mixed / nh48: only the mixed-vs-FP64 "
                         "comparison on the 64^3 elasticity / 48^3 Neo-Hookean cube; "
                         "c3/c5: large mixed-vs-FP64 Neo-Hookean runs; c4: the serving "
                         "stream only (--serve-requests, --serve-qps)")
    ap.add_argument("--large-steps", type=int, default=None,
                    help="c3/c5: run only the first N load steps (short profiled runs)")
    ap.add_argument("--serve-requests", type=int, default=500)
    ap.add_argument("--serve-qps", type=float, nargs="+", default=[40.0, 160.0, 640.0])
    ap.add_argument("--serve-budget-gb", type=float, default=0.4,
                    help="HBM budget of the extra serving run at the highest QPS (0: skip)")
    ap.add_argument("--serve-chunk", type=int, default=8, help="batched PCG iterations per poll")
    ap.add_argument("--serve-prefetch", type=int, default=2,
                    help="serving: threads preparing arrived requests")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the path's "
              "multi-GPU sharding is a design note, not a measured configuration.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from tatva_amd import solvers

    torch.cuda.set_device(0)
    solvers.set_linearization_cache(args.lin_cache)
    solvers.set_geometry_cache(not args.no_geo_cache)
    solvers.set_deferred_update(not args.no_defer)
    solvers.set_full_vectors(not args.no_full_vectors)
    if args.workload == "copy":
        # calibration of the HBM traffic counters: a plain 512 MiB device copy
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"workload": "copy", "bytes_read": src.numel() * 8,
                          "bytes_written": dst.numel() * 8, "launches": 10}), flush=True)
        return 0
    if args.workload == "valu":
        # calibration of the VALU issue rate: FP32 chains of one instruction form
        # (0 v_fma_f32 s, s; 1 v_fmac_f32; 2 v_add_f32; 3 v_mul_f32; 4 v_fma_f32 v, v;
        # 5 v_fmac_f32 s), 8 waves per SIMD
        from tatva_amd import _native
        blocks, iters, launches = 2048, 4096, 10
        out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
        st = torch.cuda.current_stream()
        for op, form in enumerate(("v_fma_f32 (scalar a, b)", "v_fmac_f32", "v_add_f32",
                                   "v_mul_f32", "v_fma_f32 (vector a, b)",
                                   "v_fmac_f32 (scalar a)")):
            run = lambda: _native.check(_native.lib().tatva_valu_calibration(
                op, blocks, iters, out.data_ptr(), st.cuda_stream))
            run()
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record(st)
            for _ in range(launches):
                run()
            t1.record(st)
            torch.cuda.synchronize()
            if not bool(torch.isfinite(out).all()):
                raise RuntimeError(f"calibration {form} produced non-finite values")
            us = t0.elapsed_time(t1) * 1e3 / launches
            n = blocks * 4 * 32 * iters
            print(json.dumps({"workload": "valu", "form": form, "op": op, "blocks": blocks,
                              "iters": iters, "launches": launches + 1,
                              "loop_wave_instructions": n, "avg_launch_us": us,
                              "g_wave_instr_per_s": n / us / 1e3}), flush=True)
        return 0
    if args.workload == "headline":
        if args.n is None:
            args.n = 160
        out = headline(args)
        if not args.no_extras:
            out["c2_mixed_vs_fp64"] = c2_extra()
            out["c1_poisson_tri3_64"] = c1_latency()
            if args.serve_requests > 0:
                out["serve_c4"] = serve_sweep(args)
        if not args.no_cpu_baseline:
            out["cpu_baseline"] = cpu_baseline_headline(args)
        print(json.dumps(out), flush=True)
        return 0
    if args.n is None:
        args.n = 64
    if args.workload in ("c3", "c5"):
        r = large_mixed(args.workload, oracle_check=not args.no_oracle,
                        upto=args.large_steps,
                        modes=("mixed",) if args.no_fp64 else ("csr", "mixed"))
        r["lin_cache"] = args.lin_cache
        r["geo_cache"] = not args.no_geo_cache
        r["deferred_cg_update"] = not args.no_defer
        r["full_length_cg_vectors"] = not args.no_full_vectors
        print(json.dumps({"workload": args.workload, **r}), flush=True)
        return 0
    if args.workload == "c4":
        print(json.dumps({"workload": args.workload, "serve": serve_sweep(args)}), flush=True)
        return 0
    if args.workload == "c2mixed":
        P = build_problem(args.n)
        nd = P["mesh"].n_nodes * 3

        def run_c2(sv):
            uu = torch.zeros(nd, dtype=torch.float64, device="cuda")
            return uu, [sv.solve(uu)]

        r = mixed_vs_fp64(P["E"], P["L"], run_c2, colored_matrix=P["cm"])
        print(json.dumps({"workload": args.workload, **r}), flush=True)
        return 0
    if args.workload == "nh48":
        NH = stretch_problem(args.nh_n)
        L0 = NH["L"].at["load"].set(NH["values"](0.0))
        r = mixed_vs_fp64(NH["E"], L0,
                          lambda sv: solve_load_steps(NH, sv, args.nh_steps, args.nh_stretch))
        print(json.dumps({"workload": args.workload, **r}), flush=True)
        return 0
    t_setup = time.perf_counter()
    P = build_problem(args.n)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    cm, solver, L = P["cm"], P["solver"], P["L"]
    n_free = cm.shape[0]
    nnz = cm.nnz
    ndof = P["mesh"].n_nodes * 3

    solvers.set_profiling(True)

    def step():
        u = torch.zeros(ndof, dtype=torch.float64, device="cuda")
        st = solver.solve(u)
        return u, st

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    u = None
    for _ in range(args.steps):
        u, st = step()
        stats.append(st)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    solvers.set_profiling(False)

    cg_total = sum(s.cg_iterations for s in stats)
    newton_total = sum(s.newton_iterations for s in stats)
    value = n_free * cg_total / elapsed
    op_ms = sum(s.op_ms for s in stats)
    op_samples = sum(s.op_samples for s in stats)
    spmv_s = (op_ms / op_samples) * 1e-3 if op_samples else float("nan")
    blocked = cm.block_view(3) is not None
    if blocked:
        # node-block view: 8 B per value, one int32 per 3x3 block, two indptr
        # reads per node row, p read once and Ap written once
        kernel = "k_pcg_bspmv<3> (node-block CSR SpMV; p.Ap shares by k_dot_shares)"
        spmv_bytes = 8 * nnz + 4 * (nnz // 9) + 8 * (n_free // 3) + 16 * n_free
    else:
        kernel = "k_sell_spmv (sliced ELL copy of the CSR, fused with p.Ap)"
        spmv_bytes = 12 * nnz + 8 * (n_free // 64) + 16 * n_free
    achieved = spmv_bytes / spmv_s / 1e9 if op_samples else None
    spmv_traffic = None
    if blocked and args.n == 64:
        spmv_traffic = profile_counters(C2_PROFILE).get("hbm_bytes_per_launch")
    converged = all(s.converged for s in stats)

    extras = {}
    if not args.no_extras:
        # mixed vs FP64 on the same C2 problem (linear: Newton = iterative refinement)
        def run_c2(sv):
            uu = torch.zeros(ndof, dtype=torch.float64, device="cuda")
            return uu, [sv.solve(uu)]

        extras["mixed_vs_fp64_c2"] = mixed_vs_fp64(P["E"], L, run_c2, colored_matrix=cm)
        # mixed vs FP64 on a load-stepped Neo-Hookean Hex8 cube (C5 form, smaller)
        NH = stretch_problem(args.nh_n)
        L0 = NH["L"].at["load"].set(NH["values"](0.0))
        extras["mixed_vs_fp64_neo_hookean"] = mixed_vs_fp64(
            NH["E"], L0, lambda sv: solve_load_steps(NH, sv, args.nh_steps, args.nh_stretch))
        extras["mixed_vs_fp64_neo_hookean"]["workload"] = (
            f"hex8 Neo-Hookean {args.nh_n}^3, mu=1, lambda=10, {args.nh_steps} load steps to "
            f"u_x={args.nh_stretch} on x=1, Newton to 1e-8 relative per step")
        del NH
        extras["c1_poisson_tri3_64"] = c1_latency()
        if args.serve_requests > 0:
            extras["serve_c4"] = serve_sweep(args)

    # checks on the final state: true linear residual and the FP64 oracle's residual
    res_free = P["E"].reduced(L).residual_full_state(u)
    rel_res = solvers.norm(res_free) / max(stats[-1].residual_history[0], 1e-300)
    oracle_rel = None
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import fem  # the checker / CPU baseline only

        mesh = P["mesh"]
        dn = P["dens"]
        prob = fem.Problem(fem.Hex8(), mesh.coords, mesh.elements,
                           fem.LinearElasticity(dn.mu, dn.lmbda), f_ext=P["f"])
        uh = u.cpu().numpy()
        r_or = prob.residual(uh)[L.free_dofs]
        r0_or = np.linalg.norm(prob.residual(np.zeros_like(uh))[L.free_dofs])
        oracle_rel = float(np.linalg.norm(r_or) / r0_or)
        # CPU baseline: oracle PCG on the same reduced stiffness (copied to the host)
        import tatva_amd.sparse as sp
        K = sp.jacfwd(P["E"].reduced(L).grad, cm)(np.zeros(n_free))
        Kh = K.to_csr()
        b = -prob.residual(np.zeros(ndof))[L.free_dofs]
        del prob
        threads = len(os.sched_getaffinity(0))
        v, its, secs = cpu_baseline(Kh, b, seconds=args.cpu_seconds)
        cpu = {"value": v, "unit": "DOF-iters/s", "cores": 1, "kind": "port",
               "sample": f"oracle Jacobi-PCG (SciPy CSR SpMV + NumPy vectors, 1 thread of "
                         f"{threads} available) on the same {n_free}-dof reduced stiffness "
                         f"({nnz} nnz), {its} iterations in {secs:.1f} s; assembly not included"}

    out = {
        "metric": "Newton DOF-iters/s + HBM GB/s (mixed vs FP64 @1e-8 rel); serve p50/p99 ms",
        "value": value,
        "unit": "DOF-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (structured mesh, closed-form loads)",
        "config": {
            "workload": f"hex8 linear elasticity {args.n}^3, FP64 colored-CSR Newton + Jacobi-PCG 1e-10",
            "elements": P["mesh"].n_elements, "nodes": P["mesh"].n_nodes, "n_free": n_free,
            "nnz": nnz, "n_colors": cm.n_colors, "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": kernel,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": spmv_traffic,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes on this line, "
                              "profiles/k_pcg_bspmv_hex8_le_64cubed.md" if spmv_traffic else None,
            "algorithmic_bytes": spmv_bytes,
            "avg_launch_us": spmv_s * 1e6 if op_samples else None,
            "samples": op_samples,
            "note": "achieved = algorithmic bytes / timed launch; the PCG streams the matrix "
                    "in alternating directions between iterations, so part of those bytes "
                    "come from the 256 MB last-level cache rather than HBM (the PMC traffic "
                    "counts every L2 miss)",
        },
        "cpu_baseline": cpu,
        "newton_iterations": newton_total,
        "cg_iterations": cg_total,
        "cg_iterations_per_step": cg_total / max(args.steps, 1),
        "converged": converged,
        "final_rel_residual": rel_res,
        "oracle_rel_residual": oracle_rel,
        "assembly_ms_per_step": sum(s.assembly_seconds for s in stats) / args.steps * 1e3,
        "pcg_ms_per_step": sum(s.cg_seconds for s in stats) / args.steps * 1e3,
        "setup_s": setup_s,
        "pattern_s": P["pattern_s"],
        "coloring_s": P["coloring_s"],
        **extras,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
