"""Benchmark of the implicit field-solve hot path on one MI355X.

Workload: the variable-coefficient Poisson problem with space charge (eps_r = 4
slab under z = 0.5 cm, eight seeded Gaussian charge blobs, 0 V / 1 kV
Dirichlet plates) on a 257^3 cell-centred grid over a 2 cm cube, solved from a
zero initial guess to a relative residual ||b - A u||_2/||b||_2 <= 1e-10 by
V(2,2) red-black Gauss-Seidel multigrid on the 257 -> 129 -> ... -> 3
hierarchy.

One step = one complete solve in the headline form: FP32 smoothing,
restriction and prolongation under an FP64 defect-correction loop (two FP32
V-cycles per FP64 residual, after a full-multigrid first step). The same
solve in plain FP64 and
the adjoint (transposed-operator) solve are timed beside it. All fields are
resident in HBM before the timed region.

Prints one JSON line (see the keys below). Also times the reference CPU path
(the reference Fortran/MPI library built from the reference sources into
oracle/_ref/, falling back to the oracle's C restatement) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK_HBM_GBPS = 8000.0
# algorithmic HBM bytes per fine cell per V-cycle (Poisson, c = 1 coefficient
# field), summed over the level hierarchy
BYTES_PER_CELL_CYCLE = {"fp64": 302.9, "fp32dc": 199.4}


def log(msg):
    """A progress line on stderr (stdout carries only the JSON line)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def plain_start_ms(case, hier, prec, steps, tol):
    """ms per solve with V-cycles from the zero guess throughout (no
    full-multigrid first pass), for comparison."""
    import torch

    from mps3d_amd import FieldSolver, to_device
    s = FieldSolver(case.n, case.length, hierarchy=hier, precision=prec, max_cycles=80, tol=tol,
                    tol_kind="rel", arith="fast", start="guess")
    s.set_case(case)
    u = to_device(np.zeros(s.shape))
    u.zero_()
    s.solve(u)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        u.zero_()
        k, r, _, st = s.solve(u, raise_on_failure=True)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    vc = s.last_vcycles()
    s.close()
    del u
    torch.cuda.empty_cache()
    return {"ms_per_solve": round(ms, 3), "cycles": int(k), "vcycles": vc}


def gpu_solves(n, steps, warmup, length=0.02, tol=1e-10, start="fmg"):
    import torch

    from mps3d_amd import FieldSolver, from_device, to_device
    from mps3d_amd import workloads as W

    case = W.varcoef_poisson(n, length=length)
    hier = "aligned" if n % 2 == 1 else "ref"
    out = {}
    sols = {}
    for prec in ("fp32dc", "fp64"):
        st_ = start if hier == "aligned" else "guess"
        s = FieldSolver(case.n, case.length, hierarchy=hier, precision=prec, max_cycles=80, tol=tol,
                        tol_kind="rel", arith="fast", start=st_)
        s.set_case(case)
        u = to_device(np.zeros(s.shape))
        for _ in range(warmup):
            u.zero_()
            s.solve(u)
        torch.cuda.synchronize()
        times, cycles, res, vcyc = [], [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            u.zero_()
            k, r, _, st = s.solve(u, raise_on_failure=True)
            ms, _ = s.last_timing()
            times.append(ms)
            cycles.append(k)
            res.append(r)
            vcyc.append(s.last_vcycles())
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / steps
        ent = {"ms_per_solve": wall, "event_ms": float(np.mean(times)), "cycles": int(np.max(cycles)),
               "relres": float(np.max(res))}
        # V-cycles as the library counts them (FP32 + DC: the FP32 cycles of
        # all defect steps; the full-multigrid first step counts as one)
        ent["vcycles"] = int(np.max(vcyc))
        ent["unknowns_per_s"] = float(np.prod(case.n)) * ent["vcycles"] / (wall * 1e-3)
        ent["solve_GBps"] = BYTES_PER_CELL_CYCLE[prec] * float(np.prod(case.n)) * ent["vcycles"] / (wall * 1e-3) / 1e9
        # hot kernels on the finest level: every launch of one solve like the
        # timed ones bracketed by HIP events on the library's stream
        # (avg_us, GBps), and the isolated hook with the Infinity Cache
        # evicted before each launch (hook_us)
        kern = {}
        for kid, label in ((0, "colour"), (1, "residual"), (2, "restrict"), (3, "prolong")):
            fp32 = prec == "fp32dc" and kid != 1
            ms_h, by, name = s.time_kernel(kid, fp32=fp32, reps=20)
            u.zero_()
            ms, nl, _ = s.time_kernel_in_solve(u, kid, fp32=fp32)
            if kid == 2 and hier == "aligned" and prec == "fp32dc":
                # the finest FP32 restriction is fused with its residual:
                # (c + 2) w + w/8 bytes per fine cell
                name, by = "k_resrestrict_planes", (1.0 + 2.0 + 0.125) * 4.0 * float(np.prod(case.n))
            kern[label] = {"name": name, "fp32": fp32, "avg_us": ms * 1e3, "launches": nl, "alg_bytes": by,
                           "GBps": by / (ms * 1e-3) / 1e9 if ms > 0 else None, "hook_us": ms_h * 1e3}
        ent["kernels"] = kern
        if prec == "fp64":
            # adjoint: A^T lambda = g with the transposed operator, same tolerance
            g = np.random.default_rng(7).standard_normal(case.n)
            lam = to_device(np.zeros(s.shape))
            gd = torch.from_numpy(np.ascontiguousarray(g.transpose(2, 1, 0))).cuda()
            s.solve_transpose(gd, lam)  # warm (builds A^T)
            ta = []
            for _ in range(max(1, steps // 2)):
                lam.zero_()
                torch.cuda.synchronize()
                t = time.perf_counter()
                ka, ra, _ = s.solve_transpose(gd, lam)
                torch.cuda.synchronize()
                ta.append((time.perf_counter() - t) * 1e3)
            ent["adjoint_ms"] = float(np.mean(ta))
            ent["adjoint_cycles"] = int(ka)
        sols[prec] = from_device(u)[1:-1, 1:-1, 1:-1]
        out[prec] = ent
        s.close()
        del u
        torch.cuda.empty_cache()
        if st_ != "guess":
            ent["plain_start"] = plain_start_ms(case, hier, prec, steps, tol)
    err = float(np.linalg.norm(sols["fp32dc"] - sols["fp64"]) / np.linalg.norm(sols["fp64"]))
    # each solution's error: against the same problem solved in FP64 to a
    # relative residual of 1e-13 (untimed)
    s = FieldSolver(case.n, case.length, hierarchy=hier, precision="fp64", max_cycles=80, tol=1e-13,
                    tol_kind="rel", arith="fast", start=start if hier == "aligned" else "guess")
    s.set_case(case)
    u = to_device(np.zeros(s.shape))
    s.solve(u)
    tight = from_device(u)[1:-1, 1:-1, 1:-1]
    s.close()
    del u
    torch.cuda.empty_cache()
    for prec in sols:
        out[prec]["err_vs_converged"] = float(np.linalg.norm(sols[prec] - tight) / np.linalg.norm(tight))
    return case, hier, out, err


def calibration_copy():
    """One device copy of a known size (hipMemcpyAsync, the
    __amd_rocclr_copyBuffer kernel): rocprofv3's FETCH_SIZE / WRITE_SIZE of
    this dispatch scale the counters of the solver's kernels to bytes."""
    import torch

    from mps3d_amd.profiling import COPY_BYTES
    src = torch.ones(COPY_BYTES, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    # the same copy timed a few more times: read + write bytes per second of
    # a plain device copy, the practical ceiling of a streaming kernel here
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbps = 2.0 * COPY_BYTES * 5 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbps


def step_kernels(n, reps=20):
    """The per-step work either side of the solves, on device arrays of the
    same grid: update_efield (mps_efield) and the transport assembly of the
    ion_elec chemistry (electron and energy *_update_transport, Poisson's,
    update_electrontemp). Mean wall time per call over `reps` calls, each
    followed by a device synchronisation (calls on device arrays return with
    their kernels queued), so this is an upper bound on the kernel time, and
    the algorithmic GB/s it implies: every array the kernel reads or writes
    once, ghosted. The fastest single call is given beside the mean: a
    call's wall time also holds the host's wake-up from the synchronisation,
    which now and then adds milliseconds."""
    import torch

    from mps3d_amd import FieldSolver
    from mps3d_amd.chemistry import etempmin_default, ion_elec

    s = FieldSolver(n, 0.02)
    s.set_bc(("ZGRD", "ZGRD", "ZGRD", "ZGRD", "DIRC", "DIRC"), (0.0, 0.0, 0.0, 0.0, 0.0, 1000.0))
    ch = ion_elec()
    s.set_chemistry(ch)
    ng = int(np.prod(s.shape))
    g = torch.Generator(device="cuda").manual_seed(20261202)

    def field(m, lo, hi):
        return torch.rand(m * ng, dtype=torch.float64, device="cuda", generator=g) * (hi - lo) + lo

    phi = field(1, 0.0, 1000.0)
    E = torch.empty(3 * ng, dtype=torch.float64, device="cuda")
    te = field(1, 5e3, 6e4)
    nd = torch.cat([field(1, 1e21, 3e22), field(1, 1e14, 1e16), field(1, 1e14, 1e16)])
    soln = field(1, 0.0, 1.0)
    eps = field(1, 1e-3, 1.0)
    out = {k: torch.empty(m * ng, dtype=torch.float64, device="cuda")
           for k, m in (("vel", 3), ("dcoeff", 1), ("reac", 1), ("source", 1), ("prod", 1))}
    sc = (1.0, 1.0, 1.0, 1.0)
    calls = {
        # name: (call, algorithmic bytes per ghosted cell)
        "efield": (lambda: s.efield(phi, 50.0, False, out=E), 8 * (1 + 3)),
        "species_transport": (lambda: s.species_transport(2, E, te, nd, sc, soln=soln, dt=1e-3, out=out),
                              8 * (3 + 1 + 2 + 1 + 3 + 1 + 1 + 1 + 1)),
        "eenergy_transport": (lambda: s.eenergy_transport(E, te, nd, sc, lscale=1.0, soln=soln, dt=1e-3, out=out),
                              8 * (3 + 1 + 2 + 1 + 3 + 1 + 1 + 1)),
        "poisson_transport": (lambda: s.poisson_transport(nd[ng:2 * ng], nd[2 * ng:], sc, soln=soln, dt=1e-3,
                                                          out=out), 8 * (1 + 1 + 1 + 3 + 1 + 1 + 1)),
        "electron_temp": (lambda: s.electron_temp(eps, nd[ng:2 * ng], 1.0, etempmin_default(), te), 8 * 3),
    }
    res = {}
    # writeoutput of two cell scalars from device tensors (the copy to the
    # host included), one piece, into a scratch directory
    import shutil
    import tempfile

    from mps3d_amd import vtk
    td = tempfile.mkdtemp()
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vtk.write_output(os.path.join(td, "out"), [phi, te], ("Potential (V)", "Electron_temp (K)"), (n, n, n),
                         (0.02,) * 3, (1, 1, 1))
        sec = time.perf_counter() - t0
        nbytes = sum(os.path.getsize(os.path.join(r, f)) for r, _, fs in os.walk(td) for f in fs)
        res["vtk_output"] = {"s_per_call": round(sec, 3), "bytes": nbytes, "MBps": round(nbytes / sec / 1e6, 1)}
    finally:
        shutil.rmtree(td, ignore_errors=True)
    for name, (fn, bpc) in calls.items():
        fn()
        torch.cuda.synchronize()
        per = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            per.append(time.perf_counter() - t0)
        us = sum(per) * 1e6 / reps
        us_min = min(per) * 1e6
        res[name] = {"us_per_call": round(us, 1), "us_min": round(us_min, 1), "alg_bytes": bpc * ng,
                     "alg_GBps": round(bpc * ng / (us * 1e-6) / 1e9, 1),
                     "alg_GBps_best": round(bpc * ng / (us_min * 1e-6) / 1e9, 1)}
    s.close()
    del phi, E, te, nd, soln, eps, out
    torch.cuda.empty_cache()
    # the whole time step (solve_manager's loop: four transport assemblies,
    # four V-cycle solves to 1e-11, the field and the temperature) on the
    # reference's ion_elec deck and on a 128^3 version of it
    from mps3d_amd.chemistry import h2_o2, ion_elec
    from mps3d_amd.plasma import H2_O2_DECK, ION_ELEC_DECK, PlasmaRun
    for tag, base, chem, n, reps, arith in (("32", ION_ELEC_DECK, ion_elec, 32, 10, "exact"),
                                            ("128", ION_ELEC_DECK, ion_elec, 128, 3, "exact"),
                                            ("h2_o2_32", H2_O2_DECK, h2_o2, 32, 10, "exact"),
                                            ("32_fast", ION_ELEC_DECK, ion_elec, 32, 10, "fast"),
                                            ("128_fast", ION_ELEC_DECK, ion_elec, 128, 3, "fast"),
                                            ("h2_o2_32_fast", H2_O2_DECK, h2_o2, 32, 10, "fast")):
        deck = dict(base)
        deck["n"] = (n, n, n)
        run = PlasmaRun(deck, chem=chem(), arith=arith)
        run.run(3)
        torch.cuda.synchronize()
        per = []
        for _ in range(reps):
            t0 = time.perf_counter()
            run.run(1)
            torch.cuda.synchronize()
            per.append((time.perf_counter() - t0) * 1e3)
        # the median step: a step's wall time now and then holds a host
        # stall of several milliseconds
        res["plasma_step_%s" % tag] = {"ms_per_step": round(float(np.median(per)), 2),
                                       "ms_mean": round(float(np.mean(per)), 2),
                                       "cycles_last_step": list(run.cycles[-1]), "p": list(deck["p"])}
        run.close()
    return res


def measured_traffic(kernel, n):
    """HBM bytes per launch of the roofline kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload, if there is
    one for this kernel and grid size."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s_%d.json" % (kernel, n))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("traffic_bytes") if str(d.get("kernel", "")).split("<")[0] == kernel else None


# ranks of the reference CPU run: the box's CPU share for one GPU is 16
# cores, whatever os.cpu_count() reports there
CPU_SHARE = 16
TWIN_N = 256


def twin_ranks():
    ncpu = min(os.cpu_count() or 1, CPU_SHARE)
    for r, p in ((64, (4, 4, 4)), (32, (2, 4, 4)), (16, (2, 2, 4)), (8, (2, 2, 2)), (4, (1, 2, 2)), (2, (1, 1, 2))):
        if ncpu >= r:
            return p
    return (1, 1, 1)


def twin_tolerance(case, tol_rel):
    """The reference's stopping rule (an absolute RMS of the residual) for a
    relative residual tol_rel: tol_rel times the RMS of b (the residual of
    the zero initial guess)."""
    from oracle import oracle as O
    P = O.Problem(case.n, case.length, (1, 1, 1), case.bc_codes, case.bc_vals, case.dcoeff, source=case.source)
    return tol_rel * O.rms(O.bvec(P))


def cpu_baseline(n=TWIN_N, tol_rel=1e-10, max_cycles=40):
    """The reference CPU path timed directly on the n^3 twin of the benchmark
    problem (even n: the reference's own hierarchy on its process grid): the
    reference Fortran/MPI library built from the reference sources into
    oracle/_ref/, one rank per core of the GPU's host share, V-cycles from
    the zero guess until its RMS residual reaches tol_rel times that of the
    zero guess (its own stopping rule). Falls back to the oracle's C
    restatement on a 32^3 sample when the reference build or mpiexec is
    absent."""
    from mps3d_amd import workloads as W
    from oracle import refrun

    ncpu = os.cpu_count() or 1
    case = W.varcoef_poisson(n)
    tol_abs = twin_tolerance(case, tol_rel)
    ranks = twin_ranks()
    if refrun.available():
        d = dict(mode=refrun.MODE_SOLVE, n=case.n, len=case.length, p=ranks, bc_codes=list(case.bc_codes),
                 bc_vals=case.bc_vals, dcoeff=case.dcoeff, source=case.source, nvcycles=max_cycles, tol=tol_abs)
        try:
            o = refrun.run(d, timeout=900)
            t = float(o["TIME"][0])
            hist = np.asarray(o["HIST"], dtype=float)
            cyc = int(np.count_nonzero(hist > 0))
            return {"value": round(t * 1e3, 1), "unit": "ms per %d^3 solve to rel. residual %.0e (reference algorithm)"
                    % (n, tol_rel), "cores": int(np.prod(ranks)), "kind": "reference", "os_cpu_count": ncpu,
                    "cycles": cyc, "converged": bool(hist[cyc - 1] <= tol_abs) if cyc else False,
                    "sample": "reference Fortran/MPI (amdflang -O3, MPICH), %d ranks on a %dx%dx%d process grid, "
                              "the %d^3 twin of the same var-coef Poisson, FP64, V(2,2) from the zero guess: "
                              "%d V-cycles in %.2f s, measured (not scaled)"
                              % (int(np.prod(ranks)), ranks[0], ranks[1], ranks[2], n, cyc, t)}, tol_abs, ranks
        except Exception as e:  # fall through to the port
            err = str(e)[:200]
    else:
        err = "reference binary or mpiexec missing"
    from oracle import oracle as O
    m, ncyc = 32, 10
    c32 = W.varcoef_poisson(m)
    P = O.Problem(c32.n, c32.length, (1, 1, 1), c32.bc_codes, c32.bc_vals, c32.dcoeff, source=c32.source)
    t = time.perf_counter()
    O.solve(P, P.zeros(), ncyc, 0.0)
    t = time.perf_counter() - t
    return {"value": round(t / (m ** 3 * ncyc) * n ** 3 * 14 * 1e3, 1), "unit": "ms per solve (scaled)",
            "cores": 1, "kind": "port", "os_cpu_count": ncpu,
            "sample": "oracle C restatement, 32^3 twin, 10 V-cycles in %.2f s, scaled to %d^3 x 14 cycles (%s)"
                      % (t, n, err)}, tol_abs, ranks


def gpu_reference_twin(tol_abs, ranks, n=TWIN_N, steps=3, max_cycles=40):
    """The GPU running the reference algorithm on the same twin and process
    grid as the CPU baseline: reference hierarchy, plain FP64, the same RMS
    stopping rule, exact arithmetic (the reference's rounding) and fast
    arithmetic. ms per solve from the zero guess."""
    import torch

    from mps3d_amd import FieldSolver, to_device
    from mps3d_amd import workloads as W

    case = W.varcoef_poisson(n)
    out = {"n": n, "p": list(ranks), "tol_rms": tol_abs}
    for arith in ("exact", "fast"):
        s = FieldSolver(case.n, case.length, hierarchy="ref", p=ranks, precision="fp64", max_cycles=max_cycles,
                        tol=tol_abs, tol_kind="rms", arith=arith)
        s.set_case(case)
        u = to_device(np.zeros(s.shape))
        s.solve(u)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            u.zero_()
            k, r, _, st = s.solve(u)
        torch.cuda.synchronize()
        out[arith] = {"ms_per_solve": round((time.perf_counter() - t0) * 1e3 / steps, 2), "cycles": int(k),
                      "status": int(st)}
        s.close()
        del u
        torch.cuda.empty_cache()
    return out


def stream_ceilings(bytes_per_array=1 << 29, reps=10):
    """Hand-written streaming kernels (mps_time_stream: 16-byte accesses,
    256 threads) over 512 MiB per array, far beyond the 256 MB Infinity
    Cache: read, write, copy, triad and the sweep's pattern (three streams
    read, one written), GB/s."""
    from mps3d_amd.solver import STREAM_KINDS, time_stream
    return {k: round(time_stream(k, bytes_per_array, reps)[0], 1) for k in STREAM_KINDS}


def coupled(steps_c3=4):
    """The coupled configurations around the field solve: the 128^3 glow
    discharge (FP64, GMRES(30) preconditioned by one V-cycle, dt 1e-11) per
    implicit step, and the adjoint gradient of the 96^3 model over 50 steps
    with checkpoints every 10 steps against its forward pass."""
    import torch

    from mps3d_amd import discharge as DS
    from mps3d_amd.coupled import GlowModel, seeded_source
    out = {}
    model, s0 = DS.glow_discharge(n=128, tol=1e-10)
    V = torch.tensor([1000.0], dtype=torch.float64, device="cuda")
    with torch.no_grad():
        st = model.run(V, 1, s0)
        torch.cuda.synchronize()
        model.step_times.clear()
        model.run(V, steps_c3, st, timed=True)
    counts = model.solve_counts()
    out["glow_discharge_128"] = {"ms_per_step": round(1e3 * float(np.median(model.step_times)), 2),
                                 "steps": steps_c3,
                                 "gmres_iterations_last": {k: v[-1] for k, v in counts.items() if k != "poisson"},
                                 "poisson_cycles_last": counts["poisson"][-1]}
    model.close()
    torch.cuda.empty_cache()
    m = GlowModel(n=96, dt=1e-11, p=(3, 3, 3), cycles=(30, 20))
    q0 = seeded_source((96, 96, 96), device="cuda")
    Vt = torch.tensor(1000.0, dtype=torch.float64, device="cuda")
    sm = torch.tensor(1.0, dtype=torch.float64, device="cuda")
    with torch.no_grad():
        m.run(Vt, sm, q0, 2)
        torch.cuda.synchronize()
        t = time.perf_counter()
        m.run(Vt, sm, q0, 50)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t
    Vg = Vt.clone().requires_grad_(True)
    sg = sm.clone().requires_grad_(True)
    qg = q0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    J = m.run(Vg, sg, qg, 50, checkpoint_every=10)
    J.backward()
    torch.cuda.synchronize()
    t_both = time.perf_counter() - t
    out["adjoint_96_50steps_ckpt10"] = {"forward_s": round(t_fwd, 3), "forward_and_gradient_s": round(t_both, 3),
                                        "adjoint_over_forward": round((t_both - t_fwd) / t_fwd, 3)}
    m.close()
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=257)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-step-kernels", action="store_true",
                    help="time only the field and transport calls around the solves (for profiling)")
    ap.add_argument("--only-solves", action="store_true",
                    help="only the timed solves and the calibration copy (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the whole 257^3 domain fits one MI355X "
              "(multi-GPU sharding is not implemented)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(3)
    torch.cuda.set_device(0)
    if args.only_step_kernels:
        print(json.dumps({"n": args.n, "step_kernels": step_kernels(args.n)}), flush=True)
        return
    case, hier, res, err = gpu_solves(args.n, args.steps, args.warmup)
    log("solves done")
    copy_gbps = calibration_copy()
    dc, f64 = res["fp32dc"], res["fp64"]
    kc = dc["kernels"]["colour"]
    traffic = measured_traffic(kc["name"], args.n)
    line = {
        "metric": "257^3 MG solve time to 1e-10: FP32+refine vs FP64, GB/s; adjoint/fwd ratio",
        "value": round(dc["ms_per_solve"], 3),
        "unit": "ms per %d^3 solve to rel. residual 1e-10 (FP32 smoothing + FP64 defect correction)" % args.n,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dc["ms_per_solve"], 3),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "fp32+fp64", "data": "synthetic",
        "config": {"workload": "varcoef_poisson_%d" % args.n, "n": args.n, "domain_m": 0.02,
                   "hierarchy": hier, "smoother": "RB-SOR V(2,2), omega 1.15 (library default on the aligned hierarchy)" if hier == "aligned" else "RB-GS V(2,2)",
                   "start": "full multigrid" if hier == "aligned" else "zero guess", "tol_rel_l2": 1e-10,
                   "seed": 20261202, "arith": "fast",
                   "fp32_cycles_per_fp64_defect": "chosen per step (the first for a nominal 0.1 per cycle, then "
                                                  "the fewest that reach the tolerance at the observed "
                                                  "contraction; at most 4)"},
        "fp64_ms": round(f64["ms_per_solve"], 3),
        "speedup_fp32dc_vs_fp64": round(f64["ms_per_solve"] / dc["ms_per_solve"], 3),
        "cycles": {"fp32dc": dc["cycles"], "fp64": f64["cycles"]},
        "vcycles": {"fp32dc": dc["vcycles"], "fp64": f64["vcycles"]},
        "plain_start": {"fp32dc": dc.get("plain_start"), "fp64": f64.get("plain_start")},
        "relres": {"fp32dc": dc["relres"], "fp64": f64["relres"]},
        "fp32dc_vs_fp64_rel_l2": err,
        "rel_l2_err_vs_converged": {"fp32dc": dc["err_vs_converged"], "fp64": f64["err_vs_converged"],
                                    "reference": "FP64 solve of the same problem to rel. residual 1e-13"},
        "unknowns_per_s": {"fp32dc": dc["unknowns_per_s"], "fp64": f64["unknowns_per_s"]},
        "solve_GBps_alg": {"fp32dc": round(dc["solve_GBps"], 1), "fp64": round(f64["solve_GBps"], 1)},
        "adjoint_ms": round(f64["adjoint_ms"], 3),
        "adjoint_fwd_ratio": round(f64["adjoint_ms"] / f64["ms_per_solve"], 3),
        "roofline": {"bound": "hbm", "kernel": kc["name"] + " (FP32 red-black sweep, finest level)",
                     "achieved": round(kc["GBps"], 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(kc["GBps"] / PEAK_HBM_GBPS, 4), "traffic": traffic,
                     "avg_us": round(kc["avg_us"], 2), "launches_timed": kc["launches"],
                     "timing": "HIP events around every finest-level launch of one solve like the timed ones",
                     "avg_us_isolated_cold_cache": round(kc["hook_us"], 2), "alg_bytes": kc["alg_bytes"],
                     "copy_GBps": round(copy_gbps, 1), "frac_of_copy": round(kc["GBps"] / copy_gbps, 4)},
        "kernels": {"fp32dc": dc["kernels"], "fp64": f64["kernels"]},
    }
    if args.only_solves:
        print(json.dumps(line), flush=True)
        return
    line["stream_GBps"] = stream_ceilings()
    ceiling = line["stream_GBps"]["sweep"]
    line["roofline"]["stream_ceiling_GBps"] = ceiling
    line["roofline"]["frac_of_stream"] = round(kc["GBps"] / ceiling, 4)
    log("stream ceilings done")
    line["step_kernels"] = step_kernels(args.n)
    log("step kernels done")
    line["coupled"] = coupled()
    line["adjoint_fwd_ratio_c4"] = line["coupled"]["adjoint_96_50steps_ckpt10"]["adjoint_over_forward"]
    log("coupled configurations done")
    if not args.no_cpu:
        cb, tol_abs, ranks = cpu_baseline()
        line["cpu_baseline"] = cb
        log("cpu baseline done")
        line["gpu_reference_twin"] = gpu_reference_twin(tol_abs, ranks)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
