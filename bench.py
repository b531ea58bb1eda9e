#!/usr/bin/env python3
"""Benchmark of the mef Biot u-p path on one MI355X.

Workload: a box of n^3 hex8 cells promoted to the reference's hexa20/8
element, with a seeded log-normal permeability field (sigma 1, correlation
length 8 cells), E ~ U[5, 50] MPa and nu ~ U[0.25, 0.35] per element, a seeded
10 % patch of 1 kPa top load, dt = 0.1 s, SQMR with Jacobi to 1e-10.  A "step"
is one backward-Euler time step of the solv macro.  The default n = 87 gives
8.67M equations, the 8.6M-dof size the metric is quoted at (with the
reference's element a 128^3-cell box has 27.6M equations; `--n 128` runs it).
`--element q1` runs the same workload with the equal-order Q1/Q1 hexa8
element (element type 7, the mesh kept as read), default n = 128: the
configuration's 128^3-cell mesh at its nominal 8.6M dofs.

The FP32+IR form (FP32 matrix values and Jacobi inside an FP64
iterative-refinement loop) runs first, then the FP64 form, each from the zero
state: W untimed warmup steps, then K timed steps -- both clamped to what
fits the wall-clock budget (--budget, default 450 s for the whole run), with
at least one timed FP32+IR step; the JSON line reports the steps actually
timed and the ones requested.  `value` is the FP32+IR form's time steps per
second.  The SpMV kernel time is measured with HIP events on the library's
stream inside the timed steps.  The CPU baseline is the reference's OpenMP
path restated in C (oracle/pm_omp.c), timed on this host at the benchmark's
own size with every available thread and with one; its per-step figure is
scaled from measured per-iteration and per-element costs (see
cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

T_START = time.perf_counter()
ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "time-steps/s at 8.6M-dof Biot u-p (FP32+IR vs FP64 @1e-10); SpMV GB/s"
# (lanes per node block, columns in flight per lane, index prefetch / runs)
# of the SpMV shapes 0..15
shape_lanes = (64, 64, 64, 64, 64, 32, 64, 64, 64, 64, 64, 64, 64, 64, 64, 32)
shape_unr = (1, 2, 3, 4, 5, 4, 3, 4, 5, 1, 2, 1, 2, 5, 1, 1)
shape_pf = (0, 0, 0, 0, 0, 0, 1, 1, 1, 2, 2, 3, 3, 4, 5, 6)


def spmv_kernel_name(prec, shape):
    """The SpMV kernel the solver launches at this shape, with its fused dot."""
    vt = "float" if prec == "fp32" else "double"
    if shape_pf[shape] == 6:
        return f"k_spmv_chg<{vt}, {shape_lanes[shape]}, {shape_unr[shape]}, 1, false>"
    if shape_pf[shape] == 5:
        return f"k_spmv_cha<{vt}, 1, false>"
    if shape_pf[shape] == 4:
        return f"k_spmv_pf<{vt}, {shape_unr[shape]}, 1, false, true>"
    if shape_pf[shape] == 3:
        return f"k_spmv_ch<{vt}, {shape_unr[shape]}, 1, false, true>"
    if shape_pf[shape] == 2:
        return f"k_spmv_ch<{vt}, {shape_unr[shape]}, 1, false, false>"
    if shape_pf[shape]:
        return f"k_spmv_pf<{vt}, {shape_unr[shape]}, 1, false, false>"
    return f"k_spmv<{vt}, {shape_lanes[shape]}, {shape_unr[shape]}, 1, false>"
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(info, value_bytes, fused_dot=False, shape=None):
    """Bytes one SpMV launch must move in the node-block layout: every value
    once, the column structure once, the per-block metadata, x once, y once,
    and w once when the solver's dot y.w is fused into it.  The column
    structure is one shared 32-bit index per column (colptr, valptr: 8 B
    each, first row and row count: 4 B each per block), or for the
    run-per-lane shapes one 32-bit word per run of <= 4 consecutive columns
    (plus chptr: 8 B per block)."""
    runs = shape is not None and shape_pf[shape] in (2, 3, 5, 6)
    cols = 4 * info["nruns"] if runs else 4 * info["ncols"]
    meta = (32 if runs else 24) * info["nblocks"]
    return value_bytes * info["nnz"] + cols + meta + (24 if fused_dot else 16) * info["neq"]


def spmv_alg_bytes(info, value_bytes):
    """Algorithmic bytes of one SpMV of the full matrix as plain CSR: every
    stored value (value_bytes) and its int32 column once, int64 row
    pointers, x read once and y written once in the value precision's
    vectors -- FP64: 12 nnz + 8 (neq + 1) + 16 neq; FP32 values (the
    refinement's inner product): 8 nnz + 8 (neq + 1) + 8 neq."""
    vec = 16 if value_bytes == 8 else 8
    return (value_bytes + 4) * info["nnz"] + 8 * (info["neq"] + 1) + vec * info["neq"]


def csr_bytes(info):
    """Bytes of the same product as a plain full CSR (int32 columns, FP64
    values, int64 row pointers, x once, y once)."""
    return 12 * info["nnz"] + 8 * (info["neq"] + 1) + 16 * info["neq"]


def measured_traffic(prec, shape, info):
    """HBM bytes per launch of the SpMV at this shape and workload, from the
    committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by a
    copy kernel of known size), or None when it was not measured for them."""
    path = ROOT / "profiles" / "spmv_87cube_hbm_counters.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if not line.startswith("<!-- traffic "):
            continue
        kv = dict(t.split("=", 1) for t in line[len("<!-- traffic "):].rstrip(" ->").split())
        if (kv.get("prec") == prec and int(kv.get("shape", -1)) == shape
                and int(kv.get("neq", -1)) == info["neq"] and int(kv.get("nnz", -1)) == info["nnz"]):
            return float(kv["bytes"])
    return None


def cpu_info():
    """The host's CPU model (lscpu's "Model name", /proc/cpuinfo) and the
    number of threads this process may use (the OpenMP team size)."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(env)) if env.isdigit() and int(env) > 0 else avail
    return model, threads, avail


def cpu_baseline(solver, deck, fp64_iters_per_step, fp64_nlit, seed=0):
    """The reference's threaded CPU path (oracle/pm_omp.c: rpsqmr_omp with
    the per-thread-buffer CSRC SpMV and reduction dots, and the coloured
    element loop of pform_pm) timed on this host at the benchmark's own size,
    with every available thread and with one.  The CSRC matrix is the device
    matrix exported in the reference's layout.  A whole step at this size
    would take an hour of CPU time, so the step is priced from measured
    parts: the time per SQMR iteration (over a few iterations) times the FP64
    form's iteration count of this run, plus the time per element of the isw 2
    and isw 4 element loops (over a sample of one colour) times the step's
    element passes."""
    from oracle import oracle as O

    model, nall, avail = cpu_info()
    t0 = time.perf_counter()
    ia, ja, ad, al = solver.export_csrc()
    ix20, idm, _ = solver.numbering()
    nnode, neq = solver.nnode, solver.neq
    fnno = np.zeros(nnode, np.int32)
    fnno[: deck.nnodev] = 1
    st = dict(ix=np.ascontiguousarray(ix20), nnode=nnode, nnodev=deck.nnodev, numel=deck.numel,
              id=np.ascontiguousarray(idm), fnno=fnno, neq=neq, nad=int(ia[-1] - 1), ia=ia, ja=ja,
              nen=deck.nen)
    m = O.OracleModel(deck, st)
    m.ad[:] = ad
    mvec = 1.0 / ad  # Jacobi, precond 2 (pre_diag, Precond.f:22-44)
    b = np.random.default_rng(seed).standard_normal(neq)
    prep_s = time.perf_counter() - t0
    out = {"model": model, "threads_available": avail}
    for label, nth, kit, nel2 in (("all", nall, 6, 2000 * nall), ("one", 1, 2, 1500)):
        t1 = time.perf_counter()
        it, _, fail = O.omp_sqmr(st, ad, al, mvec, b, tol=1e-300, maxit=kit, nthreads=nth)
        t_sq = time.perf_counter() - t1
        per_iter = t_sq / (kit + 1)  # + the start: one product and three dots
        nc, col = m.colour() if label == "all" else (nc, col)
        t_e2, n2 = m.pform_colour(col, 1, isw=2, lhs=True, rhs=True, max_elems=nel2, nthreads=nth)
        t_e4, n4 = m.pform_colour(col, 2, isw=4, lhs=False, rhs=True, max_elems=max(nel2 // 4, 1),
                                  nthreads=nth)
        per_e2, per_e4 = t_e2 / max(n2, 1), t_e4 / max(n4, 1)
        step_s = (fp64_iters_per_step * per_iter + deck.numel * (fp64_nlit * per_e2 + 2 * per_e4))
        out[label] = {"threads": nth, "s_per_sqmr_iteration": per_iter,
                      "s_per_element_isw2": per_e2, "s_per_element_isw4": per_e4,
                      "sqmr_iterations_timed": kit, "elements_timed": [n2, n4],
                      "s_per_step": step_s, "time_steps_per_s": 1.0 / step_s,
                      "sample_seconds": t_sq + t_e2 + t_e4}
        log(f"cpu baseline ({label}, {nth} threads): {per_iter:.3f} s per SQMR iteration, "
            f"{per_e2 * 1e3:.3f} / {per_e4 * 1e3:.3f} ms per element (isw 2 / 4); "
            f"{step_s:.0f} s per step")
    del m, ia, ja, ad, al
    a = out["all"]
    return {
        "value": a["time_steps_per_s"],
        "unit": "time-steps/s",
        "cores": a["threads"],
        "kind": "port",
        "model": model,
        "sample": (f"OpenMP restatement of the reference's CPU path (rpsqmr_omp, "
                   f"matvec_csrc_sym_pm_omp, dot_par_omp, the coloured pform_pm loop) on the "
                   f"benchmark box itself, {a['threads']} threads on {model}: "
                   f"{a['sqmr_iterations_timed']} SQMR iterations and a sample of "
                   f"{a['elements_timed'][0]} isw 2 / {a['elements_timed'][1]} isw 4 elements "
                   f"timed; scaled to a step by the FP64 form's {fp64_iters_per_step} "
                   f"iterations and {fp64_nlit} + 2 element passes measured in this run"),
        "scaled_per_step": True,
        "one_core": out["one"],
        "all_cores": a,
        "prep_s": prep_s,
    }


def batch_form(args):
    """The batching form: nbox independent boxes (n^3 cells, n ~ U[nmin, nmax])
    as one block-diagonal system, every box with its own stop tests; K steps
    of FP64 and of FP32+IR.  Prints one JSON line (box-steps/s)."""
    from poromec_amd import problem as P
    from poromec_amd.solver import MefSolver, SolverConfig

    nen = 8 if args.element == "q1" else 20
    decks = P.batch_decks(nbox=args.nbox, seed=args.seed, nmin=args.nmin, nmax=args.nmax, nen=nen)
    big, off = P.concat_decks(decks)
    t0 = time.perf_counter()
    s = MefSolver(big)
    s.batch_define(off)
    if args.batch_eager:  # launch the iteration chunks one kernel at a time
        s.batch_graphs(False)
    setup_s = time.perf_counter() - t0
    s.pform(2, True, False, big.dt, big.dt)
    shapes = {name: s.tune_spmv(use32, reps=5)[0] for use32, name in ((False, "fp64"), (True, "fp32"))}
    log(f"batch ({args.element}): {args.nbox} boxes, neq {s.neq}, setup {setup_s:.1f} s, "
        f"shapes {shapes}")
    out = {}
    forms = [f for f in (("fp64", SolverConfig(tol=1e-10)),
                         ("fp32ir", SolverConfig(tol=1e-10, precision="fp32ir")))
             if f[0] in args.batch_forms.split(",")]
    for name, cfg in forms:
        z4, zn = np.zeros((s.nnode, 4)), np.zeros(s.nnode)
        s.set_state(u=z4, u0=z4, dp=zn, porosity=zn, t=0.0)
        for _ in range(args.warmup):
            s.step(cfg)
        t0 = time.perf_counter()
        its = []
        for _ in range(args.steps):
            s.step(cfg)
            its.append(s.batch_iters().tolist())
        wall = time.perf_counter() - t0
        it = np.array(its)
        out[name] = {"box_steps_per_s": args.nbox * args.steps / wall, "ms_per_step": 1e3 * wall / max(args.steps, 1),
                     "iters_per_box_step": {"min": int(it.min()), "mean": float(it.mean()), "max": int(it.max())}}
        log(f"  {name}: {out[name]}")
    s.close()
    # the same boxes one at a time (a sample of them), for comparison
    seq = None
    if args.seq_sample > 0:
        k = min(args.seq_sample, args.nbox)
        wall = 0.0
        for i, d in enumerate(decks[:k]):
            if i % 32 == 0:
                log(f"  one box at a time: box {i} of {k}, {wall:.1f} s of steps so far")
            one = MefSolver(d)
            cfg = SolverConfig(tol=1e-10)
            for _ in range(args.warmup):
                one.step(cfg)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                one.step(cfg)
            wall += time.perf_counter() - t0
            one.close()
        seq = {"boxes_timed": k, "box_steps_per_s": k * args.steps / wall,
               "s_per_step_all_boxes": wall / max(args.steps, 1)}
        log(f"  one box at a time (fp64, first {k} boxes): {seq}")
    print(json.dumps({"form": "batch", "element": args.element, "graphs": not args.batch_eager,
                      "nbox": args.nbox, "n_range": [args.nmin, args.nmax],
                      "neq": s.neq, "steps": args.steps, "warmup": args.warmup,
                      "setup_s": setup_s, "spmv_shapes": shapes, **out,
                      "one_box_at_a_time_fp64": seq}), flush=True)


def adjoint_form(args):
    """The differentiation form: misfit of pore-pressure histories at nsens
    seeded sensor nodes against targets from a second seeded field, over
    nsteps steps; the adjoint gradient w.r.t. per-element log k and log E
    (one material per element) timed, then checked against central finite
    differences along ndir seeded random directions."""
    from poromec_amd import problem as P
    from poromec_amd.adjoint import MarchMisfit
    from poromec_amd.solver import MefSolver, SolverConfig

    n = args.n if args.n is not None else 87
    nen = 8 if args.element == "q1" else 20
    deck = P.heterogeneous_box_deck(n, seed=args.seed, dt=0.1, nsteps=args.nsteps,
                                    corr_len=min(8.0, n / 2), nen=nen)
    cfg = SolverConfig(tol=1e-13, maxit=200000)
    rng = np.random.default_rng(args.seed + 1)
    probe = MefSolver(deck)
    _, idm, _ = probe.numbering()
    free_p = np.nonzero(idm[: deck.nnodev, 3] > 0)[0]
    sens = [(int(v), 3) for v in rng.choice(free_p, size=min(args.nsens, len(free_p)), replace=False)]
    tdeck = P.heterogeneous_box_deck(n, seed=args.seed + 100, dt=0.1, nsteps=args.nsteps,
                                     corr_len=min(8.0, n / 2), nen=nen)
    probe.set_materials(tdeck)
    obs = np.zeros((args.nsteps, len(sens)))
    for k in range(args.nsteps):
        probe.step(cfg)
        u = probe.state()["u"]
        obs[k] = [u[a, b] for a, b in sens]
    probe.close()
    m = MarchMisfit(deck, sens, obs, nsteps=args.nsteps, cfg=cfg, ckpt_every=args.ckpt)
    lE, lk = (t.numpy() for t in m.params())
    t0 = time.perf_counter()
    J, gE, gk = m._evaluate(lE, lk)
    t_adj = time.perf_counter() - t0
    log(f"adjoint {n}^3 ({args.element}), {args.nsteps} steps, checkpoints every {args.ckpt}: "
        f"J {J:.12e}, forward plus adjoint {t_adj:.1f} s")
    g = np.concatenate([gE, gk])
    h, errs = args.fd_h, []
    for _ in range(args.ndir):
        v = rng.standard_normal(2 * len(lE))
        jp = m._evaluate(lE + h * v[: len(lE)], lk + h * v[len(lE):])[0]
        jm = m._evaluate(lE - h * v[: len(lE)], lk - h * v[len(lE):])[0]
        fd = (jp - jm) / (2 * h)
        errs.append(abs(fd - g @ v) / max(abs(fd), 1e-300))
        log(f"  direction {len(errs)}: fd {fd:.12e} adjoint {g @ v:.12e} rel {errs[-1]:.2e}")
    print(json.dumps({"form": "adjoint", "element": args.element, "n": n, "numel": deck.numel,
                      "neq": m.solver.neq,
                      "nsteps": args.nsteps, "nsens": len(sens), "ckpt_every": args.ckpt,
                      "J": J, "forward_plus_adjoint_s": t_adj, "fd_h": h, "ndir": args.ndir,
                      "fd_rel_err_max": max(errs) if errs else None,
                      "fd_rel_err_median": float(np.median(errs)) if errs else None}), flush=True)


def tform_form(args):
    """The pnsf stress recovery (pm_tform) on the n^3 box: vertex total,
    Terzaghi and Biot stresses and Darcy flux of a smooth state, timed per
    call (the kernel, no host copies), against its compulsory HBM bytes; the
    oracle's tform_pm restatement timed on a small box beside it."""
    import ctypes as C

    from oracle import oracle as O
    from poromec_amd import problem as P
    from poromec_amd.solver import MefSolver

    def state(deck, nnode):
        ix = np.arange(nnode, dtype=np.float64)
        u = np.stack([1e-3 * np.sin(ix * 1e-3), 1e-3 * np.cos(ix * 7e-4), -1e-3 * np.sin(ix * 3e-4),
                      np.cos(ix * 1e-4)], axis=1)
        u[deck.nnodev:, 3] = 0.0
        dp = 1e-2 * np.sin(ix * 5e-4)
        dp[deck.nnodev:] = 0.0
        return u, dp

    n = args.n if args.n is not None else 87
    deck = P.heterogeneous_box_deck(n, seed=args.seed, dt=0.1, nsteps=1)
    s = MefSolver(deck)
    u, dp = state(deck, s.nnode)
    s.set_state(u=u, u0=u, dp=dp, t=0.0)
    tx = s.tform()[0]  # first call builds the node -> (element, corner) lists
    null = C.c_void_p()
    reps = max(args.spmv_reps, 5)
    s.lib.pm_tform(s.ctx, null, null, null, null)
    t0 = time.perf_counter()
    for _ in range(reps):
        rc = s.lib.pm_tform(s.ctx, null, null, null, null)
        assert rc == 0, rc
    ms = (time.perf_counter() - t0) / reps * 1e3
    npairs = 8 * deck.numel
    nbytes = (s.nnode * (4 + 1 + 21) * 8 + deck.nnodev * 3 * 8 + deck.numel * (20 + 1) * 4
              + (s.nnode + 1) * 8 + npairs * 4)
    log(f"tform {n}^3: {ms:.3f} ms per call, {nbytes / ms / 1e6:.0f} GB/s compulsory")
    # oracle restatement on a small box, per element
    m_small = max(4, min(n, 12))
    dsm = P.heterogeneous_box_deck(m_small, seed=args.seed, dt=0.1, nsteps=1)
    m = O.OracleModel(dsm)
    us, dps = state(dsm, m.st["nnode"])
    m.u[:] = us
    m.dp[:] = dps
    t0 = time.perf_counter()
    k = 0
    while time.perf_counter() - t0 < 2.0 or k < 2:
        m.tform()
        k += 1
    cpu_s = (time.perf_counter() - t0) / k
    cpu_per_elem = cpu_s / dsm.numel
    print(json.dumps({"form": "tform", "n": n, "numel": deck.numel, "nnode": s.nnode,
                      "ms_per_call": ms, "compulsory_bytes": nbytes,
                      "GBps_compulsory": nbytes / ms / 1e6, "frac_of_8TBps": nbytes / ms / 8e9,
                      "elements_per_s": deck.numel / ms * 1e3,
                      "cpu_oracle": {"box": m_small, "s_per_call": cpu_s,
                                     "elements_per_s": 1.0 / cpu_per_elem, "cores": 1,
                                     "kind": "port"},
                      "max_abs_tx": float(np.abs(tx).max())}), flush=True)
    s.close()


def vprop_form(args):
    """The variable-property element (elmt18_pm, every conseq flag on)
    against elmt17_pm on the n^3 box: one FP64 step each from the same
    state, with the step's element and solver time; the oracle's step with
    and without the properties timed on a small box beside it."""
    from oracle import oracle as O
    from poromec_amd import problem as P
    from poromec_amd.solver import MefSolver, SolverConfig

    def deck(n):
        d = P.heterogeneous_box_deck(n, seed=args.seed, dt=0.1, patch_fraction=0.1, nsteps=1)
        d.e[:, 5], d.e[:, 7] = 2000.0, 0.3
        return d

    n = args.n if args.n is not None else 87
    cfg = SolverConfig(tol=1e-10, stop_crit=1)
    out = {"form": "vprop", "n": n}
    for label, ie in (("elmt17", None), ("elmt18", 18)):
        d = deck(n)
        s = MefSolver(d)
        if ie:
            s.set_vprop([18] * d.e.shape[0], (1, 1, 1, 1))
        info = s.step(cfg)
        out[label] = {"ms_total": info["ms_total"], "ms_element": info["ms_element"],
                      "ms_solve": info["ms_solve"], "nlit": info["nlit"], "iters": info["iters"]}
        log(f"{label} {n}^3: step {info['ms_total']:.1f} ms, elements {info['ms_element']:.1f} ms, "
            f"solves {info['ms_solve']:.1f} ms, iterations {info['iters']}")
        out["numel"] = d.numel
        s.close()
    m_small = max(3, min(n, 6))
    cpu = {}
    for label, kw in (("elmt17", {}), ("elmt18", {"ie": [18], "vprop": (1, 1, 1, 1)})):
        dsm = deck(m_small)
        dsm.e = dsm.e[:1]
        dsm.ix[:, 20] = 1
        m = O.OracleModel(dsm, stop_crit=1, **kw)
        t0 = time.perf_counter()
        m.step()
        cpu[label] = time.perf_counter() - t0
    out["cpu_oracle"] = {"box": m_small, "s_per_step": cpu, "cores": 1, "kind": "port"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--budget", type=float, default=float(os.environ.get("PM_BENCH_BUDGET_S", 450)),
                    help="wall-clock seconds for the whole run (setup, tuner, both forms, CPU "
                         "leg): warmup and timed steps are clamped to what fits, at least one "
                         "timed step of the FP32+IR form")
    ap.add_argument("--margin", type=float, default=20.0,
                    help="seconds kept free at the end of the budget")
    ap.add_argument("--cpu-reserve", type=float, default=60.0,
                    help="seconds the CPU baseline leg is allowed for")
    ap.add_argument("--only", choices=["fp32ir", "fp64"], default=None,
                    help="time one form only (the 128^3 box: a step of each takes minutes); "
                         "--only fp64 reports the FP64 form's time steps per second as value")
    ap.add_argument("--live-every", type=int, default=16,
                    help="time every k-th solver SpMV with HIP events")
    ap.add_argument("--n", type=int, default=None,
                    help="box edge in cells (default 87 for hexa20/8, 128 for Q1/Q1: 8.6M dofs "
                         "either way)")
    ap.add_argument("--element", choices=["hexa20", "q1"], default="hexa20",
                    help="hexa20: the reference's hexa20/8 element (maxno 20); q1: the "
                         "equal-order Q1/Q1 hexa8 element (element type 7, maxno 8), whose "
                         "vertex-only dofs are the configurations' nominal counts")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--spmv-reps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", choices=["precision", "batch", "adjoint", "tform", "vprop"],
                    default="precision",
                    help="precision (default, the metric's workload), batch or adjoint form")
    ap.add_argument("--nbox", type=int, default=256)
    ap.add_argument("--nmin", type=int, default=12)
    ap.add_argument("--nmax", type=int, default=32)
    ap.add_argument("--seq-sample", type=int, default=8,
                    help="batch form: boxes also solved one at a time, for comparison")
    ap.add_argument("--batch-forms", default="fp64,fp32ir",
                    help="batch form: the forms stepped as one block-diagonal system")
    ap.add_argument("--batch-eager", action="store_true",
                    help="batch form: launch the iteration chunks kernel by kernel instead of "
                         "replaying their captured graph")
    ap.add_argument("--nsteps", type=int, default=50, help="adjoint form: march length")
    ap.add_argument("--nsens", type=int, default=64)
    ap.add_argument("--ckpt", type=int, default=10)
    ap.add_argument("--ndir", type=int, default=20)
    ap.add_argument("--fd-h", type=float, default=1e-6)
    ap.add_argument("--no-q1", action="store_true",
                    help="skip the Q1/Q1 128^3 leg the default run adds when time is left")
    ap.add_argument("--q1-reserve", type=float, default=90.0,
                    help="seconds of budget the Q1/Q1 128^3 leg needs to start")
    ap.add_argument("--spmv-only", action="store_true",
                    help="set up, then only time the FP64 / FP32 SpMV and a copy kernel "
                         "(the run profiled with hardware counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this path has no multi-GPU form "
              "(it would shard by domain decomposition, one subdomain per GPU)",
              file=sys.stderr)
        sys.exit(2)

    if args.form == "batch":
        return batch_form(args)
    if args.form == "adjoint":
        return adjoint_form(args)
    if args.form == "tform":
        return tform_form(args)
    if args.form == "vprop":
        return vprop_form(args)

    from poromec_amd import problem as P
    from poromec_amd.solver import MefSolver, SolverConfig

    q1 = args.element == "q1"
    n = args.n if args.n is not None else (128 if q1 else 87)
    budget = args.budget

    def left():
        return budget - (time.perf_counter() - T_START)

    t0 = time.perf_counter()
    deck = P.heterogeneous_box_deck(n, seed=args.seed, dt=0.1, patch_fraction=0.1, nsteps=50,
                                    nen=8 if q1 else 20)
    log(f"deck {n}^3 ({args.element}): {deck.numel} elements, {deck.nnodev} vertices "
        f"({time.perf_counter()-t0:.1f} s)")
    t0 = time.perf_counter()
    solver = MefSolver(deck)
    setup_s = time.perf_counter() - t0
    info = solver.matrix_info()
    log(f"setup {setup_s:.1f} s: nnode {solver.nnode}, neq {solver.neq}, nnz {info['nnz']}, "
        f"blocks {info['nblocks']}")

    if args.spmv_only:
        solver.pform(2, True, False, deck.dt, deck.dt)
        shapes = {}
        for use32, name in ((False, "fp64"), (True, "fp32")):
            best, ms = solver.tune_spmv(use32, reps=5)
            shapes[name] = {"best": best, "ms": [round(v, 4) for v in ms]}
            log(f"SpMV shapes {name}: " + ", ".join(f"{v:.3f}" for v in ms) + f" ms -> shape {best}")
        ms64 = solver.bench_spmv(False, args.spmv_reps)
        ms32 = solver.bench_spmv(True, args.spmv_reps)
        copy_bytes = 1 << 30
        msc = solver.bench_copy(copy_bytes, args.spmv_reps)
        sbytes = 4 << 30
        stream = {}
        for kind, name, moved in ((0, "copy_32B_gridstride", 2), (1, "read_4x16B", 1),
                                  (2, "read_4x16B_nontemporal", 1), (3, "copy_4x16B", 2),
                                  (4, "write_4x16B", 1), (5, "read_8x16B", 1)):
            ms = solver.bench_stream(kind, sbytes, args.spmv_reps)
            stream[name] = {"ms": ms, "GBps": moved * sbytes / ms / 1e6}
            log(f"stream {name}: {ms:.3f} ms, {moved * sbytes / ms / 1e6:.0f} GB/s")
        print(json.dumps({"spmv_fp64_ms": ms64, "spmv_fp32_ms": ms32,
                          "spmv_fp64_bytes": spmv_bytes(info, 8, shape=shapes["fp64"]["best"]),
                          "spmv_fp32_bytes": spmv_bytes(info, 4, shape=shapes["fp32"]["best"]),
                          "copy_ms": msc,
                          "copy_bytes_read": copy_bytes, "copy_bytes_written": copy_bytes,
                          "copy_GBps": 2 * copy_bytes / msc / 1e6, "tuned_shapes": shapes,
                          "stream_4GiB": stream,
                          **info}), flush=True)
        return

    # SpMV shape tuner on this matrix (lanes per node block x columns in flight)
    solver.pform(2, True, False, deck.dt, deck.dt)
    shapes = {}
    for use32, name in ((False, "fp64"), (True, "fp32")):
        best, ms = solver.tune_spmv(use32, reps=5)
        shapes[name] = {"best": best, "ms": [round(v, 4) for v in ms]}
        log(f"SpMV shapes {name}: " + ", ".join(f"{v:.3f}" for v in ms) + f" ms -> shape {best}")

    cfg64 = SolverConfig(precision="fp64", tol=1e-10, stop_crit=1)
    cfg32 = SolverConfig(precision="fp32ir", tol=1e-10, ir_inner_tol=1e-5, stop_crit=1)
    margin = args.margin

    def zero_state():
        z4, zn = np.zeros((solver.nnode, 4)), np.zeros(solver.nnode)
        solver.set_state(u=z4, u0=z4, dp=zn, porosity=zn, t=0.0)

    def march(cfg, max_steps, warmup, more_warmup, more_timed):
        """From the zero state: up to `warmup` untimed steps, then timed steps,
        each next one only while more_warmup(w, est) / more_timed(w, k, est)
        says the rest of the run still fits the budget (est: the seconds of
        one step so far); always at least one timed step.  Warmup steps are
        the march's first steps, so the timed ones are the steps after them.
        Each step is timed alone (pm_step returns after the device has
        finished it); the state is read back after each timed step, outside
        its time."""
        zero_state()
        done_w, est = 0, None
        while done_w < warmup:
            t1 = time.perf_counter()
            inf = solver.step(cfg)
            est = time.perf_counter() - t1
            done_w += 1
            log(f"  {cfg.precision} warmup step {done_w - 1}: {est:.1f} s, iters {inf['iters']}")
            if done_w < warmup and not more_warmup(done_w, est):
                break
        solver.spmv_timing(args.live_every)  # HIP events around every k-th solver SpMV
        walls, infos, states = [], [], []
        while len(walls) < max_steps:
            t1 = time.perf_counter()
            inf = solver.step(cfg)
            walls.append(time.perf_counter() - t1)
            infos.append(inf)
            st = solver.state()
            states.append((st["u"].copy(), st["dp"].copy()))
            log(f"  {cfg.precision} timed step {len(walls) - 1}: {walls[-1]:.1f} s, "
                f"iters {inf['iters']}, true_rel {max(inf['true_rel']):.2e}")
            est = sum(walls) / len(walls)
            if not more_timed(done_w, len(walls), est):
                break
        live = solver.spmv_timing(0)
        return done_w, walls, infos, states, live

    # FP32+IR first: it is `value`.  Its steps leave room for the FP64 form's
    # same warmup and at least as many timed steps (an FP64 step estimated at
    # 1.8 FP32+IR steps until one is measured) and for the CPU leg.
    cpu_reserve = 0.0 if args.no_cpu else args.cpu_reserve
    r64 = 0.0 if args.only == "fp32ir" else 1.8
    if args.only == "fp64":  # the FP64 form stands in the FP32+IR form's place
        cfg32 = cfg64
    w32, walls32, infos32, st32, live32 = march(
        cfg32, args.steps, args.warmup,
        lambda w, est: left() - margin >= 2 * est + r64 * est * (w + 2) + cpu_reserve,
        lambda w, k, est: left() - margin >= est + r64 * est * (w + k + 1) + cpu_reserve)
    est32 = sum(walls32) / len(walls32)
    walls64, infos64, st64, live64, w64 = [], [], [], None, 0
    if args.only is not None:
        log(f"only the {args.only} form timed")
    elif left() - margin >= r64 * est32 * (w32 + 1) + cpu_reserve:
        w64, walls64, infos64, st64, live64 = march(
            cfg64, len(walls32), w32, lambda w, est: True,
            lambda w, k, est: left() - margin >= est + cpu_reserve)
    else:
        log(f"FP64 form skipped: {left():.0f} s left of the {budget:.0f} s budget")

    # error of the FP32+IR histories against the FP64 form, step by step
    nv = deck.nnodev
    err_u = err_p = None
    if st64:
        err_u = err_p = 0.0
        for (u64, _), (u32, _) in zip(st64, st32):
            su = max(np.abs(u64[:, :3]).max(), 1e-300)
            sp = max(np.abs(u64[:nv, 3]).max(), 1e-300)
            err_u = max(err_u, float(np.abs(u32[:, :3] - u64[:, :3]).max() / su))
            err_p = max(err_p, float(np.abs(u32[:nv, 3] - u64[:nv, 3]).max() / sp))

    # SpMV kernel time inside the timed steps (HIP events on the library's
    # stream around every live_every-th solver SpMV), per value precision
    b32 = spmv_bytes(info, 4, fused_dot=True, shape=shapes["fp32"]["best"])
    b64 = spmv_bytes(info, 8, fused_dot=True, shape=shapes["fp64"]["best"])
    if args.only == "fp64":  # the timed form's SpMV is the FP64 one
        b32, live32 = b64, {"fp32_ms": live32["fp64_ms"], "fp32_n": live32["fp64_n"]}
    ms32 = live32["fp32_ms"] / max(live32["fp32_n"], 1)
    ms64 = live64["fp64_ms"] / max(live64["fp64_n"], 1) if live64 else None
    alg32 = spmv_alg_bytes(info, 8 if args.only == "fp64" else 4)
    alg64 = spmv_alg_bytes(info, 8)
    first = "fp64" if args.only == "fp64" else "fp32"
    log(f"SpMV in the timed steps (every {args.live_every}th launch timed): {first} {ms32:.3f} ms x "
        f"{live32['fp32_n']}" + (f", fp64 {ms64:.3f} ms x {live64['fp64_n']}" if ms64 else ""))

    steps32, steps64 = len(walls32), len(walls64)
    it32 = sum(sum(i["iters"]) for i in infos32) / steps32
    ms_step32 = 1e3 * sum(walls32) / steps32
    rel32 = max(r for i in infos32 for r in i["true_rel"])
    fp64 = None
    if steps64:
        it64 = sum(sum(i["iters"]) for i in infos64) / steps64
        nlit64 = max(i["nlit"] for i in infos64)
        ms_step64 = 1e3 * sum(walls64) / steps64
        fp64 = {"time_steps_per_s": 1e3 / ms_step64, "ms_per_step": ms_step64, "steps": steps64,
                "warmup": w64, "krylov_iters_per_step": it64,
                "krylov_iters_per_s": it64 / (ms_step64 / 1e3),
                "max_true_rel": max(r for i in infos64 for r in i["true_rel"]),
                "spmv_ms": ms64, "spmv_launches_timed": live64["fp64_n"],
                "spmv_GBps_algorithmic": alg64 / ms64 / 1e6,
                "spmv_GBps_compulsory": b64 / ms64 / 1e6}

    cpu = None
    if not args.no_cpu and fp64 is not None:
        try:
            cpu = cpu_baseline(solver, deck, int(round(fp64["krylov_iters_per_step"])), nlit64)
        except Exception as exc:  # the oracle is test infrastructure: optional here
            log(f"cpu baseline skipped: {exc}")

    traffic32 = (measured_traffic("fp64", shapes["fp64"]["best"], info) if args.only == "fp64"
                 else measured_traffic("fp32", shapes["fp32"]["best"], info))
    out = {
        "metric": METRIC,
        "value": 1e3 / ms_step32,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": steps32,
        "warmup": w32,
        "ms_per_step": ms_step32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.only == "fp64" else "f32-spmv/f64-ir",
        "data": "synthetic",
        "config": {
            "workload": (f"Biot u-p {'Q1/Q1 hexa8' if q1 else 'hexa20/8'} box {n}^3 cells "
                         f"({solver.neq} equations"
                         + (", the metric's 8.6M-dof size" if n == (128 if q1 else 87) else "")
                         + f"), heterogeneous log-k, 10% patch load, dt 0.1 s, SQMR+Jacobi "
                         f"1e-10, {'FP64' if args.only == 'fp64' else 'FP32+IR'}"),
            "n": n, "element": args.element, "numel": deck.numel, "nnode": solver.nnode,
            "neq": solver.neq,
            "nnz": info["nnz"], "dt": deck.dt, "solver": "sqmr", "precond": "diag",
            "tol": 1e-10, "setup_s": round(setup_s, 2),
        },
        "requested": {"steps": args.steps, "warmup": args.warmup, "budget_s": budget},
        "fp32ir": None if args.only == "fp64" else {"time_steps_per_s": 1e3 / ms_step32, "ms_per_step": ms_step32,
                   "inner_iters_per_step": it32,
                   "krylov_iters_per_s": it32 / (ms_step32 / 1e3), "max_true_rel": rel32,
                   "max_rel_err_vs_fp64": {"u": err_u, "p": err_p},
                   "spmv_ms": ms32, "spmv_launches_timed": live32["fp32_n"],
                   "spmv_GBps_algorithmic": alg32 / ms32 / 1e6,
                   "spmv_GBps_compulsory": b32 / ms32 / 1e6},
        "fp64": fp64 if args.only != "fp64" else {
            "time_steps_per_s": 1e3 / ms_step32, "ms_per_step": ms_step32, "steps": steps32,
            "warmup": w32, "krylov_iters_per_step": it32,
            "krylov_iters_per_s": it32 / (ms_step32 / 1e3), "max_true_rel": rel32,
            "spmv_ms": ms32, "spmv_launches_timed": live32["fp32_n"],
            "spmv_GBps_algorithmic": alg32 / ms32 / 1e6, "spmv_GBps_compulsory": b32 / ms32 / 1e6},
        "speedup_fp32ir_vs_fp64": (fp64["ms_per_step"] / ms_step32) if fp64 else None,
        "spmv": {"tuned_shapes": shapes, "fp32_alg_bytes": alg32, "fp64_alg_bytes": alg64,
                 "fp32_compulsory_bytes": b32, "fp64_compulsory_bytes": b64},
        "roofline": {"bound": "hbm",
                     "kernel": (spmv_kernel_name("fp64", shapes["fp64"]["best"])
                                + " (FP64 node-block SpMV with the fused q.t dot)")
                     if args.only == "fp64" else
                     (spmv_kernel_name("fp32", shapes["fp32"]["best"])
                      + " (FP32 node-block SpMV with the fused q.t dot)"),
                     "achieved": alg32 / ms32 / 1e6, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": alg32 / ms32 / 1e6 / HBM_PEAK_GBPS,
                     "traffic": traffic32,
                     "algorithmic_bytes": alg32,
                     "compulsory_bytes": b32,
                     "achieved_compulsory": b32 / ms32 / 1e6,
                     "frac_compulsory": b32 / ms32 / 1e6 / HBM_PEAK_GBPS,
                     "achieved_traffic": (traffic32 / ms32 / 1e6) if traffic32 else None},
        "cpu_baseline": cpu,
    }
    # the configuration's own 128^3 mesh at its nominal 8.6M dofs (the Q1/Q1
    # element), one step of each form, when the budget still has room for it
    if not q1 and args.n is None and not args.no_q1 and args.only is None:
        if left() - margin >= args.q1_reserve:
            solver.close()
            out["q1q1_128cube"] = q1_nominal_leg(args, cfg32, cfg64, left, margin)
        else:
            log(f"Q1/Q1 128^3 leg skipped: {left():.0f} s left of the {budget:.0f} s budget")
    out["wall_s"] = time.perf_counter() - T_START
    print(json.dumps(out), flush=True)


def q1_nominal_leg(args, cfg32, cfg64, left, margin):
    """One FP32+IR and one FP64 step, each from the zero state, of the
    128^3-cell box with the equal-order Q1/Q1 element (8,586,756 dofs before
    restraints): the benchmark configuration's mesh at its nominal size."""
    from poromec_amd import problem as P
    from poromec_amd.solver import MefSolver
    t0 = time.perf_counter()
    d = P.heterogeneous_box_deck(128, seed=args.seed, dt=0.1, patch_fraction=0.1, nsteps=50,
                                 nen=8)
    s = MefSolver(d)
    info = s.matrix_info()
    s.pform(2, True, False, d.dt, d.dt)
    shapes = {name: s.tune_spmv(use32, reps=5)[0] for use32, name in ((False, "fp64"), (True, "fp32"))}
    res = {"workload": f"Biot u-p Q1/Q1 hexa8 box 128^3 cells ({s.neq} equations), "
                       "heterogeneous log-k, 10% patch load, dt 0.1 s, SQMR+Jacobi 1e-10",
           "neq": s.neq, "nnz": info["nnz"], "setup_s": time.perf_counter() - t0,
           "spmv_shapes": shapes}
    for name, cfg in (("fp32ir", cfg32), ("fp64", cfg64)):
        if name == "fp64" and left() - margin < 1.8 * res["fp32ir"]["s_per_step"]:
            log(f"Q1/Q1 128^3 FP64 step skipped: {left():.0f} s left")
            break
        z4, zn = np.zeros((s.nnode, 4)), np.zeros(s.nnode)
        s.set_state(u=z4, u0=z4, dp=zn, porosity=zn, t=0.0)
        s.spmv_timing(args.live_every)
        t1 = time.perf_counter()
        inf = s.step(cfg)
        wall = time.perf_counter() - t1
        live = s.spmv_timing(0)
        key = "fp32" if name == "fp32ir" else "fp64"
        ms = live[key + "_ms"] / max(live[key + "_n"], 1)
        comp = spmv_bytes(info, 4 if name == "fp32ir" else 8, fused_dot=True,
                          shape=shapes[key])
        res[name] = {"s_per_step": wall, "time_steps_per_s": 1.0 / wall,
                     "krylov_iters": sum(inf["iters"]), "max_true_rel": max(inf["true_rel"]),
                     "spmv_ms": ms, "spmv_GBps_compulsory": comp / ms / 1e6}
        log(f"  Q1/Q1 128^3 {name}: {wall:.1f} s, iters {inf['iters']}, SpMV {ms:.3f} ms")
    if "fp64" in res:
        res["speedup_fp32ir_vs_fp64"] = res["fp64"]["s_per_step"] / res["fp32ir"]["s_per_step"]
    s.close()
    return res


if __name__ == "__main__":
    main()
