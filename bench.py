#!/usr/bin/env python3
"""Benchmark of the mef Biot u-p path on one MI355X.

Workload: a box of n^3 hex8 cells promoted to the reference's hexa20/8
element, with a seeded log-normal permeability field (sigma 1, correlation
length 8 cells), E ~ U[5, 50] MPa and nu ~ U[0.25, 0.35] per element, a seeded
10 % patch of 1 kPa top load, dt = 0.1 s, SQMR with Jacobi to 1e-10.  A "step"
is one backward-Euler time step of the solv macro.  The default n = 87 gives
8.67M equations, the 8.6M-dof size the metric is quoted at (with the
reference's element a 128^3-cell box has 27.6M equations; `--n 128` runs it).

The FP64 form and the FP32+IR form (FP32 matrix values and Jacobi inside an
FP64 iterative-refinement loop) each run W untimed steps, then exactly K
timed steps; `value` is the FP32+IR form's time steps per second.  The SpMV
kernel time is measured with HIP events on the library's stream.  The CPU
baseline is the oracle (a scalar C restatement of the reference) timed on a
small box of the same workload on this host and scaled to the benchmark box.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "time-steps/s at 8.6M-dof Biot u-p (FP32+IR vs FP64 @1e-10); SpMV GB/s"
# (lanes per node block, columns in flight per lane, index prefetch / runs)
# of the SpMV shapes 0..14
shape_lanes = (64, 64, 64, 64, 64, 32, 64, 64, 64, 64, 64, 64, 64, 64, 64)
shape_unr = (1, 2, 3, 4, 5, 4, 3, 4, 5, 1, 2, 1, 2, 5, 1)
shape_pf = (0, 0, 0, 0, 0, 0, 1, 1, 1, 2, 2, 3, 3, 4, 5)


def spmv_kernel_name(prec, shape):
    """The SpMV kernel the solver launches at this shape, with its fused dot."""
    vt = "float" if prec == "fp32" else "double"
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
    runs = shape is not None and shape_pf[shape] in (2, 3, 5)
    cols = 4 * info["nruns"] if runs else 4 * info["ncols"]
    meta = (32 if runs else 24) * info["nblocks"]
    return value_bytes * info["nnz"] + cols + meta + (24 if fused_dot else 16) * info["neq"]


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


def run_form(solver, cfg, warmup, steps, keep_states=False, live_every=16):
    from poromec_amd.solver import MefSolver  # noqa: F401

    z4 = np.zeros((solver.nnode, 4))
    zn = np.zeros(solver.nnode)
    solver.set_state(u=z4, u0=z4, dp=zn, porosity=zn, t=0.0)
    states, infos = [], []
    for k in range(warmup):
        info = solver.step(cfg)
        log(f"  {cfg.precision} warmup step {k}: {info['ms_total']:.0f} ms, iters {info['iters']}")
        if keep_states is not None:
            st = solver.state()
            states.append((st["u"].copy(), st["dp"].copy()))
    solver.spmv_timing(live_every)  # HIP events around every k-th SpMV of the timed steps
    t0 = time.perf_counter()
    for k in range(steps):
        info = solver.step(cfg)
        infos.append(info)
        log(f"  {cfg.precision} timed step {k}: {info['ms_total']:.0f} ms, iters {info['iters']}")
        if keep_states is not None:
            # state readback happens after the step returned; excluded below
            pass
    wall = time.perf_counter() - t0
    spmv = solver.spmv_timing(0)
    if keep_states is not None and steps:
        st = solver.state()
        states.append((st["u"].copy(), st["dp"].copy()))
    return wall, infos, states, spmv


def cpu_baseline(n_sample, n_bench_numel, bench_nnz, gpu_iters_per_step, seed=0):
    """The oracle timed on an n_sample^3 box of the same generator; scaled to
    the benchmark box by solver work (iterations x stored entries) and
    element work (elements x element passes)."""
    from oracle import oracle as O
    from poromec_amd import problem as P

    deck = P.heterogeneous_box_deck(n_sample, seed=seed, dt=0.1, patch_fraction=0.1)
    m = O.OracleModel(deck)
    nnz_full = 2 * m.st["nad"] + m.st["neq"]
    t0 = time.perf_counter()
    m.assemble(isw=2, lhs=True, rhs=True)
    t_elem = time.perf_counter() - t0
    t0 = time.perf_counter()
    info = m.step()
    t_step = time.perf_counter() - t0
    iters = sum(info["iters"])
    passes = 3  # isw 4, isw 2 (matrix), isw 2 (residual check); porosity is lighter
    t_solve = max(t_step - passes * t_elem, 1e-9)
    per_iter_nnz = t_solve / max(iters, 1) / nnz_full
    per_elem_pass = t_elem / deck.numel
    est = per_iter_nnz * bench_nnz * gpu_iters_per_step + per_elem_pass * n_bench_numel * passes
    return {
        "value": 1.0 / est,
        "unit": "time-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle (scalar C restatement) on a {n_sample}^3 box of the same generator: "
                   f"one step took {t_step:.2f} s ({iters} SQMR iterations, {nnz_full} stored "
                   f"entries, element pass {t_elem:.2f} s); scaled to the benchmark box by "
                   f"iterations x entries ({gpu_iters_per_step} iterations/step measured on the "
                   f"GPU) and elements x 3 element passes"),
        "sample_seconds": t_step + t_elem,
    }


def batch_form(args):
    """The batching form: nbox independent boxes (n^3 cells, n ~ U[nmin, nmax])
    as one block-diagonal system, every box with its own stop tests; K steps
    of FP64 and of FP32+IR.  Prints one JSON line (box-steps/s)."""
    from poromec_amd import problem as P
    from poromec_amd.solver import MefSolver, SolverConfig

    decks = P.batch_decks(nbox=args.nbox, seed=args.seed, nmin=args.nmin, nmax=args.nmax)
    big, off = P.concat_decks(decks)
    t0 = time.perf_counter()
    s = MefSolver(big)
    s.batch_define(off)
    setup_s = time.perf_counter() - t0
    s.pform(2, True, False, big.dt, big.dt)
    shapes = {name: s.tune_spmv(use32, reps=5)[0] for use32, name in ((False, "fp64"), (True, "fp32"))}
    log(f"batch: {args.nbox} boxes, neq {s.neq}, setup {setup_s:.1f} s, shapes {shapes}")
    out = {}
    for name, cfg in (("fp64", SolverConfig(tol=1e-10)),
                      ("fp32ir", SolverConfig(tol=1e-10, precision="fp32ir"))):
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
        for d in decks[:k]:
            one = MefSolver(d)
            cfg = SolverConfig(tol=1e-10)
            for _ in range(args.warmup):
                one.step(cfg)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                one.step(cfg)
            wall += time.perf_counter() - t0
            one.close()
        seq = {"boxes_timed": k, "box_steps_per_s": k * args.steps / wall}
        log(f"  one box at a time (fp64, first {k} boxes): {seq}")
    print(json.dumps({"form": "batch", "nbox": args.nbox, "n_range": [args.nmin, args.nmax],
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

    n = args.n
    deck = P.heterogeneous_box_deck(n, seed=args.seed, dt=0.1, nsteps=args.nsteps,
                                    corr_len=min(8.0, n / 2))
    cfg = SolverConfig(tol=1e-13, maxit=200000)
    rng = np.random.default_rng(args.seed + 1)
    probe = MefSolver(deck)
    _, idm, _ = probe.numbering()
    free_p = np.nonzero(idm[: deck.nnodev, 3] > 0)[0]
    sens = [(int(v), 3) for v in rng.choice(free_p, size=min(args.nsens, len(free_p)), replace=False)]
    tdeck = P.heterogeneous_box_deck(n, seed=args.seed + 100, dt=0.1, nsteps=args.nsteps,
                                     corr_len=min(8.0, n / 2))
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
    g = np.concatenate([gE, gk])
    h, errs = args.fd_h, []
    for _ in range(args.ndir):
        v = rng.standard_normal(2 * len(lE))
        jp = m._evaluate(lE + h * v[: len(lE)], lk + h * v[len(lE):])[0]
        jm = m._evaluate(lE - h * v[: len(lE)], lk - h * v[len(lE):])[0]
        fd = (jp - jm) / (2 * h)
        errs.append(abs(fd - g @ v) / max(abs(fd), 1e-300))
        log(f"  direction {len(errs)}: fd {fd:.12e} adjoint {g @ v:.12e} rel {errs[-1]:.2e}")
    print(json.dumps({"form": "adjoint", "n": n, "numel": deck.numel, "neq": m.solver.neq,
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

    n = args.n
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

    n = args.n
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
    ap.add_argument("--n", type=int, default=87, help="box edge in cells")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--spmv-reps", type=int, default=20)
    ap.add_argument("--cpu-sample-n", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", choices=["precision", "batch", "adjoint", "tform", "vprop"],
                    default="precision",
                    help="precision (default, the metric's workload), batch or adjoint form")
    ap.add_argument("--nbox", type=int, default=256)
    ap.add_argument("--nmin", type=int, default=12)
    ap.add_argument("--nmax", type=int, default=32)
    ap.add_argument("--seq-sample", type=int, default=8,
                    help="batch form: boxes also solved one at a time, for comparison")
    ap.add_argument("--nsteps", type=int, default=50, help="adjoint form: march length")
    ap.add_argument("--nsens", type=int, default=64)
    ap.add_argument("--ckpt", type=int, default=10)
    ap.add_argument("--ndir", type=int, default=20)
    ap.add_argument("--fd-h", type=float, default=1e-6)
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

    n = args.n
    t0 = time.perf_counter()
    deck = P.heterogeneous_box_deck(n, seed=args.seed, dt=0.1, patch_fraction=0.1, nsteps=50)
    log(f"deck {n}^3: {deck.numel} elements, {deck.nnodev} vertices ({time.perf_counter()-t0:.1f} s)")
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
        print(json.dumps({"spmv_fp64_ms": ms64, "spmv_fp32_ms": ms32,
                          "spmv_fp64_bytes": spmv_bytes(info, 8, shape=shapes["fp64"]["best"]),
                          "spmv_fp32_bytes": spmv_bytes(info, 4, shape=shapes["fp32"]["best"]),
                          "copy_ms": msc,
                          "copy_bytes_read": copy_bytes, "copy_bytes_written": copy_bytes,
                          "copy_GBps": 2 * copy_bytes / msc / 1e6, "tuned_shapes": shapes,
                          **info}), flush=True)
        return

    # SpMV shape tuner on this matrix (lanes per node block x columns in flight)
    solver.pform(2, True, False, deck.dt, deck.dt)
    shapes = {}
    for use32, name in ((False, "fp64"), (True, "fp32")):
        best, ms = solver.tune_spmv(use32, reps=5)
        shapes[name] = {"best": best, "ms": [round(v, 4) for v in ms]}
        log(f"SpMV shapes {name}: " + ", ".join(f"{v:.3f}" for v in ms) + f" ms -> shape {best}")

    cfg64 = SolverConfig(precision="fp64", tol=1e-10)
    cfg32 = SolverConfig(precision="fp32ir", tol=1e-10, ir_inner_tol=1e-5)

    wall64, infos64, st64, live64 = run_form(solver, cfg64, args.warmup, args.steps,
                                             keep_states=True)
    wall32, infos32, st32, live32 = run_form(solver, cfg32, args.warmup, args.steps,
                                             keep_states=True)

    # error of the FP32+IR histories against the FP64 form, step by step
    nv = deck.nnodev
    err_u = err_p = 0.0
    for (u64, _), (u32, _) in zip(st64, st32):
        su = max(np.abs(u64[:, :3]).max(), 1e-300)
        sp = max(np.abs(u64[:nv, 3]).max(), 1e-300)
        err_u = max(err_u, np.abs(u32[:, :3] - u64[:, :3]).max() / su)
        err_p = max(err_p, np.abs(u32[:nv, 3] - u64[:nv, 3]).max() / sp)

    # SpMV kernel time inside the timed steps (HIP events on the library's
    # stream around every solver SpMV), per value precision
    b64 = spmv_bytes(info, 8, fused_dot=True, shape=shapes["fp64"]["best"])
    b32 = spmv_bytes(info, 4, fused_dot=True, shape=shapes["fp32"]["best"])
    ms64 = live64["fp64_ms"] / max(live64["fp64_n"], 1)
    ms32 = live32["fp32_ms"] / max(live32["fp32_n"], 1)
    gbps64, gbps32 = b64 / ms64 / 1e6, b32 / ms32 / 1e6
    log(f"SpMV in the timed steps (every 16th launch timed): fp64 {ms64:.3f} ms x "
        f"{live64['fp64_n']} ({gbps64:.0f} GB/s), fp32 {ms32:.3f} ms x {live32['fp32_n']} "
        f"({gbps32:.0f} GB/s)")
    iso64 = solver.bench_spmv(False, args.spmv_reps)
    iso32 = solver.bench_spmv(True, args.spmv_reps)
    log(f"SpMV back to back: fp64 {iso64:.3f} ms, fp32 {iso32:.3f} ms")

    steps = args.steps
    it64 = sum(sum(i["iters"]) for i in infos64) / max(steps, 1)
    it32 = sum(sum(i["iters"]) for i in infos32) / max(steps, 1)
    ms_step64 = 1e3 * wall64 / max(steps, 1)
    ms_step32 = 1e3 * wall32 / max(steps, 1)
    rel64 = max((r for i in infos64 for r in i["true_rel"]), default=0.0)
    rel32 = max((r for i in infos32 for r in i["true_rel"]), default=0.0)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.cpu_sample_n, deck.numel, info["nnz"], int(round(it64)))
            log(f"cpu baseline: {cpu['value']:.3e} steps/s ({cpu['sample_seconds']:.1f} s of CPU work)")
        except Exception as exc:  # the oracle is optional on the benchmark host
            log(f"cpu baseline skipped: {exc}")

    out = {
        "metric": METRIC,
        "value": 1e3 / ms_step32,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-spmv/f64-ir",
        "data": "synthetic",
        "config": {
            "workload": f"Biot u-p hexa20/8 box {n}^3, heterogeneous log-k, 10% patch load, "
                        f"SQMR+Jacobi 1e-10, FP32+IR",
            "n": n, "numel": deck.numel, "nnode": solver.nnode, "neq": solver.neq,
            "nnz": info["nnz"], "dt": deck.dt, "solver": "sqmr", "precond": "diag",
            "tol": 1e-10, "setup_s": round(setup_s, 2),
        },
        "fp64": {"time_steps_per_s": 1e3 / ms_step64, "ms_per_step": ms_step64,
                 "krylov_iters_per_step": it64,
                 "krylov_iters_per_s": it64 / (ms_step64 / 1e3), "max_true_rel": rel64},
        "fp32ir": {"time_steps_per_s": 1e3 / ms_step32, "ms_per_step": ms_step32,
                   "inner_iters_per_step": it32,
                   "krylov_iters_per_s": it32 / (ms_step32 / 1e3), "max_true_rel": rel32,
                   "max_rel_err_vs_fp64": {"u": err_u, "p": err_p}},
        "speedup_fp32ir_vs_fp64": ms_step64 / ms_step32,
        "spmv": {"fp64_ms": ms64, "fp32_ms": ms32, "fp64_GBps": gbps64, "fp32_GBps": gbps32,
                 "fp64_launches": live64["fp64_n"], "fp32_launches": live32["fp32_n"],
                 "fp64_bytes": b64, "fp32_bytes": b32,
                 "back_to_back_fp64_ms": iso64, "back_to_back_fp32_ms": iso32,
                 "tuned_shapes": shapes,
                 "csr_equivalent_fp64_GBps": csr_bytes(info) / ms64 / 1e6},
        "roofline": {"bound": "hbm",
                     "kernel": spmv_kernel_name("fp32", shapes["fp32"]["best"])
                               + " (FP32 node-block SpMV with the fused q.t dot)",
                     "achieved": gbps32, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": gbps32 / HBM_PEAK_GBPS,
                     "traffic": measured_traffic("fp32", shapes["fp32"]["best"], info)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
