"""bench.py's wall-clock budget, without a device: the solver is replaced by a
stand-in whose steps advance a simulated clock by fixed durations, so the
admission rules of the default run (warmup and timed steps of each form,
the CPU leg's reserve, the Q1/Q1 128^3 leg) can be checked against the
budget for step times the benchmark box has shown and for slower ones.
The run must end inside the budget, time at least one FP32+IR step, and
report the steps it timed and the ones asked for.  Two steps are always run
when a warmup is asked for (the first warmup step, before any step time is
known, and one timed step), so steps longer than half the budget overrun it
by design; the default workload's steps take about a minute."""
import importlib.util
import json
import sys
import types
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


class Clock:
    def __init__(self):
        self.t = 0.0

    def perf_counter(self):
        return self.t


def _load_bench(clock):
    spec = importlib.util.spec_from_file_location("bench_budget_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.time = types.SimpleNamespace(perf_counter=clock.perf_counter)
    mod.T_START = clock.perf_counter()
    return mod


def _fake_solver_class(clock, seconds):
    """seconds(nen, precision) -> simulated duration of one step."""

    class FakeSolver:
        def __init__(self, deck):
            self.nen = deck.nen
            self.nnode = 64
            self.neq = 200
            self.timing = 0
            self.launches = {"fp32": 0, "fp64": 0}

        def matrix_info(self):
            return {"nnz": 40000, "ncols": 12000, "nruns": 3500, "nblocks": 64, "neq": self.neq}

        def pform(self, *args):
            pass

        def tune_spmv(self, use32, reps=5):
            return 14, [1.0] * 16

        def set_state(self, **kw):
            pass

        def step(self, cfg):
            clock.t += seconds(self.nen, cfg.precision)
            if self.timing:
                self.launches["fp32" if cfg.precision == "fp32ir" else "fp64"] += 100
            return {"iters": [1000], "true_rel": [3e-11], "nlit": 1}

        def state(self):
            return {"u": np.zeros((self.nnode, 4)), "dp": np.zeros(self.nnode)}

        def spmv_timing(self, every):
            self.timing = every
            out = {"fp32_ms": 2.0 * self.launches["fp32"], "fp32_n": self.launches["fp32"],
                   "fp64_ms": 3.0 * self.launches["fp64"], "fp64_n": self.launches["fp64"]}
            if every == 0:
                self.launches = {"fp32": 0, "fp64": 0}
            return out

        def close(self):
            pass

    return FakeSolver


def _run(monkeypatch, capsys, argv, seconds):
    import poromec_amd.problem as P
    import poromec_amd.solver as S
    clock = Clock()
    bench = _load_bench(clock)
    monkeypatch.setattr(S, "MefSolver", _fake_solver_class(clock, seconds))
    monkeypatch.setattr(P, "heterogeneous_box_deck",
                        lambda n, nen=20, **kw: types.SimpleNamespace(
                            n=n, nen=nen, numel=n ** 3, nnodev=(n + 1) ** 3, dt=0.1))
    monkeypatch.setattr(sys, "argv", ["bench.py", *argv])
    bench.main()
    line = capsys.readouterr().out.strip().splitlines()[-1]
    return json.loads(line), clock.t


def _box_seconds(nen, precision):
    # the benchmark box's step times: hexa20/8 87^3 and Q1/Q1 128^3
    if nen == 20:
        return 62.6 if precision == "fp32ir" else 85.0
    return 19.1 if precision == "fp32ir" else 27.1


def test_driver_command_fits_the_budget(monkeypatch, capsys):
    out, t = _run(monkeypatch, capsys, ["--gpus", "1", "--steps", "20", "--warmup", "5",
                                        "--no-cpu"], _box_seconds)
    assert t <= out["requested"]["budget_s"] == 450.0
    assert out["requested"] == {"steps": 20, "warmup": 5, "budget_s": 450.0}
    assert 1 <= out["steps"] <= 20 and 0 <= out["warmup"] <= 5
    assert out["fp64"] is not None and out["fp64"]["steps"] == out["steps"]
    assert out["value"] == pytest.approx(1.0 / 62.6)
    assert out["speedup_fp32ir_vs_fp64"] == pytest.approx(85.0 / 62.6)
    assert "q1q1_128cube" in out
    assert out["q1q1_128cube"]["fp64"]["s_per_step"] == pytest.approx(27.1)
    assert out["wall_s"] == pytest.approx(t)


def test_slow_steps_still_time_one_step_and_skip_the_rest(monkeypatch, capsys):
    out, t = _run(monkeypatch, capsys, ["--steps", "20", "--warmup", "5", "--no-cpu"],
                  lambda nen, prec: 400.0 if prec == "fp32ir" else 700.0)
    assert out["steps"] == 1 and out["warmup"] == 1
    assert out["fp64"] is None  # no room left for an FP64 step
    assert "q1q1_128cube" not in out
    assert t == pytest.approx(800.0)  # the two steps that always run


def test_no_q1_and_explicit_size_leave_the_leg_out(monkeypatch, capsys):
    out, _ = _run(monkeypatch, capsys, ["--no-cpu", "--no-q1"], _box_seconds)
    assert "q1q1_128cube" not in out
    out, _ = _run(monkeypatch, capsys, ["--no-cpu", "--n", "32"], _box_seconds)
    assert "q1q1_128cube" not in out and out["config"]["n"] == 32
