"""Command-line driver shared by the two mirror modules: the flags, the
warm-up / timed-run sequence and the log file of the reference scripts'
``__main__`` blocks (vlasov1D1V.py and vlasov2D2V.py), so the reference's
log parsers (``Grid: ...``, ``Steps: N``, ``Elapsed time: X [s]``) read
these logs unchanged.  torch.compile / dynamo reports have no counterpart
here (the kernels are ahead-of-time HIP code)."""
from __future__ import annotations

import argparse
import math
import pathlib
import time


def parse(argv, two_d: bool):
    ap = argparse.ArgumentParser(add_help=True)
    if two_d:
        for axis in ("nx", "ny", "nvx", "nvy"):
            ap.add_argument(f"-{axis}", nargs="?", type=int, default=32)
        ap.add_argument("-lx", nargs="?", type=float, default=4.0 * math.pi)
        ap.add_argument("-ly", nargs="?", type=float, default=4.0 * math.pi)
        ap.add_argument("-Vx_max", nargs="?", type=float, default=5.0)
        ap.add_argument("-Vy_max", nargs="?", type=float, default=5.0)
        ap.add_argument("-nbiter", nargs="?", type=int, default=10)
    else:
        ap.add_argument("-nx", nargs="?", type=int, default=128)
        ap.add_argument("-nv", nargs="?", type=int, default=128)
        ap.add_argument("-lx", nargs="?", type=float, default=4.0 * math.pi)
        ap.add_argument("-vmax", nargs="?", type=float, default=5.0)
        ap.add_argument("-nbiter", nargs="?", type=int, default=1000)
    ap.add_argument("-diag_steps", nargs="?", type=int, default=10)
    ap.add_argument("-dt", nargs="?", type=float, default=0.05)
    ap.add_argument("-out_dir", nargs="?", type=str, default="data_python")
    ap.add_argument("-physics_mode", action="store_true", default=False)
    ap.add_argument("-epsilon", nargs="?", type=float, default=0.001)
    ap.add_argument("-dtype", type=str, default="float64", choices=["float32", "float64"])
    ap.add_argument("-solver", nargs="?", type=int, default=0, choices=[0, 1, 2])
    return ap.parse_args(argv)


def run(solver, log_name: str, memory) -> str:
    """Warm-up run, fresh initial state, timed run; appends the reference's
    log lines and writes the log to ``log_name`` in the working directory.
    Returns the log text."""
    get_memory_stats, reset_peak_memory, format_memory = memory
    state = solver.initialize()
    solver._append_log("Warming up (kernels are compiled ahead of time)...")
    solver.run(state)
    solver._append_log("Warm-up done. Starting timed run...")
    state = solver.initialize()
    reset_peak_memory(solver.device)
    before = get_memory_stats(solver.device)
    start = time.time()
    solver.run(state)
    elapsed = time.time() - start
    after = get_memory_stats(solver.device)
    solver._append_log(f"Elapsed time: {elapsed:.4f} [s]")
    if before is not None:
        msg = "\nMemory statistics\n"
        msg += f"Allocated before benchmark : {format_memory(before['allocated'])}\n"
        msg += f"Reserved before benchmark  : {format_memory(before['reserved'])}\n"
        msg += f"Peak allocated             : {format_memory(after['peak_allocated'])}\n"
        msg += f"Peak reserved              : {format_memory(after['peak_reserved'])}\n"
        solver._append_log(msg)
    pathlib.Path(log_name).write_text(solver.output_log)
    return solver.output_log
