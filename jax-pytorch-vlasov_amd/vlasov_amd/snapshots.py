"""Raw binary snapshots in the layout of the reference's Kokkos solver, and
a restart reader.

The Kokkos drivers dump every diagnostic as headerless C-order arrays
(``to_binary_file``, simulations/vlasov2d_2v/kokkos/src/vlasov_poisson.hpp:
935-955; ``IO::to_binary``, common/kokkos/io_utils.hpp:123-135), one file per
variable: ``fn_SSSSSS.bin`` (f), ``f0_SSSSSS.bin``, ``rho_SSSSSS.bin``,
``phi_SSSSSS.bin`` with the iteration zero-filled to six digits.  The
reference's 1D post-processing reads that layout back
(``collect_binary_files`` / ``read_binary``,
simulations/vlasov1d_1v/postscript/postscript.py:104-180), inferring the
element type from the file size.  Files written here follow it exactly, so
those scripts read them unchanged; ``read_snapshot`` restores a state from
them (or from the Kokkos solver's own dumps).

The density files hold the Kokkos solver's quantity, the full density
rho = sum_v f dv (``compute_rho``, simulations/vlasov1d_1v/kokkos/src/
vlasov_poisson.hpp:418), not the PyTorch drivers' delta_rho = rho - 1:
the drivers write ``kokkos_rho(delta_rho)`` and convert back on restart with
``delta_rho_of(rho)``.
"""
from __future__ import annotations

import pathlib

import numpy as np
import torch

VARIABLES = ("fn", "f0", "rho", "phi")


def kokkos_rho(delta_rho):
    """Full density rho = delta_rho + 1, the Kokkos snapshots' rho."""
    return delta_rho + 1.0


def delta_rho_of(rho):
    """delta_rho = rho - 1 of a Kokkos-convention density."""
    return rho - 1.0


def snapshot_path(out_dir, name: str, step: int, suffix: str = "bin") -> pathlib.Path:
    return pathlib.Path(out_dir) / f"{name}_{int(step):06d}.{suffix}"


def write_snapshot(out_dir, step: int, fn, f0, rho, phi) -> list:
    """Write the four arrays (tensors or arrays) as raw C-order files; returns
    the paths written.  rho is the full density (``kokkos_rho``)."""
    out = pathlib.Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    paths = []
    for name, value in zip(VARIABLES, (fn, f0, rho, phi)):
        arr = value.detach().cpu().contiguous().numpy() if isinstance(value, torch.Tensor) else np.asarray(value)
        path = snapshot_path(out, name, step)
        np.ascontiguousarray(arr).tofile(path)
        paths.append(path)
    return paths


def _dtype_from_size(path: pathlib.Path, count: int):
    size = path.stat().st_size
    for dt in (np.float64, np.float32):
        if size == count * np.dtype(dt).itemsize:
            return np.dtype(dt)
    raise ValueError(f"{path} holds {size} bytes, which is not {count} float64 or float32 values")


def read_snapshot(out_dir, step: int, shapes: dict, dtype=None) -> dict:
    """Arrays of one snapshot.  ``shapes`` maps each of fn, f0, rho, phi to
    its shape; the element type comes from the file size (as the reference's
    reader infers it) unless ``dtype`` is given.  A file whose size does not
    fit its shape raises ValueError."""
    out = {}
    for name in VARIABLES:
        path = snapshot_path(out_dir, name, step)
        count = int(np.prod(shapes[name]))
        dt = np.dtype(dtype) if dtype is not None else _dtype_from_size(path, count)
        arr = np.fromfile(path, dtype=dt)
        if arr.size != count:
            raise ValueError(f"{path} holds {arr.size} elements, but shape {tuple(shapes[name])} was expected")
        out[name] = arr.reshape(shapes[name])
    return out


def snapshot_steps(out_dir) -> list:
    """Iterations for which every variable has a file, sorted."""
    out = pathlib.Path(out_dir)
    sets = []
    for name in VARIABLES:
        sets.append({int(p.stem.rsplit("_", 1)[1]) for p in out.glob(f"{name}_*.bin")})
    return sorted(set.intersection(*sets)) if sets else []
