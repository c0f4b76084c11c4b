"""Raw binary snapshots (the Kokkos solver's layout) on the CPU: names,
headerless C-order content, element type from the file size, size checks.
The layout is the one the reference's readers parse
(simulations/vlasov1d_1v/postscript/postscript.py:104-180)."""
import numpy as np
import pytest

from vlasov_amd import snapshots


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_write_read_roundtrip(tmp_path, dtype):
    rng = np.random.default_rng(0)
    arrays = {"fn": rng.random((4, 3, 5, 6)).astype(dtype), "f0": rng.random((5, 6)).astype(dtype),
              "rho": rng.random((4, 3)).astype(dtype), "phi": rng.random((4, 3)).astype(dtype)}
    paths = snapshots.write_snapshot(tmp_path, 20, arrays["fn"], arrays["f0"], arrays["rho"], arrays["phi"])
    assert [p.name for p in paths] == ["fn_000020.bin", "f0_000020.bin", "rho_000020.bin", "phi_000020.bin"]
    # headerless, C order: exactly what numpy.fromfile(...).reshape(shape) reads
    for name, arr in arrays.items():
        raw = np.fromfile(tmp_path / f"{name}_000020.bin", dtype=dtype)
        assert raw.size == arr.size
        np.testing.assert_array_equal(raw.reshape(arr.shape), arr)
    back = snapshots.read_snapshot(tmp_path, 20, {k: v.shape for k, v in arrays.items()})
    for name, arr in arrays.items():
        assert back[name].dtype == np.dtype(dtype)
        np.testing.assert_array_equal(back[name], arr)
    assert snapshots.snapshot_steps(tmp_path) == [20]


def test_size_mismatch_raises(tmp_path):
    snapshots.write_snapshot(tmp_path, 3, np.zeros((2, 5)), np.zeros(5), np.zeros(2), np.zeros(2))
    with pytest.raises(ValueError):
        snapshots.read_snapshot(tmp_path, 3, {"fn": (2, 6), "f0": (5,), "rho": (2,), "phi": (2,)})
    with pytest.raises(ValueError):
        snapshots.read_snapshot(tmp_path, 3, {"fn": (2, 5), "f0": (5,), "rho": (2,), "phi": (2,)}, dtype=np.float32)


def test_steps_need_every_variable(tmp_path):
    snapshots.write_snapshot(tmp_path, 0, np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2))
    snapshots.write_snapshot(tmp_path, 10, np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2))
    (tmp_path / "phi_000010.bin").unlink()
    assert snapshots.snapshot_steps(tmp_path) == [0]


def test_density_files_hold_the_kokkos_full_density(tmp_path):
    """rho_*.bin holds rho = sum_v f dv (the Kokkos solver's compute_rho), not
    the PyTorch drivers' delta_rho = rho - 1: a uniform f of unit density
    gives a file of ones, and a hand-built Kokkos-convention file reads back
    as delta_rho = rho - 1."""
    nx, nv, vmax = 8, 16, 4.0
    dv = 2 * vmax / nv
    f = np.full((nx, nv), 1.0 / (nv * dv))
    delta = f.sum(axis=1) * dv - 1.0  # the drivers' convention
    snapshots.write_snapshot(tmp_path, 0, f, f[0], snapshots.kokkos_rho(delta), np.zeros(nx))
    raw = np.fromfile(tmp_path / "rho_000000.bin", dtype=np.float64)
    np.testing.assert_allclose(raw, 1.0, rtol=0, atol=1e-15)
    kokkos = 1.0 + 0.01 * np.cos(2 * np.pi * np.arange(nx) / nx)  # a Kokkos dump's full density
    kokkos.tofile(tmp_path / "rho_000007.bin")
    for name, arr in (("fn", f), ("f0", f[0]), ("phi", np.zeros(nx))):
        np.ascontiguousarray(arr).tofile(tmp_path / f"{name}_000007.bin")
    back = snapshots.read_snapshot(tmp_path, 7, {"fn": f.shape, "f0": (nv,), "rho": (nx,), "phi": (nx,)})
    np.testing.assert_array_equal(snapshots.delta_rho_of(back["rho"]), kokkos - 1.0)
