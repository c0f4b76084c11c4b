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
