"""The Kokkos-semantics oracle against what the mathematics fixes."""
import math

import numpy as np

from oracle import kokkos_oracle as ko
from oracle import vlasov_oracle as vo


def _product_basis(x):
    """Lagrange basis at nodes -2..3 from its definition."""
    nodes = np.arange(-2, 4, dtype=np.float64)
    out = []
    for j in range(6):
        b = np.ones_like(x)
        for m in range(6):
            if m != j:
                b = b * (x - nodes[m]) / (nodes[j] - nodes[m])
        out.append(b)
    return out


def test_horner_basis_is_the_lagrange_basis():
    x = np.linspace(0.0, 1.0, 1001, endpoint=False)
    for a, b in zip(ko.kokkos_weights(x), _product_basis(x)):
        assert np.max(np.abs(a - b)) <= 1e-15
    w = ko.kokkos_weights(np.array([0.0]))
    assert [float(v[0]) for v in w] == [0.0, 0.0, 1.0, 0.0, 0.0, 0.0]


def test_horner_basis_is_exact_on_quintics_and_sums_to_one():
    x = np.linspace(0.0, 1.0, 97, endpoint=False)
    w = ko.kokkos_weights(x)
    assert np.max(np.abs(sum(w) - 1.0)) <= 1e-15
    nodes = np.arange(-2, 4, dtype=np.float64)
    for p in range(6):
        approx = sum(w[k] * nodes[k] ** p for k in range(6))
        assert np.max(np.abs(approx - x ** p)) <= 1e-13 * max(1.0, 3.0 ** p)


def test_whole_cell_shifts_are_exact_rolls():
    rng = np.random.default_rng(0)
    n, L = 32, 8.0
    dx = L / n
    grid = np.arange(n) * dx
    rows = rng.random((5, n))
    for cells in (1, 3, -2, -7, 40, -45):  # also beyond one box length
        vel = np.full(5, cells * dx / 0.5)  # vel * dt = cells * dx exactly (dx a power of two)
        out = ko.advect_rows_kokkos(rows, vel, 0.5, grid, L, 0.0, 1.0 / dx, True)
        np.testing.assert_array_equal(out, np.roll(rows, cells, axis=1))


def test_dirichlet_shifts_zero_fill_and_leave_the_box():
    rng = np.random.default_rng(1)
    n, vmax = 32, 4.0
    dv = 2 * vmax / n
    grid = -vmax + np.arange(n) * dv
    rows = rng.random((3, n))
    out = ko.advect_rows_kokkos(rows, np.full(3, 2 * dv), 1.0, grid, 2 * vmax, -vmax, 1.0 / dv, False)
    ref = np.zeros_like(rows)
    ref[:, 2:] = rows[:, :-2]
    np.testing.assert_array_equal(out, ref)
    out = ko.advect_rows_kokkos(rows, np.full(3, -3 * vmax), 1.0, grid, 2 * vmax, -vmax, 1.0 / dv, False)
    assert np.count_nonzero(out) == 0


def test_agrees_with_the_pytorch_form_on_smooth_periodic_data():
    n, L = 64, 4 * math.pi
    dx = L / n
    grid = np.arange(n) * dx
    f = np.sin(0.5 * grid)[None, :] + 0.3 * np.cos(1.5 * grid)[None, :] * np.ones((4, 1))
    vel = np.array([0.37, -1.3, 2.9, -0.05])
    dt = 0.1
    a = ko.advect_rows_kokkos(f, vel, dt, grid, L, 0.0, 1.0 / dx, True)
    b = vo.advect_rows(f, vel, dt, dx, periodic=True)
    assert np.max(np.abs(a - b)) <= 1e-13
