"""Host-side checks of the serving engine (no GPU): request validation."""
import types

import numpy as np

from sparsesolv_amd.serve import rhs_error


def _matrix(n, dtype):
    return types.SimpleNamespace(n=n, dtype=dtype)


def test_rhs_of_the_right_size_and_type_is_accepted():
    assert rhs_error(np.zeros(5), _matrix(5, np.float64)) is None
    assert rhs_error(np.zeros(5, np.float32), _matrix(5, np.float64)) is None     # converted, not refused
    assert rhs_error(np.zeros(5), _matrix(5, np.complex128)) is None             # real rhs, complex system
    assert rhs_error(np.zeros(5, np.complex64), _matrix(5, np.complex128)) is None


def test_rhs_of_another_length_or_shape_is_refused():
    assert "shape" in rhs_error(np.zeros(4), _matrix(5, np.float64))
    assert "shape" in rhs_error(np.zeros(10), _matrix(5, np.float64))
    assert "shape" in rhs_error(np.zeros((5, 1)), _matrix(5, np.float64))


def test_complex_rhs_for_a_real_system_is_refused():
    assert "complex" in rhs_error(np.zeros(5, np.complex128), _matrix(5, np.float64))


def test_non_numeric_rhs_is_refused():
    assert "numeric" in rhs_error(np.array(["a"] * 5), _matrix(5, np.float64))
