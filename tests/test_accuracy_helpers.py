"""CPU checks of the accuracy-report helpers."""
import numpy as np
import pytest

from accuracy import hausdorff


def test_hausdorff_point_sets():
    a = np.array([[0.0, 0, 0], [1, 0, 0]])
    b = np.array([[0.0, 0, 0], [1, 0, 0.5]])
    assert hausdorff(a, b) == pytest.approx(0.5)
    assert hausdorff(b, a) == pytest.approx(0.5)
    assert hausdorff(a, a) == 0.0
    # one-sided excess counts: an extra far point in b
    c = np.vstack([b, [[0.0, 3.0, 0.0]]])
    assert hausdorff(a, c) == pytest.approx(3.0)
    assert hausdorff(a, np.zeros((0, 3))) == float("inf")
