"""shiftsSolver(acceptedEqns=...) at the boundary (CPU only: every case below
is decided before a plan is made).

The reference zeroes the weights of rejected equations in the minimisation
(zorro/zorro.py:1390, :1413) and then multiplies the weights of all M
equations by the errors of the Maccepted accepted ones (:1485-1487); numpy
cannot broadcast (M,) against (Maccepted,) unless the two agree or one is 1.
"""
import numpy as np
import pytest

from zorro_amd.registrator import ImageRegistrator


def _trimats(n=6):
    corr = np.triu(np.full((n, n), 0.5, np.float32), 1)
    shifts = np.zeros((n, n, 2), np.float32)
    peaksig = np.triu(np.full((n, n), 8.0, np.float32), 1)
    return shifts, corr, peaksig


def _reg(n=6):
    reg = ImageRegistrator()
    reg.images = np.zeros((n, 64, 64), np.float32)
    return reg


def test_rejected_equations_fail_like_the_reference():
    shifts, corr, peaksig = _trimats()
    m = 5 * 6 // 2                       # pairs of the 6 frames = equations
    acc = np.ones(m, bool)
    acc[[2, 7]] = False
    with pytest.raises(ValueError, match="broadcast"):
        _reg().shiftsSolver(shifts, corr, peaksig, acceptedEqns=acc)
    # numpy's (M,) x (0,) fails the same way
    with pytest.raises(ValueError, match="broadcast"):
        _reg().shiftsSolver(shifts, corr, peaksig, acceptedEqns=np.zeros(m, bool))


def test_single_accepted_equation_is_refused():
    shifts, corr, peaksig = _trimats()
    acc = np.zeros(15, bool)
    acc[4] = True
    with pytest.raises(NotImplementedError):
        _reg().shiftsSolver(shifts, corr, peaksig, acceptedEqns=acc)


def test_wrong_length_is_a_value_error():
    shifts, corr, peaksig = _trimats()
    with pytest.raises(ValueError, match="entries"):
        _reg().shiftsSolver(shifts, corr, peaksig, acceptedEqns=np.ones(14, bool))


def test_equation_count_follows_the_nonzero_correlations():
    # a pair with zero correlation is not an equation (:1282)
    shifts, corr, peaksig = _trimats()
    corr[0, 3] = 0.0
    with pytest.raises(ValueError, match="14 equations"):
        _reg().shiftsSolver(shifts, corr, peaksig, acceptedEqns=np.ones(15, bool))
