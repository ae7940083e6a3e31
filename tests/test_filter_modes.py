"""filterMode parsing of the boundary (CPU only, no library calls).

alignImageStack runs hotpixFilter when 'hot' occurs anywhere in filterMode
(zorro/zorro.py:1595-1596) and then applies the comma-separated tokens in
order (:1692-1711).  make_params keeps that order in zr_params.filters and
rejects tokens the reference would skip silently.
"""
import pytest

from zorro_amd import _lib as L
from zorro_amd.engine import make_params, parse_filter_mode


def _steps(p):
    return [p.filters[i] for i in range(p.n_filters)]


def test_tokens_in_order_with_case_and_spaces():
    p = make_params(filterMode="Dose, gaussLP", pixelsize=0.1)
    assert _steps(p) == [L.ZR_FILTER_DOSE, L.ZR_FILTER_GAUSSLP]
    assert p.do_dose == 1 and p.do_hot == 0


def test_dosenorm_is_its_own_step():
    p = make_params(filterMode="dosenorm", pixelsize=0.1)
    assert _steps(p) == [L.ZR_FILTER_DOSENORM] and p.do_dose == 1


def test_gausslp_alone_has_no_dose_step():
    p = make_params(filterMode="gausslp")
    assert _steps(p) == [L.ZR_FILTER_GAUSSLP] and p.do_dose == 0


def test_hot_anywhere_runs_the_outlier_filter():
    assert parse_filter_mode("hot,dose") == (["dose"], True)
    assert parse_filter_mode("dose, hotpix") == (["dose"], True)
    assert parse_filter_mode(None) == ([], False)
    assert parse_filter_mode("dose,") == (["dose"], False)


def test_background_is_a_step_and_unknown_tokens_raise():
    p = make_params(filterMode="dose,background", pixelsize=0.1)
    assert _steps(p) == [L.ZR_FILTER_DOSE, L.ZR_FILTER_BACKGROUND] and p.do_dose == 1
    assert make_params(filterMode="background").do_dose == 0
    with pytest.raises(ValueError):
        make_params(filterMode="dsoe", pixelsize=0.1)
    with pytest.raises(NotImplementedError):
        parse_filter_mode("dose,dose,gausslp,dosenorm,gausslp")
