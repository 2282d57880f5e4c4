"""CPU pins of the oracle's principal-symbol adjoint placement (no GPU).

* 1-D d'Alembert closed forms of the reference's pairing diagnostic
  (examples/diagnostics/msgb_adjoint_pairing.py:186-199): P f = f(c t) / 2 and
  P* h = window h / (2 c) on the same grid.  This pins the adjoint's start
  time t_start = k_t / (rho bl_t), its amplitude (the boundary amplitude,
  B^-1, the c^2 source weight and the c^-2 image weight) and the forward's
  amplitude and positions; a plausible slip -- rho dropped from t_start or
  from the amplitude -- is planted in the oracle and must fail the check.
* 2-D four-edge composition: the oracle's four single-face adjoints (corner
  sensors weighted 1/2, column faces solved with the spatial axes exchanged)
  back-project the forward data of a two-packet p0 onto p0 itself
  (cosine > 0.9).  This pins the tangential placement x_T = k / (rho bl) and
  the axis exchange; a dropped rho in x_T is planted and must fail.
"""

import numpy as np
import pytest

from adjoint_cases import (DALEMBERT_ADJ, DALEMBERT_FWD, Dalembert1D, FourEdge2D, best_fit,
                           cpp_field_sum, dalembert_ok)
from oracle import adjoint as oadj


@pytest.fixture(scope="module")
def dalembert_default():
    case = Dalembert1D()
    return case, case.oracle_forward()


@pytest.mark.parametrize("c,length", [(1.0, 1.0), (1500.0, 0.0128)])
def test_dalembert_forward_closed_form(c, length):
    case = Dalembert1D(c=c, length=length)
    pf = case.oracle_forward()
    fit = best_fit(case.forward_closed_form, pf)
    assert dalembert_ok(fit, DALEMBERT_FWD), fit


@pytest.mark.parametrize("c,length", [(1.0, 1.0), (1500.0, 0.0128)])
def test_dalembert_adjoint_closed_form(c, length):
    case = Dalembert1D(c=c, length=length)
    y = case.oracle_forward()
    q, _ = case.oracle_adjoint(y)
    fit = best_fit(case.adjoint_closed_form(y), q)
    assert dalembert_ok(fit, DALEMBERT_ADJ), fit


def test_dalembert_detects_rho_dropped_from_start_time(dalembert_default, monkeypatch):
    case, y = dalembert_default
    monkeypatch.setattr(oadj, "start_time", lambda k, support: k / (support / 2))
    q, _ = case.oracle_adjoint(y)
    fit = best_fit(case.adjoint_closed_form(y), q)
    assert not dalembert_ok(fit, DALEMBERT_ADJ), fit
    assert fit[2] < 0.5


def test_dalembert_detects_rho_dropped_from_amplitude(dalembert_default, monkeypatch):
    case, y = dalembert_default
    amp = oadj.boundary_amplitude
    monkeypatch.setattr(oadj, "boundary_amplitude",
                        lambda L, support, N, sigma: amp(L, support / 2, N, sigma))
    q, _ = case.oracle_adjoint(y)
    fit = best_fit(case.adjoint_closed_form(y), q)
    assert not dalembert_ok(fit, DALEMBERT_ADJ), fit


def test_dalembert_detects_a_half_sample_shift(dalembert_default, monkeypatch):
    """The time index offset by half a sample: t_start = (k + 1/2) / (rho bl)."""
    case, y = dalembert_default
    monkeypatch.setattr(oadj, "start_time", lambda k, support: (k + 0.5) / support)
    q, _ = case.oracle_adjoint(y)
    assert not dalembert_ok(best_fit(case.adjoint_closed_form(y), q), DALEMBERT_ADJ)


def test_four_edge_faces_cover_each_sensor_once():
    """Weights over the four faces add to one per sensor; each face is one
    edge in increasing tangential order."""
    N = (6, 9)
    faces = oadj.four_edge_faces(N, (1.0, 1.0))
    S = 2 * N[1] + 2 * (N[0] - 2)
    tot = np.zeros(S)
    for f in faces:
        tot[f["members"]] += f["weights"]
    assert np.allclose(tot, 1.0)
    mask = np.zeros(N, dtype=bool)
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = True
    rc = np.argwhere(mask)
    assert np.array_equal(rc[faces[0]["members"]], np.stack([np.zeros(9), np.arange(9)], 1))
    assert np.array_equal(rc[faces[3]["members"]], np.stack([np.arange(6), np.full(6, 8)], 1))
    assert [len(f["members"]) for f in faces] == [9, 9, 6, 6]


def test_fourier_crop_keeps_a_band_limited_signal():
    """The crop resamples a band-limited record: a slow sinusoid comes back
    on the coarse grid, scaled by sqrt(Nt / nt_data) (the example's unitary
    FFT pair)."""
    nt, ntd = 150, 64
    t = np.arange(nt) / nt
    x = np.cos(2 * np.pi * 3 * t)[:, None] * np.ones((1, 5))
    got = oadj.fourier_crop_time(x, ntd)
    td = np.arange(ntd) / ntd
    ref = np.sqrt(nt / ntd) * np.cos(2 * np.pi * 3 * td)[:, None] * np.ones((1, 5))
    assert np.max(np.abs(got - ref)) < 1e-12


@pytest.fixture(scope="module")
def four_edge_case():
    return FourEdge2D(N=(64, 64), seed=0)


def _cos(a, b):
    return float(np.vdot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)))


def test_four_edge_adjoint_back_projects_onto_p0(four_edge_case):
    case = four_edge_case
    img, info = oadj.four_edge_adjoint(case.data, case.N, case.dx, case.c, case.ts, 128, 4000,
                                       field_sum=cpp_field_sum)
    assert _cos(img, case.p0) > 0.9
    # each face alone sees part of the wave field only
    assert all(len(b[0]) == 4000 for b in info)


def test_four_edge_adjoint_detects_rho_dropped_from_x_T(four_edge_case, monkeypatch):
    case = four_edge_case
    monkeypatch.setattr(oadj, "detector_coordinate", lambda k, support: k / (support / 2))
    img, _ = oadj.four_edge_adjoint(case.data, case.N, case.dx, case.c, case.ts, 128, 4000,
                                    field_sum=cpp_field_sum)
    assert _cos(img, case.p0) < 0.5


def test_four_edge_adjoint_detects_unexchanged_column_faces(four_edge_case, monkeypatch):
    """Column faces solved without the axis exchange (their image not
    transposed back) no longer back-project onto p0."""
    case = four_edge_case
    faces = oadj.four_edge_faces

    def no_swap(N, dx):
        out = faces(N, dx)
        for f in out[2:]:
            f["swap"] = False
        return out
    monkeypatch.setattr(oadj, "four_edge_faces", no_swap)
    img, _ = oadj.four_edge_adjoint(case.data, case.N, case.dx, case.c, case.ts, 128, 4000,
                                    field_sum=cpp_field_sum)
    assert _cos(img, case.p0) < 0.8
