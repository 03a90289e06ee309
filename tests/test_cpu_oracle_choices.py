"""Pins the two discrete choices of the reference path that no closed form fixes:
the window of getAsymto_moments (FourierTransform.cpp:289-315) and the
causality patch of FourierTransform_ (FourierTransform.cpp:145-175).

Each case is constructed so that its outcome is known without the oracle, and
so that a one-line change of the rule (a dropped positivity test, <= for <,
the scan started at the first centre, the interpolation taken from the wrong
neighbour, +1e-10 for -1e-10) gives a different outcome."""
import numpy as np
import pytest

from oracle import oracle as O
from moment_cases import brute_scan, noisy_two_pole, scan_rule

PI = 3.1415926535897932384626


# ----------------------------------------------------------- window scan rule
def test_scan_first_strict_minimum():
    # two equal minima: the first stays (a <= comparison would move to index 3)
    assert O.select_window([3.0, 1.0, 2.0, 1.0, 4.0], [1, 1, 1, 1, 1]) == 1


def test_scan_skips_windows_failing_positivity():
    # the smallest spread (index 1) fails the positivity tests: the next admissible one wins
    assert O.select_window([3.0, 0.5, 1.0, 2.0, 4.0], [1, 0, 1, 1, 1]) == 2


def test_scan_last_centre_wins_when_nothing_beats_it():
    # the last centre is the start value even when it fails the tests itself: every
    # admissible spread is above it, so it stays (a scan started at centre 0, or an
    # argmin over admissible centres only, returns 2)
    assert O.select_window([3.0, 5.0, 2.5, 4.0, 2.0], [1, 1, 1, 1, 0]) == 4
    assert O.select_window([3.0, 5.0, 2.5, 4.0, 2.0], [1, 1, 1, 1, 1]) == 4


def test_scan_nothing_admissible():
    assert O.select_window([0.1, 0.2, 0.3], [0, 0, 0]) == 2


# ------------------------------------------------- windows of the whole fit
def test_windows_all_rejected_fall_back_to_last():
    """m1 = 1 - 1.2 < 0: every window fails the positivity test, so every moment takes the
    last centre, although the spreads have their minimum well inside the range."""
    nf, beta = 200, 50.0
    fw = noisy_two_pole(nf, beta, 1.2, 3.0, 1e-4, seed=1)
    var, ok, Nv = brute_scan(fw, beta)
    ncent = var.shape[1]
    assert ok.sum() == 0
    _, win = O.asymto_moments(fw, beta, return_windows=True)
    assert win == [ncent - 1 + Nv] * 4
    for q in range(4):  # without the positivity filter the choice would differ
        assert int(np.argmin(var[q])) < ncent - 1 - 5


def test_windows_positivity_filter_against_brute_force():
    """The lowest-spread windows fail the second positivity test (Vi m3 Vi - (Vi m2 Vi)^2
    not positive): the window is the first strict minimum among the admissible ones,
    found here by an independent lstsq scan; the minimum is clear (relative margin
    >= 1e-3 over the runner-up), so rounding cannot move it."""
    nf, beta = 200, 50.0
    fw = noisy_two_pole(nf, beta, 0.8, 1.0, 3e-4, seed=2)
    var, ok, Nv = brute_scan(fw, beta)
    assert 0 < ok.sum() < len(ok)
    _, win = O.asymto_moments(fw, beta, return_windows=True)
    for q in range(4):
        sel = scan_rule(var[q], ok)
        adm = np.sort(var[q][ok.astype(bool)])
        assert (adm[1] - adm[0]) / adm[0] > 1e-3
        assert win[q] == sel + Nv, (q, win, sel + Nv)
        # the unfiltered argmin is an inadmissible window elsewhere
        unf = int(np.argmin(var[q]))
        assert not ok[unf] and unf != sel
        # dropping the second test admits other windows and changes the choice
        m1ok = np.array([1] * len(ok))
        assert scan_rule(var[q], m1ok) != sel


# ----------------------------------------------------------- causality patch
def _gt(n_tau, d, values, seed=0):
    """G(tau_t) = U diag(values[t]) U^H, one seeded unitary U for all t."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return np.array([q @ np.diag(v) @ q.conj().T for v in values])


def test_patch_one_positive_interior_point():
    n_tau, d, beta = 10, 3, 5.0
    vals = [[-0.5 - 0.01 * t, -0.3, -0.2 - 0.02 * t] for t in range(n_tau + 1)]
    vals[4] = [0.1, 0.2, 0.05]  # every eigenvalue >= 0 at tau_4
    gt = _gt(n_tau, d, vals)
    out, check = O.causality_patch(gt, beta)
    assert check == 1
    # t = 4: interpolation between t-1 = 3 and t1 = 5 at tau_4, i.e. the midpoint
    assert np.max(np.abs(out[4] - 0.5 * (gt[3] + gt[5]))) < 1e-15
    others = [t for t in range(n_tau + 1) if t != 4]
    assert np.array_equal(out[others], gt[others])


def test_patch_two_positive_points():
    n_tau, d, beta = 12, 2, 4.0
    vals = [[-0.6, -0.4] for _ in range(n_tau + 1)]
    vals[5] = [0.3, 0.2]
    vals[6] = [0.9, 0.8]  # large: the first interpolant (toward t = 6) stays positive
    gt = _gt(n_tau, d, vals, seed=1)
    out, check = O.causality_patch(gt, beta)
    assert check == 1
    # t = 5: t1 = 6 gives (G4 + G6)/2 with eigenvalues (0.15, 0.2) >= 0, so t1 = 7:
    # G4 + (G7 - G4) (tau5 - tau4)/(tau7 - tau4) = G4 + (G7 - G4)/3 = G4 (constant) < 0
    p5 = gt[4] + (gt[7] - gt[4]) / 3.0
    assert np.max(np.abs(out[5] - p5)) < 1e-15
    # t = 6 (after 5 was patched): midpoint of the patched G5 and G7
    assert np.max(np.abs(out[6] - 0.5 * (p5 + gt[7]))) < 1e-15


def test_patch_no_negative_continuation():
    n_tau, d, beta = 8, 2, 3.0
    # G(tau_4) barely negative, G(tau_t >= 5) large and positive: every interpolant
    # from tau_4 (or from the -1e-10 I a patched point becomes) stays positive
    vals = [[-0.5, -0.2] for _ in range(n_tau + 1)]
    vals[4] = [-0.01, -0.01]
    for t in range(5, n_tau + 1):
        vals[t] = [10.0, 12.0]
    gt = _gt(n_tau, d, vals, seed=2)
    out, check = O.causality_patch(gt, beta)
    assert check == 1
    for t in range(5, n_tau + 1):
        assert np.array_equal(out[t], -1e-10 * np.eye(d))
    assert np.array_equal(out[:5], gt[:5])


def test_patch_endpoints_are_never_interpolated():
    n_tau, d, beta = 6, 2, 3.0
    vals = [[-0.5, -0.2] for _ in range(n_tau + 1)]
    vals[0] = [0.1, 0.2]
    gt = _gt(n_tau, d, vals, seed=3)
    out, check = O.causality_patch(gt, beta)
    assert check == 1
    assert np.array_equal(out[0], -1e-10 * np.eye(d))
    assert np.array_equal(out[1:], gt[1:])


def test_patch_leaves_causal_data_alone():
    n_tau, d, beta = 6, 3, 3.0
    vals = [[-0.5, -0.2, -1e-3] for _ in range(n_tau + 1)]
    gt = _gt(n_tau, d, vals, seed=4)
    out, check = O.causality_patch(gt, beta)
    assert check == 0 and np.array_equal(out, gt)
