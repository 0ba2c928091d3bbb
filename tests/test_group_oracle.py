"""The oracle's group averaging (post_block, fishertrans, inv_fishertrans and
the threshold of the averaged matrix; W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:36-50,
548-558, 683-703, 824-887), pinned by closed forms, numpy and brute force on
small inputs.  CPU only."""
import math

import numpy as np
import pytest

import orc


def test_fisher_is_atanh():
    for r in np.linspace(-0.99, 0.99, 67, dtype=np.float32):
        assert orc.fisher(float(r)) == pytest.approx(math.atanh(float(r)), rel=2e-6, abs=1e-7)


def test_fisher_at_one_moves_below_one():
    # r == 1 becomes the float nearest 1 - 1e-6 (CorMat.cpp:39); the result is
    # finite and equals atanh of that float up to the float quotient's rounding
    r1 = float(np.float32(1.0 - 1e-6))
    z = orc.fisher(1.0)
    assert math.isfinite(z)
    assert z == pytest.approx(math.atanh(r1), rel=1e-6)
    assert orc.fisher(-1.0) == -math.inf


def test_inverse_round_trip():
    for r in np.linspace(-0.95, 0.95, 39, dtype=np.float32):
        assert orc.inv_fisher(orc.fisher(float(r))) == pytest.approx(float(r), abs=2e-6)
    assert orc.inv_fisher(-math.inf) == -1.0


def _subjects(S, N, seed):
    rng = np.random.default_rng(seed)
    M1 = N * (N - 1) // 2
    return [np.tanh(rng.normal(0, 0.5, M1)) for _ in range(S)]


def test_mean_matches_numpy():
    N = 60
    subs = _subjects(5, N, 1)
    avg = orc.group_average(subs, N, fisher=False)
    ref = np.mean(np.stack([s.astype(np.float32).astype(np.float64) for s in subs]), axis=0)
    assert np.max(np.abs(avg - ref)) < 1e-6


def test_fisher_mean_matches_numpy():
    N = 60
    subs = _subjects(4, N, 2)
    avg = orc.group_average(subs, N, fisher=True)
    ref = np.tanh(np.mean(np.arctanh(np.stack([s.astype(np.float32).astype(np.float64) for s in subs])), axis=0))
    assert np.max(np.abs(avg - ref)) < 5e-6


def test_two_equal_subjects_give_the_subject_exactly():
    N = 40
    s = _subjects(1, N, 3)[0]
    avg = orc.group_average([s, s], N, fisher=False)
    assert np.array_equal(avg, s.astype(np.float32))


def test_nan_and_out_of_range_values_add_nothing():
    N = 30
    a, b = _subjects(2, N, 4)
    b = b.copy()
    b[::7] = np.nan
    b[3::11] = 1.5
    b[5::13] = -1.25
    avg = orc.group_average([a, b], N, fisher=False)
    skip = np.isnan(b) | (b > 1) | (b < -1)
    assert np.array_equal(avg[skip], (a[skip].astype(np.float32) / np.float32(2)))
    both = (a.astype(np.float32) + b.astype(np.float32)) / np.float32(2)
    assert np.array_equal(avg[~skip], both[~skip])


def _brute_threshold(packed, N, thr):
    A = np.zeros((N, N))
    Vm = np.zeros((N, N))
    t = 0
    for i in range(N):
        for j in range(i + 1, N):
            if packed[t] > thr:
                A[i, j] = A[j, i] = 1
                Vm[i, j] = Vm[j, i] = packed[t]
            t += 1
    R = [0]
    C, V = [], []
    for i in range(N):
        for j in range(N):
            if A[i, j]:
                C.append(j)
                V.append(Vm[i, j])
        R.append(len(C))
    return np.array(R), np.array(C), np.array(V)


def test_threshold_packed_matches_brute_force():
    N = 45
    avg = orc.group_average(_subjects(3, N, 5), N, fisher=True)
    og = orc.threshold_packed(avg, N, 0.3)
    R, C, V = _brute_threshold(avg.astype(np.float64), N, 0.3)
    assert np.array_equal(og.R, R)
    assert np.array_equal(og.C, C)
    assert np.array_equal(og.V64, V)
