"""Pins the CPU oracle (oracle/bna_oracle.c) before anything is compared to it.

Each check is chosen so a plausible slip (a dropped term, a wrong index or
sign, a transposed operand) fails it: known answers held in the reference's own
EC test program, closed forms, brute force with numpy at small sizes, and the
literal restatement of the reference's tie loop against its closed form.
"""
import numpy as np
import pytest

import orc


def setup_module(_):
    assert orc.olib().orc_cpu_has_fma(), "oracle needs an FMA-capable host CPU"


def _csr_from_dense(A):
    N = A.shape[0]
    R = np.zeros(N + 1, np.int64)
    cols, vals = [], []
    for i in range(N):
        nz = np.nonzero(A[i])[0]
        cols.extend(nz.tolist())
        vals.extend(A[i, nz].tolist())
        R[i + 1] = R[i] + len(nz)
    return R, np.array(cols, np.int32), np.array(vals, np.float32)


# ----------------------------------------------------------------- z-norm
def test_znorm_closed_forms():
    rng = np.random.default_rng(1)
    L, N = 57, 40
    x = rng.normal(size=(L, N)).astype(np.float32) * 3 + 1
    z64, z32, mu, sd, valid = orc.znorm(x, 0)
    assert valid.all()
    xd = x.T.astype(np.float64)
    np.testing.assert_allclose(mu, xd.mean(axis=1), rtol=0, atol=1e-13)
    np.testing.assert_allclose(sd, np.linalg.norm(xd - xd.mean(axis=1, keepdims=True), axis=1), rtol=1e-13)
    np.testing.assert_allclose(z64.mean(axis=1), 0, atol=1e-15)
    np.testing.assert_allclose(np.linalg.norm(z64, axis=1), 1, atol=1e-14)
    assert np.array_equal(z32, z64.astype(np.float32))
    # z(2x) == z(x) bit for bit: doubling is exact through every step
    z64b, *_ = orc.znorm(2 * x, 0)
    assert np.array_equal(z64, z64b)
    # voxel-major input gives the same result
    z64c, *_ = orc.znorm(np.ascontiguousarray(x.T), 1)
    assert np.array_equal(z64, z64c)


def test_znorm_constant_and_nonfinite_voxels_are_invalid():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(30, 6)).astype(np.float32)
    x[:, 1] = 4.5  # constant series: norm 0 -> 0/0
    x[3, 4] = np.nan
    x[7, 5] = np.inf
    _, _, _, sd, valid = orc.znorm(x, 0)
    assert valid.tolist() == [1, 0, 1, 1, 0, 0]
    assert sd[1] == 0.0


# ---------------------------------------------------------------- FP64 r
def test_dot64_closed_forms():
    t = np.arange(240)
    a = np.sin(2 * np.pi * 3 * t / 240)
    b = np.cos(2 * np.pi * 3 * t / 240)
    za, zb = a / np.linalg.norm(a), b / np.linalg.norm(b)
    assert abs(orc.dot64(za, zb)) < 1e-15
    assert orc.dot64(za, za) == pytest.approx(1.0, abs=1e-15)
    assert orc.dot64(za, -za) == pytest.approx(-1.0, abs=1e-15)
    x = np.random.default_rng(3).normal(size=(1, 100)).astype(np.float32)
    y = (2 * x.astype(np.float64) + 1).astype(np.float32)
    z = orc.znorm(np.vstack([x, y]), 1)[0]
    assert orc.dot64(z[0], z[1]) == pytest.approx(1.0, abs=1e-7)


@pytest.mark.parametrize("L", [1, 2, 63, 64, 65, 200, 1200])
def test_dot64_matches_numpy(L):
    rng = np.random.default_rng(L)
    a, b = rng.normal(size=L), rng.normal(size=L)
    assert orc.dot64(a, b) == pytest.approx(float(np.dot(a, b)), rel=1e-13, abs=1e-13)
    # every element takes part once: a one-hot probe picks out b[l]
    for l in {0, L // 2, L - 1}:
        e = np.zeros(L)
        e[l] = 1.0
        assert orc.dot64(e, b) == b[l]


def test_pearson_against_numpy_corrcoef():
    x = np.random.default_rng(4).normal(size=(50, 25)).astype(np.float32)
    z64, *_ = orc.znorm(x, 0)
    cc = np.corrcoef(x.T.astype(np.float64))
    packed = orc.all_pairs(z64, np.ones(25, np.uint8))
    iu = np.triu_indices(25, 1)
    np.testing.assert_allclose(packed, cc[iu], atol=1e-14)


# -------------------------------------------------------- threshold graph
def test_threshold_graph_brute_force_and_invariants():
    from cuda_newest20160331_amd import synth_factor_model

    x = synth_factor_model(300, 120, 8, 0.7, 11)
    z64, _, _, _, valid = orc.znorm(x, 0)
    thr = 0.3
    g = orc.build_threshold(z64, valid, thr)
    cc = np.corrcoef(x.T.astype(np.float64))
    np.fill_diagonal(cc, -np.inf)
    assert np.min(np.abs(cc[np.isfinite(cc)] - thr)) > 1e-9  # no pair sits on the threshold
    A = cc > thr
    assert g.nnz == A.sum() and g.nnz % 2 == 0
    for i in range(300):
        row = g.C[g.R[i]:g.R[i + 1]]
        assert np.array_equal(row, np.nonzero(A[i])[0])
        np.testing.assert_allclose(g.V64[g.R[i]:g.R[i + 1]], cc[i, row], atol=1e-13)
    assert np.array_equal(g.V, g.V64.astype(np.float32))


def test_threshold_excludes_invalid_voxels_even_below_zero():
    x = np.random.default_rng(5).normal(size=(40, 10)).astype(np.float32)
    x[:, 3] = 1.0
    z64, _, _, _, valid = orc.znorm(x, 0)
    g = orc.build_threshold(z64, valid, -2.0)  # every valid pair is an edge
    assert g.R[4] - g.R[3] == 0
    assert g.nnz == 9 * 8
    assert 3 not in set(g.C.tolist())


# ----------------------------------------------------------------- density
def test_density_k_reference_arithmetic():
    # C2: N = 45000, 1% -> Clength = 20,249,550, k = 10,124,775
    assert orc.density_k(45000, 1.0) == 10124775
    for N, s in [(5000, 1.0), (45000, 0.5), (60000, 2.5), (123457, 0.37), (10, 50.0)]:
        M1 = N * (N - 1) // 2
        prod = np.float32(np.float32(2 * M1) * np.float32(s))  # long long * float in float
        cl = int(float(prod) / 100.0)
        cl += cl % 2
        assert orc.density_k(N, s) == cl // 2


def test_density_selection_literal_loop_equals_closed_form():
    rng = np.random.default_rng(6)
    N = 60
    M1 = N * (N - 1) // 2
    for trial in range(20):
        packed = rng.integers(0, 12, size=M1).astype(np.float64) / 11.0  # heavy ties
        packed[rng.integers(0, M1, size=30)] = np.nan
        nv = int(np.sum(~np.isnan(packed)))
        for k in [1, 7, nv // 3, nv // 2, nv]:
            a = orc.select_packed(packed, N, k, literal=True)
            b = orc.select_packed(packed, N, k, literal=False)
            assert a.nnz == b.nnz == 2 * k
            assert np.array_equal(a.R, b.R) and np.array_equal(a.C, b.C)
            vals = np.sort(packed[~np.isnan(packed)])[::-1]
            assert a.r_star == vals[k - 1]


def test_density_graph_full_sort_small():
    from cuda_newest20160331_amd import synth_factor_model

    x = synth_factor_model(120, 60, 8, 0.7, 3)
    z64, _, _, _, valid = orc.znorm(x, 0)
    k = orc.density_k(120, 5.0)
    g = orc.build_density(z64, valid, k)
    cc = np.corrcoef(x.T.astype(np.float64))
    iu = np.triu_indices(120, 1)
    order = np.argsort(-cc[iu], kind="stable")
    top = set(zip(iu[0][order[:k]].tolist(), iu[1][order[:k]].tolist()))
    got = set()
    for i in range(120):
        for j in g.C[g.R[i]:g.R[i + 1]]:
            if i < j:
                got.add((i, int(j)))
    assert got == top


# ------------------------------------------------------------------ degree
def test_degree():
    A = np.array([[0, 0.5, 0.25], [0.5, 0, 0], [0.25, 0, 0]], np.float64)
    R, Cc, V = _csr_from_dense(A)
    assert orc.degree(R, V, 3).tolist() == [0.75, 0.5, 0.25]
    assert orc.degree(R, None, 3).tolist() == [2.0, 1.0, 1.0]


# --------------------------------------------------------------------- EC
def test_ec_reference_known_answer_4x4():
    # B/src/cuda_EC/cuda_EC/ecGPU.cu:217-233: the 4x4 test matrix of the
    # reference's EC program (binary).  Leading eigenpair by numpy eigh:
    # lambda = 2.170086486626, v = (0.5227207, 0.2818452, 0.6116285, 0.5227207).
    R = np.array([0, 2, 3, 6, 8])
    Cc = np.array([2, 3, 2, 0, 1, 3, 0, 2], np.int32)
    y, lam, it = orc.eigcent(R, Cc, None, 4, out_mode=1)
    np.testing.assert_allclose(y, [0.5227207, 0.2818452, 0.6116285, 0.5227207], atol=2e-6)
    assert lam == pytest.approx(2.170086486626, abs=1e-5)
    x, lam2, it2 = orc.eigcent(R, Cc, None, 4, out_mode=0)
    np.testing.assert_allclose(x, [1.1343491, 0.6116283, 1.3272869, 1.1343491], atol=5e-6)
    assert it == it2 == 32
    A = np.zeros((4, 4))
    for i in range(4):
        A[i, Cc[R[i]:R[i + 1]]] = 1
    w, v = np.linalg.eigh(A)
    assert lam == pytest.approx(w[-1], abs=1e-5)


def test_ec_closed_forms_complete_and_cycle():
    n = 9
    A = np.ones((n, n)) - np.eye(n)
    R, Cc, V = _csr_from_dense(A)
    x, lam, it = orc.eigcent(R, Cc, V, n, out_mode=0)
    np.testing.assert_allclose(x, (n - 1) / np.sqrt(n), rtol=1e-6)
    assert lam == pytest.approx(n - 1, rel=1e-6)
    y, _, _ = orc.eigcent(R, Cc, None, n, out_mode=1)
    np.testing.assert_allclose(y, 1 / np.sqrt(n), rtol=1e-6)
    m = 11  # odd cycle: 2-regular, not bipartite
    A = np.zeros((m, m))
    for i in range(m):
        A[i, (i + 1) % m] = A[(i + 1) % m, i] = 1
    R, Cc, V = _csr_from_dense(A)
    x, lam, it = orc.eigcent(R, Cc, V, m, out_mode=0, max_iter=100000, tol=1e-9)
    np.testing.assert_allclose(x, 2 / np.sqrt(m), rtol=1e-5)


def test_ec_weighted_random_graph_vs_eigh():
    rng = np.random.default_rng(7)
    n = 80
    W = rng.random((n, n)) * (rng.random((n, n)) < 0.2)
    W = np.triu(W, 1)
    W = (W + W.T).astype(np.float32).astype(np.float64)
    R, Cc, V = _csr_from_dense(W)
    x, lam, it = orc.eigcent(R, Cc, V, n, out_mode=0)
    w, v = np.linalg.eigh(W)
    v1 = np.abs(v[:, -1])
    assert lam == pytest.approx(w[-1], rel=1e-6)
    np.testing.assert_allclose(x / np.linalg.norm(x), v1, atol=1e-5)


# ------------------------------------------------------- reference CPU path
def test_reference_cpu_path_close_to_fp64():
    from cuda_newest20160331_amd import synth_factor_model

    x = synth_factor_model(200, 150, 8, 0.7, 9)
    z64, z32, _, _, valid = orc.znorm(x, 0)
    r_ref = orc.ref_corr_rows(z32, 0, 200)
    r64 = orc.all_pairs(z64, valid)
    assert np.max(np.abs(r_ref.astype(np.float64) - r64)) < 2e-6
    # the threaded timing path counts the same edges as the packed values
    e, _ = orc.ref_cpu_rows(z32, 0, 200, 0.3, nthreads=3)
    assert e == int(np.sum(r_ref > np.float32(0.3)))


def test_reference_cpu_path_bit_exact_numpy_restatement():
    # CorMat.cpp:954-965: each product a[l] * b[l] rounded to float, summed in
    # double in l order, the sum stored as float.  numpy: float32 products, a
    # sequential float64 accumulation (add.accumulate keeps the order), a cast.
    # A float accumulator, a double product or another order all differ.
    from cuda_newest20160331_amd import synth_factor_model

    x = synth_factor_model(120, 97, 6, 0.7, 314)
    _, z32, _, _, _ = orc.znorm(x, 0)
    N, L = z32.shape
    got = orc.ref_corr_rows(z32, 3, 40)
    want = []
    for i in range(3, 40):
        p = (z32[i][None, :] * z32[i + 1:]).astype(np.float32)  # float products
        s = np.add.accumulate(p.astype(np.float64), axis=1)[:, -1]  # double, in l order
        want.append(s.astype(np.float32))
    want = np.concatenate(want)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the pin is sharp: a float accumulator differs somewhere
    wrong = np.concatenate([np.add.accumulate((z32[i][None, :] * z32[i + 1:]).astype(np.float32), axis=1)[:, -1]
                            for i in range(3, 40)])
    assert not np.array_equal(got, wrong)
