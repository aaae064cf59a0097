"""Pins the CPU oracle (test infrastructure) before it is trusted as the checker.

Pins used:
  * the paper's worked examples and hand-derived known answers (golden JSON);
  * exhaustive small cases checked against invariants computed independently
    (pure Python, not the C checker): length, survivor multiset, in-place
    invariant, no removed item left;
  * agreement of two differently structured implementations (sequential
    canonical vs the T-slice Redzone_CPU restatement) for every T;
  * the O(n) comparator against numpy.delete (a library stable remove);
  * planted mistakes in the oracle must each be caught by these checks.
"""
import itertools
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")


def _cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def _invariants_py(A0, out, R):
    """Independent pure-Python invariant check; returns a list of failures."""
    n, k = len(A0), len(R)
    sz = n - k
    fails = []
    rs = set(int(r) for r in R)
    survivors = sorted(int(A0[i]) for i in range(n) if i not in rs)
    if sorted(int(v) for v in out[:sz]) != survivors:
        fails.append("multiset")
    for i in range(sz):
        if i not in rs and int(out[i]) != int(A0[i]):
            fails.append(f"moved {i}")
            break
    removed = set(int(A0[r]) for r in rs)
    if any(int(v) in removed for v in out[:sz]):
        fails.append("removed item left")
    return fails


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c["name"])
def test_known_answers_canonical(oracle, case):
    A = np.arange(case["n"], dtype=np.int32)
    R = np.array(case["R"], dtype=np.int32)
    assert oracle.canonical(A, R) == oracle.OK
    assert A[: case["n"] - len(case["R"])].tolist() == case["expect"]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c["name"])
@pytest.mark.parametrize("threads", [1, 2, 3, 5, 8])
def test_known_answers_redzone_cpu(oracle, case, threads):
    A = np.arange(case["n"], dtype=np.int32)
    R = np.array(case["R"], dtype=np.int32)
    st, _ = oracle.redzone_cpu(A, R, threads=threads)
    assert st == oracle.OK
    assert A[: case["n"] - len(case["R"])].tolist() == case["expect"]


def test_exhaustive_small(oracle):
    """Every ordered removal list on n <= 6 items (2,372 lists)."""
    count = 0
    for n in range(0, 7):
        for k in range(0, n + 1):
            for R in itertools.permutations(range(n), k):
                A0 = (np.arange(n, dtype=np.int32) * 7 + 3)
                Rn = np.array(R, dtype=np.int32)
                A1 = A0.copy()
                assert oracle.canonical(A1, Rn) == oracle.OK
                assert _invariants_py(A0, A1, Rn) == [], (n, R)
                for T in (1, 2, 3):
                    A2 = A0.copy()
                    st, _ = oracle.redzone_cpu(A2, Rn.copy(), threads=T)
                    assert st == oracle.OK
                    assert np.array_equal(A1[: n - k], A2[: n - k]), (n, R, T)
                count += 1
    assert count == 2372


@pytest.mark.parametrize("esz", [4, 8, 16])
def test_redzone_cpu_matches_canonical(oracle, esz):
    rng = np.random.default_rng(esz)
    for trial in range(40):
        n = int(rng.integers(1, 20000))
        k = int(rng.integers(0, n + 1)) if trial % 3 else int(rng.integers(0, max(1, n // 50)))
        R = rng.permutation(n)[:k].astype(np.int32)
        A0 = rng.integers(-2 ** 31, 2 ** 31, size=(n, esz // 4), dtype=np.int64).astype(np.int32)
        A1 = A0.copy()
        assert oracle.canonical(A1, R) == oracle.OK
        for T in (1, 4, 7, 33):
            A2 = A0.copy()
            st, _ = oracle.redzone_cpu(A2, R.copy(), threads=T)
            assert st == oracle.OK
            assert np.array_equal(A1[: n - k], A2[: n - k]), (n, k, T)
        bad, fb = oracle.check(A0, A1, R)
        assert bad == 0, (bad, fb)


def test_checker_matches_python_invariants(oracle):
    rng = np.random.default_rng(5)
    for trial in range(200):
        n = int(rng.integers(1, 60))
        k = int(rng.integers(0, n + 1))
        R = rng.permutation(n)[:k].astype(np.int32)
        A0 = np.arange(n, dtype=np.int32) * 3 + 1
        out = A0.copy()
        oracle.canonical(out, R)
        if trial % 2 and n - k >= 2:  # plant a defect in the output
            i, j = rng.choice(n - k, 2, replace=False)
            if out[i] != out[j]:
                out[i], out[j] = out[j], out[i]
        bad, _ = oracle.check(A0, out, R)
        assert (bad == 0) == (_invariants_py(A0, out, R) == []), (n, k, bad)


def test_remove_stable_matches_numpy_delete(oracle):
    rng = np.random.default_rng(9)
    for trial in range(60):
        n = int(rng.integers(1, 50000))
        k = int(rng.integers(0, n + 1))
        R = rng.permutation(n)[:k].astype(np.int32)
        A0 = rng.integers(0, 2 ** 31, n).astype(np.int32)
        want = np.delete(A0, R)
        for threads in (1, 3, 8):
            A = A0.copy()
            st, _ = oracle.remove_stable(A, R, threads=threads)
            assert st == oracle.OK
            assert np.array_equal(A[: n - k], want)
            bad, _ = oracle.check(A0, A, R, mode=1)
            assert bad == 0


def test_input_errors(oracle):
    A = np.arange(10, dtype=np.int32)
    assert oracle.canonical(A.copy(), np.array([3, 3], np.int32)) in (oracle.EDUP, oracle.OK)
    assert oracle.canonical(A.copy(), np.array([9, 9], np.int32)) == oracle.EDUP
    assert oracle.canonical(A.copy(), np.array([10], np.int32)) == oracle.ERANGE
    assert oracle.canonical(A.copy(), np.array([-1], np.int32)) == oracle.ERANGE
    bad, _ = oracle.check(A, A, np.array([1, 1], np.int32))
    assert bad == 8


@pytest.mark.parametrize("mutation", [1, 2, 3, 4])
def test_planted_mistakes_are_caught(oracle, mutation):
    """Each planted mistake fails the known answers or the invariants or the
    agreement with the Redzone_CPU restatement."""
    caught = False
    for case in _cases():
        A = np.arange(case["n"], dtype=np.int32)
        R = np.array(case["R"], dtype=np.int32)
        oracle.canonical(A, R, mutation=mutation)
        if A[: case["n"] - len(case["R"])].tolist() != case["expect"]:
            caught = True
    rng = np.random.default_rng(mutation)
    for trial in range(50):
        n = int(rng.integers(2, 3000))
        k = int(rng.integers(1, n))
        R = rng.permutation(n)[:k].astype(np.int32)
        A0 = np.arange(n, dtype=np.int32)
        A1 = A0.copy()
        oracle.canonical(A1, R, mutation=mutation)
        A2 = A0.copy()
        oracle.redzone_cpu(A2, R.copy(), threads=3)
        if not np.array_equal(A1[: n - k], A2[: n - k]) or oracle.check(A0, A1, R)[0] != 0:
            caught = True
    assert caught


def test_count_tail(oracle):
    R = np.array([9, 2, 8, 7, 0], np.int32)
    assert oracle.count_tail(R, 10) == 3
