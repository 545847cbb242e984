"""Shared test inputs and tolerances.

Tolerance: the reference's own rule, |a - b| <= tol * max(1, |a|, |b|) with
tol = 1e-12 (double) and 1e-5 (float)
(tests/rpp/cpu/single_thread/ops/cpu_typed_ft_ops_test_helper.hpp:33-36,
261-273).
"""
from __future__ import annotations

import numpy as np

TOL = {np.float64: 1e-12, np.float32: 1e-5}

# Shape sweep of the reference's numeric ft_mul tests
# (tests/rpp/cpu/single_thread/ops/ft_mul_tests.cpp:324-371).
SHAPES = [(1, 8), (2, 7), (2, 8), (2, 9), (3, 4), (3, 5), (3, 6), (3, 7), (4, 3), (4, 4), (4, 5),
          (5, 2), (5, 3), (5, 4), (5, 5), (6, 4), (7, 4), (16, 2), (16, 3), (17, 2), (256, 1)]

# Degree-window cases at depth 4 (tests/rpp/ft_degree_range_cases.hpp:19-91).
FT_MUL_CASES = [
    ("positive_minima", (0, 4), (1, 2), (1, 2)),
    ("disjoint_operands", (0, 4), (1, 1), (2, 2)),
    ("disjoint_operands_reversed", (0, 4), (2, 2), (1, 1)),
    ("lhs_unit_only", (0, 4), (0, 0), (2, 3)),
    ("rhs_unit_only", (0, 4), (2, 3), (0, 0)),
    ("both_unit_only", (0, 4), (0, 0), (0, 0)),
    ("below_product", (0, 1), (1, 2), (1, 2)),
    ("above_product", (3, 4), (1, 1), (1, 1)),
    ("product_beyond_depth", (0, 4), (3, 4), (2, 3)),
    ("unit_output_lhs_absent", (0, 0), (1, 2), (0, 2)),
    ("unit_output_rhs_absent", (0, 0), (0, 2), (1, 2)),
    ("interior_output", (1, 3), (1, 1), (1, 1)),
]

FT_FMA_CASES = [
    ("addend_below_product", (0, 4), (0, 1), (1, 2), (1, 2)),
    ("addend_above_product", (0, 4), (3, 4), (1, 1), (1, 1)),
    ("gap_between_addend_and_product", (0, 4), (0, 0), (2, 2), (2, 2)),
    ("disjoint_product_operands", (0, 4), (1, 1), (1, 1), (2, 2)),
    ("disjoint_product_operands_reversed", (0, 4), (1, 1), (2, 2), (1, 1)),
    ("lhs_unit_only", (0, 4), (1, 1), (0, 0), (2, 3)),
    ("rhs_unit_only", (0, 4), (1, 1), (2, 3), (0, 0)),
    ("both_unit_only", (0, 4), (0, 0), (0, 0), (0, 0)),
    ("below_all_contributions", (0, 1), (2, 3), (1, 2), (1, 2)),
    ("above_all_contributions", (3, 4), (1, 2), (1, 1), (1, 1)),
    ("only_addend_contributes", (0, 4), (1, 2), (3, 4), (2, 3)),
    ("unit_output_lhs_absent", (0, 0), (1, 2), (1, 2), (0, 2)),
    ("unit_output_rhs_absent", (0, 0), (1, 2), (0, 2), (1, 2)),
    ("unit_output_addend_only", (0, 0), (0, 0), (1, 2), (0, 2)),
    ("interior_output", (1, 3), (3, 3), (1, 1), (1, 1)),
]

FT_INPLACE_CASES = [
    ("full_views", (0, 4), (0, 4), (0, 4)),
    ("positive_minima", (1, 4), (1, 2), (1, 2)),
    ("disjoint_ranges", (2, 4), (1, 1), (0, 0)),
    ("unit_multiplier", (1, 4), (0, 0), (2, 2)),
    ("unit_destination", (0, 0), (0, 0), (0, 0)),
    ("unit_destination_missing_operands", (0, 0), (1, 2), (1, 2)),
    ("low_output_missing_addend", (0, 2), (0, 1), (3, 4)),
    ("high_output_missing_addend", (3, 4), (0, 1), (0, 2)),
    ("product_above_output", (0, 2), (3, 4), (3, 4)),
    ("product_below_output", (3, 4), (0, 1), (0, 1)),
    ("interior_output", (1, 3), (1, 1), (3, 3)),
    ("single_high_degree", (4, 4), (0, 1), (1, 2)),
]


# Range cases of the shuffle adjoints at depth 4 (out, op, arg)
# (tests/rpp/shuffle_adjoint_test_helper.hpp:19-36) and the two extra cases of
# the half-shuffle adjoints (half_shuffle_adjoint_test_helper.hpp:86-108).
SHUFFLE_ADJ_CASES = [
    ("full_views", (0, 4), (0, 4), (0, 4)),
    ("unit_operator", (0, 4), (0, 0), (1, 2)),
    ("positive_operator_minimum", (0, 4), (1, 2), (2, 3)),
    ("single_degree_operands", (0, 4), (1, 1), (3, 3)),
    ("operator_above_argument", (0, 4), (3, 4), (0, 2)),
    ("output_below_support", (0, 1), (0, 1), (3, 4)),
    ("output_above_support", (3, 4), (1, 2), (1, 2)),
    ("unit_argument", (0, 4), (0, 0), (0, 0)),
    ("unit_argument_missing_operator", (0, 4), (1, 2), (0, 0)),
    ("unit_output_present", (0, 0), (2, 2), (2, 2)),
    ("unit_output_absent", (0, 0), (1, 1), (2, 2)),
    ("interior_output", (1, 3), (0, 1), (2, 3)),
    ("single_high_output", (4, 4), (0, 0), (4, 4)),
    ("mixed_words", (1, 2), (1, 2), (2, 4)),
    ("unit_cotangent_only", (0, 4), (0, 4), (0, 0)),
]


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    scale = np.maximum(1.0, np.maximum(np.abs(a), np.abs(b)))
    return float(np.max(np.abs(a - b) / scale))


def assert_close(actual, expected, dtype, what=""):
    tol = TOL[np.dtype(dtype).type]
    err = rel_err(actual, expected)
    assert np.all(np.isfinite(np.asarray(actual)[np.isfinite(np.asarray(expected))])), f"{what}: non-finite"
    assert err <= tol, f"{what}: relative error {err:.3e} > {tol:.0e}"


def assert_exact(actual, expected, dtype, what=""):
    """Within the reference's tolerance, and equal bit for bit (NaN where the
    expected value is NaN)."""
    a, e = np.asarray(actual), np.asarray(expected)
    fa, fe = np.nan_to_num(a, nan=0.0), np.nan_to_num(e, nan=0.0)
    assert_close(fa, fe, dtype, what)
    assert np.array_equal(np.isnan(a), np.isnan(e)), f"{what}: NaN positions differ"
    assert np.array_equal(fa, fe), f"{what}: not bit-identical (max rel {rel_err(fa, fe):.3e})"


def degree_begin(width, depth):
    out = [0]
    p = 1
    for _ in range(depth + 1):
        out.append(out[-1] + p)
        p *= width
    return out


def random_batch(rng, n, dim, dtype, scale=1.0, unit=True):
    """iid U[-1,1] coefficients, scaled so products stay O(1); degree 0 set to 1."""
    x = rng.uniform(-1.0, 1.0, size=(n, dim)) * scale
    if unit:
        x[:, 0] = 1.0
    return x.astype(dtype)


def damped_batch(rng, n, width, depth, dtype):
    """U[-1,1] scaled by min(1, 16/sqrt(dim)) and 0.5^k per degree, the input
    distribution of the reference GPU tests (gpu_block_test_helper.cuh:107-157)."""
    db = degree_begin(width, depth)
    dim = db[-1]
    x = rng.uniform(-1.0, 1.0, size=(n, dim)) * min(1.0, 16.0 / np.sqrt(dim))
    for k in range(depth + 1):
        x[:, db[k]:db[k + 1]] *= 0.5 ** k
    return x.astype(dtype)


def window_mask(width, depth, window):
    db = degree_begin(width, depth)
    m = np.zeros(db[-1], dtype=bool)
    lo, hi = window
    if lo <= hi:
        m[db[lo]:db[hi + 1]] = True
    return m
