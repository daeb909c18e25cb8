"""CPU checks of the exact form's number format (tests/oz_model.py): digit ranges, exact
representation, truncation bounds, and the accuracy class of a product."""
import numpy as np

import oz_model as M


def test_digits_range_and_reconstruction():
    g = np.random.default_rng(1)
    x = g.standard_normal(20000) * np.exp2(g.integers(-30, 30, 20000))
    E = M.line_exps(x.astype(complex)[None, :], 1)[0]
    d = M.digits(x, E)
    assert d[0].min() >= -64 and d[0].max() <= 64
    for k in (1, 2, 3):
        assert d[k].min() >= -64 and d[k].max() <= 63
    # truncation: at most half a unit of the last digit, 2^(E-28)
    err = np.abs(M.value(d, E) - x)
    assert err.max() <= np.ldexp(1.0, int(E) - 28)
    # values already on the grid are exact, including the line's extremes
    y = np.ldexp(np.array([-(2 ** 27), 2 ** 27 - 1, 12345, -1, 0], float), int(E) - 27)
    assert np.array_equal(M.value(M.digits(y, E), E), y)


def test_exponent_of_zero_line_and_line_bound():
    z = np.zeros((3, 5), complex)
    z[1, 2] = 3.0 - 0.5j
    E = M.line_exps(z, 1)
    assert E[0] == 0 and E[2] == 0 and E[1] == 2  # 3 < 2^2
    assert (np.abs(z.real).max(1) < np.exp2(E.astype(float))).all()


def test_product_matches_fp64_to_the_format_accuracy():
    g = np.random.default_rng(2)
    for D in (32, 256):
        A = g.standard_normal((D, D)) + 1j * g.standard_normal((D, D))
        B = g.standard_normal((D, D)) + 1j * g.standard_normal((D, D))
        R = A @ B
        C = M.product(A, B, M.line_exps(A, 1), M.line_exps(B, 0))
        err = np.abs(C - R).max() / np.abs(R).max()
        assert err < 4e-8, (D, err)
        # dropping order 4 is a measurably worse format (a planted truncation is caught)
        C3 = M.product(A, B, M.line_exps(A, 1), M.line_exps(B, 0), maxo=3)
        assert np.abs(C3 - R).max() / np.abs(R).max() > 2 * err


def test_round_trip_bound():
    g = np.random.default_rng(3)
    C = (g.standard_normal((64, 256)) + 1j * g.standard_normal((64, 256))) * np.exp2(g.integers(-20, 20, (64, 1)))
    R = M.round_trip(C)
    rmax = np.maximum(np.abs(C.real), np.abs(C.imag)).max(1, keepdims=True)
    assert (np.abs(R.real - C.real) <= 2.0 ** -27 * rmax).all()
    assert (np.abs(R.imag - C.imag) <= 2.0 ** -27 * rmax).all()
