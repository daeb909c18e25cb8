"""CPU checks of the exact form's number format (tests/oz_model.py): digit ranges, exact
representation, truncation bounds, and the accuracy class of a product."""
import numpy as np

import oz_model as M


def test_digits_range_and_reconstruction():
    g = np.random.default_rng(1)
    x = g.standard_normal(20000) * np.exp2(g.integers(-30, 30, 20000))
    E = M.line_exps(x.astype(complex)[None, :], 1)[0]
    d = M.digits(x, E)
    assert d[0].min() >= -65 and d[0].max() <= 65
    for k in (1, 2, 3):
        assert d[k].min() >= -127 and d[k].max() <= 127
    # integer digits of the rounded value: the reconstruction is exactly rint(x 2^(30 - E))
    v = ((d[0] * 255 + d[1]) * 255 + d[2]) * 255 + d[3]
    assert np.array_equal(v, np.rint(np.ldexp(x, 30 - int(E))).astype(np.int64))
    # truncation: at most half a unit of the last digit, 2^(E-31)
    err = np.abs(M.value(d, E) - x)
    assert err.max() <= np.ldexp(1.0, int(E) - 31)
    # values already on the grid are exact, including the line's extremes
    y = np.ldexp(np.array([-(2 ** 30), 2 ** 30 - 1, 12345, -1, 0], float), int(E) - 30)
    assert np.array_equal(M.value(M.digits(y, E), E), y)


def test_digit_split_against_brute_force():
    # every v in a window around each power of 255 and the extremes splits into the unique balanced
    # representation (checked by plain division with remainder in [-127, 127])
    vs = np.concatenate([np.arange(c - 300, c + 300) for c in (0, 255, 255 ** 2, 255 ** 3, 2 ** 30 - 300)])
    vs = np.concatenate([vs, -vs])
    d = M.digits(np.ldexp(vs.astype(float), -30), 0)
    for v, d0, d1, d2, d3 in zip(vs.tolist(), *[x.tolist() for x in d]):
        r = []
        w = v
        for _ in range(3):
            m = ((w + 127) % 255) - 127
            r.append(m)
            w = (w - m) // 255
        assert [d3, d2, d1, d0] == r + [w], v


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
        assert err < 3e-8, (D, err)
        # dropping order 3 is a measurably worse format (a planted truncation is caught)
        C2 = M.product(A, B, M.line_exps(A, 1), M.line_exps(B, 0), maxo=2)
        assert np.abs(C2 - R).max() / np.abs(R).max() > 10 * err


def test_round_trip_bound():
    g = np.random.default_rng(3)
    C = (g.standard_normal((64, 256)) + 1j * g.standard_normal((64, 256))) * np.exp2(g.integers(-20, 20, (64, 1)))
    R = M.round_trip(C)
    rmax = np.maximum(np.abs(C.real), np.abs(C.imag)).max(1, keepdims=True)
    assert (np.abs(R.real - C.real) <= 2.0 ** -30 * rmax).all()
    assert (np.abs(R.imag - C.imag) <= 2.0 ** -30 * rmax).all()
