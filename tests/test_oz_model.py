"""CPU checks of the exact form's number format (tests/oz_model.py): digit ranges, exact
representation, truncation bounds, and the accuracy class of a product."""
import numpy as np

import oz_model as M


def test_digits_range_and_reconstruction():
    g = np.random.default_rng(1)
    x = g.standard_normal(20000) * np.exp2(g.integers(-30, 30, 20000))
    E = M.line_exps(x.astype(complex)[None, :], 1)[0]
    d = M.digits(x, E)
    for k in range(4):
        assert d[k].min() >= -127 and d[k].max() <= 127
    # integer digits of the rounded value: the reconstruction is exactly rint(2^(30 - E) x SIGMA),
    # here from exact rationals
    from fractions import Fraction
    v = ((d[0] * 255 + d[1]) * 255 + d[2]) * 255 + d[3]
    for i in range(0, x.size, 97):
        q = Fraction(float(x[i])) * Fraction(63, 32) * Fraction(2) ** (30 - int(E))
        fl, rem = divmod(q, 1)
        r = fl + (1 if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1) else 0)
        assert v[i] == r, i
    # truncation: at most half a unit of the last digit, 2^(E-31) / SIGMA (and the one rounding of
    # the decode)
    err = np.abs(M.value(d, E) - x)
    assert (err <= np.ldexp(1.0, int(E) - 31) / M.SIGMA + np.abs(x) * 2.0 ** -52).all()
    # values already on the grid are exact, including the line's extremes
    vmax = int(M.SIGMA * 2 ** 30)
    y = np.ldexp(np.array([-vmax, vmax, vmax - 1, 2 ** 30, 12345, -1, 0], float), int(E) - 30) / M.SIGMA
    assert np.array_equal(M.value(M.digits(y, E), E), y)
    # the leading digit reaches the whole int8 range (a plain 2^30 scale stops at 65)
    assert np.array_equal(M.digits(y, E)[0][:2], [-127, 127])


def test_digit_split_against_brute_force():
    # every v in a window around each power of 255 and the extremes splits into the unique balanced
    # representation (checked by plain division with remainder in [-127, 127])
    vmax = int(M.SIGMA * 2 ** 30)
    vs = np.concatenate([np.arange(c - 300, c + 300) for c in (0, 255, 255 ** 2, 255 ** 3, 2 ** 30, vmax - 300)])
    vs = np.concatenate([vs, -vs])
    d = M.digits(np.ldexp(vs.astype(float), -30) / M.SIGMA, 0)
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
        assert err < 1e-8, (D, err)
        # dropping order 3 is a measurably worse format (a planted truncation is caught)
        C2 = M.product(A, B, M.line_exps(A, 1), M.line_exps(B, 0), maxo=2)
        assert np.abs(C2 - R).max() / np.abs(R).max() > 10 * err


def test_round_trip_bound():
    g = np.random.default_rng(3)
    C = (g.standard_normal((64, 256)) + 1j * g.standard_normal((64, 256))) * np.exp2(g.integers(-20, 20, (64, 1)))
    R = M.round_trip(C)
    rmax = np.maximum(np.abs(C.real), np.abs(C.imag)).max(1, keepdims=True)
    bound = 2.0 ** -30 / M.SIGMA * rmax * (1 + 2.0 ** -40)
    assert (np.abs(R.real - C.real) <= bound).all()
    assert (np.abs(R.imag - C.imag) <= bound).all()


def test_half_rows_round_trip():
    # rows whose halves have unequal exponents (the right half scaled down by 2^0 .. 2^9 per row):
    # the smaller half under the row's exponent after one more rounding, v -> rint(v 2^-k)
    from fractions import Fraction
    g = np.random.default_rng(4)
    n = 256
    C = g.standard_normal((64, 2 * n)) + 1j * g.standard_normal((64, 2 * n))
    C[:, n:] *= np.exp2(-g.integers(0, 10, (64, 1)))
    C[5, n:] = 0  # an all-zero half
    R = M.half_rows_round_trip(C)
    E = M.line_exps(C, 1)
    unit = np.ldexp(1.0, E - 30)[:, None] / M.SIGMA
    Eh = M.line_exps(C[:, n:], 1)
    same = Eh == E
    assert same.any() and (~same).any()
    # a row whose halves share the exponent is the plain round trip
    assert np.array_equal(R[same], M.round_trip(C)[same])
    # otherwise within 1/2 + 1/4 of a unit of the row's scale (the second rounding at k >= 1)
    for part in ("real", "imag"):
        err = np.abs(getattr(R, part) - getattr(C, part))
        assert (err <= 0.75 * unit * (1 + 2.0 ** -40)).all()
    # against exact rationals: v = rint(x SIGMA 2^(30 - E_h)), then rint(v 2^(E_h - E)), ties to even
    def rne(q):
        fl, rem = divmod(q, 1)
        return fl + (1 if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1) else 0)
    for r in np.nonzero(~same)[0][:8]:
        for j in range(n, 2 * n, 37):
            x = float(C[r, j].real)
            v = rne(Fraction(x) * Fraction(63, 32) * Fraction(2) ** (30 - int(Eh[r])))
            v2 = rne(v * Fraction(2) ** (int(Eh[r]) - int(E[r])))
            assert R[r, j].real == float(v2 * Fraction(2) ** (int(E[r]) - 30) / Fraction(63, 32)), (r, j)


def test_cf_rows_against_rationals():
    # the closing-factor form: rint(x 2^(30 - E)) ties to even under the row exponent, and with
    # halves the smaller half through rint(v 2^(E_h - E)); checked with exact rationals, and the
    # ties of the second rounding go to even (a planted truncation would fail both)
    from fractions import Fraction
    g = np.random.default_rng(5)
    n = 64
    C = g.standard_normal((16, 2 * n)) + 1j * g.standard_normal((16, 2 * n))
    C[:, n:] *= np.exp2(-g.integers(0, 6, (16, 1)))
    def rne(q):
        fl, rem = divmod(q, 1)
        return fl + (1 if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1) else 0)
    re, im, E = M.cf_rows(C)
    for r in range(16):
        for j in range(0, 2 * n, 7):
            assert re[r, j] == rne(Fraction(float(C[r, j].real)) * Fraction(2) ** (30 - int(E[r])))
    re2, im2, E2 = M.cf_rows(C, halves=True)
    assert np.array_equal(E, E2)
    Eh = M.line_exps(C[:, n:], 1)
    for r in range(16):
        for j in range(n, 2 * n, 5):
            v = rne(Fraction(float(C[r, j].imag)) * Fraction(2) ** (30 - int(Eh[r])))
            assert im2[r, j] == rne(v * Fraction(2) ** (int(Eh[r]) - int(E[r]))), (r, j)
    # rows whose halves share the exponent are the plain form
    same = Eh == E
    assert same.any() and (~same).any()
    assert np.array_equal(re2[same], re[same]) and np.array_equal(im2[same], im[same])
    # halves of an exact tie: 2^k v + 2^(k-1) goes to the even neighbour
    v = np.array([[1.0, 3.0, 5.0, -3.0]])
    assert np.array_equal(np.rint(np.ldexp(v, -1)), [[0.0, 2.0, 2.0, -2.0]])
