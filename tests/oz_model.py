"""numpy model of the exact form's number format and product (test infrastructure).

A real value x of a stored line with exponent E (every |x| of the line below 2^E) is four
balanced base-255 digits: v = rint(2^(30-E) x SIGMA), v = D0 255^3 + D1 255^2 + D2 255 + D3
with every digit in [-127, 127] (SIGMA = 63/32, so |v| <= SIGMA 2^30 < (255^4 - 1) / 2). A product
sums the digit pairs of total order <= 3 in integers, forms S = ((acc_0 255 + acc_1) 255 + acc_2)
255 + acc_3 (exact), multiplies it by UNSCALE = rn(255^3 / SIGMA^2) in FP64 (one rounding) and
scales by 2^(E_A + E_B - 60). Every step is exact or one IEEE rounding, so the HIP kernel must
reproduce this model bit for bit.
"""
import numpy as np

NDIG = 4
MAXO = 3
BASE = 255
FRAC = 30  # bits of a line below its exponent, before the factor SIGMA
SIGMA = 1.96875  # 63/32: the leading digit spans [-127, 127]
UNSCALE = 16581375.0 * 1024.0 / 3969.0  # rn(255^3 / SIGMA^2)
_MIN = np.iinfo(np.int64).min
_MB = 8290000
_BIAS = 255 * _MB  # about 2^31: keeps every intermediate of the digit split non-negative


def exps_of(x):
    """frexp exponent of each nonzero value (|x| < 2^e), the int64 minimum for zeros."""
    _, e = np.frexp(x)
    return np.where(x == 0, _MIN, e.astype(np.int64))


def line_exps(z, axis):
    """Exponent of each line of a complex array: max over re and im along `axis` (0 if all zero)."""
    e = np.maximum(exps_of(z.real), exps_of(z.imag)).max(axis=axis)
    return np.where(e == _MIN, 0, e)


def digits(x, E):
    """Balanced base-255 digits of x under exponent E (broadcast), most significant first."""
    # the exact product x SIGMA 2^(30-E) (at most 59 significant bits: exact in the x86 extended
    # format) rounded once to the nearest integer, ties to even
    xl = np.asarray(x, dtype=np.longdouble)
    v = np.rint(np.ldexp(xl, FRAC - np.asarray(E)) * np.longdouble(SIGMA)).astype(np.int64)
    d = [None] * NDIG
    for k in range(NDIG - 1, 0, -1):
        u = v + 127 + _BIAS
        q = u // BASE
        d[k] = u - BASE * q - 127
        v = q - _MB
    d[0] = v
    return d


def value(d, E):
    """The value a digit set stands for."""
    v = ((d[0] * BASE + d[1]) * BASE + d[2]) * BASE + d[3]
    return np.ldexp(v.astype(np.float64), np.asarray(E) - FRAC) / SIGMA


def product(A, B, EA, EB, maxo=MAXO):
    """C = A B in the exact form: A [M,K] with row exponents EA [M], B [K,N] with column exponents
    EB [N] (complex128 inputs, digitised here)."""
    dar, dai = digits(A.real, EA[:, None]), digits(A.imag, EA[:, None])
    dbr, dbi = digits(B.real, EB[None, :]), digits(B.imag, EB[None, :])
    re = [0] * (MAXO + 1)
    im = [0] * (MAXO + 1)
    for a in range(NDIG):
        for b in range(NDIG):
            o = a + b
            if o > maxo:
                continue
            re[o] = re[o] + dar[a] @ dbr[b] - dai[a] @ dbi[b]
            im[o] = im[o] + dar[a] @ dbi[b] + dai[a] @ dbr[b]

    def comb(acc):
        acc = [np.asarray(x, dtype=np.int64) for x in acc]
        s = acc[0]
        for o in range(1, MAXO + 1):
            s = s * BASE + acc[o]
        return s.astype(np.float64) * UNSCALE  # s exact (|s| < 2^53), one rounding

    e = EA[:, None] + EB[None, :] - 2 * FRAC
    return np.ldexp(comb(re), e) + 1j * np.ldexp(comb(im), e)


def half_rows_round_trip(C):
    """C written back as a P = 128 product writes its digit output, decoded: each half of a row
    (columns [0, n/2) and [n/2, n)) digitised under its own exponent, then the half with the
    smaller one brought to the row's exponent E, v -> rint(v 2^-k) with k = E - E_half (ties to
    even, exact in FP64 for |v| < 2^31)."""
    n = C.shape[1] // 2
    E = line_exps(C, 1)[:, None]
    out = np.empty_like(C)
    for h in range(2):
        X = C[:, h * n:(h + 1) * n]
        Eh = line_exps(X, 1)[:, None]  # (0 for an all-zero half, whose digits are zero)
        k = E - Eh
        parts = []
        for x in (X.real, X.imag):
            d = digits(x, Eh)
            v = ((d[0] * BASE + d[1]) * BASE + d[2]) * BASE + d[3]
            v2 = np.rint(np.ldexp(v.astype(np.float64), -k)).astype(np.int64)
            parts.append(np.ldexp(v2.astype(np.float64), E - FRAC) / SIGMA)
        out[:, h * n:(h + 1) * n] = parts[0] + 1j * parts[1]
    return out


def cf_values(x, E):
    """Closing-factor form of x under exponent E (broadcast): rint(x 2^(30-E)), ties to even (the
    scaling is exact)."""
    return np.rint(np.ldexp(np.asarray(x, dtype=np.float64), FRAC - np.asarray(E))).astype(np.int64)


def cf_rows(C, halves=False):
    """C in the closing-factor form as a product writes it: one exponent per row (E, returned), and
    with halves (P = 128) each half of a row under its own exponent first, then brought to E by
    v -> rint(v 2^-k), k = E - E_half, ties to even. Returns (re ints, im ints, E)."""
    E = line_exps(C, 1)[:, None]
    if not halves:
        return cf_values(C.real, E), cf_values(C.imag, E), E[:, 0]
    n = C.shape[1] // 2
    re = np.empty(C.shape, np.int64)
    im = np.empty(C.shape, np.int64)
    for h in range(2):
        X = C[:, h * n:(h + 1) * n]
        Eh = line_exps(X, 1)[:, None]
        for out, x in ((re, X.real), (im, X.imag)):
            v = cf_values(x, Eh)
            out[:, h * n:(h + 1) * n] = np.rint(np.ldexp(v.astype(np.float64), Eh - E)).astype(np.int64)
    return re, im, E[:, 0]


def round_trip(C, axis=1):
    """C written back as digits with one exponent per line (rows: axis=1) and decoded."""
    E = line_exps(C, axis)
    E = E[:, None] if axis == 1 else E[None, :]
    return value(digits(C.real, E), E) + 1j * value(digits(C.imag, E), E)
