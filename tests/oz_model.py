"""numpy model of the exact form's number format and product (test infrastructure).

A real value x of a stored line with exponent E (every |x| of the line below 2^E) is four
balanced base-128 digits: v = rint(x 2^(27-E)), v = D0 128^3 + D1 128^2 + D2 128 + D3 with
D1..D3 in [-64, 63] and D0 in [-64, 64]. A product sums the digit pairs of total order <= 4 in
integers and scales the result by 2^(E_A + E_B - 40). Integer sums and power-of-two scaling are
exact, so the HIP kernel must reproduce this model bit for bit.
"""
import numpy as np

NDIG = 4
MAXO = 4
_MIN = np.iinfo(np.int64).min


def exps_of(x):
    """frexp exponent of each nonzero value (|x| < 2^e), the int64 minimum for zeros."""
    _, e = np.frexp(x)
    return np.where(x == 0, _MIN, e.astype(np.int64))


def line_exps(z, axis):
    """Exponent of each line of a complex array: max over re and im along `axis` (0 if all zero)."""
    e = np.maximum(exps_of(z.real), exps_of(z.imag)).max(axis=axis)
    return np.where(e == _MIN, 0, e)


def digits(x, E):
    """Balanced base-128 digits of x under exponent E (broadcast), most significant first."""
    v = np.rint(np.ldexp(x, 27 - np.asarray(E))).astype(np.int64)
    w = v + 1056832  # 64 (1 + 128 + 128^2)
    return [w >> 21, ((w >> 14) & 127) - 64, ((w >> 7) & 127) - 64, (w & 127) - 64]


def value(d, E):
    """The value a digit set stands for."""
    v = ((d[0] * 128 + d[1]) * 128 + d[2]) * 128 + d[3]
    return np.ldexp(v.astype(np.float64), np.asarray(E) - 27)


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
        hi = acc[0] * 128 + acc[1]
        mid = acc[2] * 128 + acc[3]
        return hi.astype(np.float64) * 2097152.0 + mid.astype(np.float64) * 128.0 + acc[4].astype(np.float64)

    e = EA[:, None] + EB[None, :] - 40
    return np.ldexp(comb(re), e) + 1j * np.ldexp(comb(im), e)


def round_trip(C, axis=1):
    """C written back as digits with one exponent per line (rows: axis=1) and decoded."""
    E = line_exps(C, axis)
    E = E[:, None] if axis == 1 else E[None, :]
    return value(digits(C.real, E), E) + 1j * value(digits(C.imag, E), E)
