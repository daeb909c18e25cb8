"""The spin tables regenerated from closed-form gammas reproduce the reference's 20 arrays.

Pinned by SHA-256 digests of the reference arrays (include/constants_*.h, serialised as
little-endian int16 / float32 with -0.0 -> +0.0). When the reference tree is present, the
arrays are also compared value by value against the header text.
"""
import os
import re

import numpy as np
import pytest

from oracle import tables

REF = "/root/reference/include"


@pytest.mark.parametrize("name", sorted(tables.REFERENCE_DIGESTS))
def test_digest(name):
    t = tables.all_tables()[name]
    shape, dig = tables.REFERENCE_DIGESTS[name]
    assert t.shape == shape
    assert tables.digest(t) == dig


def _parse_header_arrays():
    out = {}
    for f in sorted(os.listdir(REF)):
        if not (f.startswith("constants_") and f.endswith(".h")):
            continue
        txt = open(os.path.join(REF, f)).read()
        for m in re.finditer(r"(short int|float)\s+(\w+)((?:\[\d+\])+)\s*=\s*\{([^}]*)\}", txt):
            kind, name, dims, body = m.groups()
            shape = tuple(int(x) for x in re.findall(r"\d+", dims))
            vals = [v for v in body.replace("\n", "").split(",") if v.strip()]
            dt = np.int16 if kind == "short int" else np.float32
            out[name] = np.array([float(v) for v in vals], dtype=dt).reshape(shape)
    return out


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_against_header_text():
    ref = _parse_header_arrays()
    gen = tables.all_tables()
    assert set(tables.REFERENCE_DIGESTS) <= set(ref)
    for name in tables.REFERENCE_DIGESTS:
        np.testing.assert_array_equal(gen[name], ref[name], err_msg=name)


def test_row_counts():
    t = tables.all_tables()
    counts = {k: v.shape[0] for k, v in t.items()}
    assert counts["spinIndices_rho"] == 768 and counts["spinIndices_rhopi"] == 768
    assert counts["spinIndices_pipi_square"] == 256 and counts["spinIndices_sigma"] == 64
    assert counts["spinIndices_sigmapi_triangle"] == 128


def test_planted_mistake_fails():
    """A sign flip in one coefficient changes the digest (the check has teeth)."""
    t = tables.all_tables()["coef_rho"].copy()
    t[5, 0] = -t[5, 0]
    assert tables.digest(t) != tables.REFERENCE_DIGESTS["coef_rho"][1]
