"""Device field arithmetic against the reference's golden vectors and the
bounds of the 29-bit form, through the C ABI (zk_field_op_device,
zk_field29_op_device).

* The standard form (8 x 32-bit Montgomery, the product-scanning asm product
  and the C++ CIOS form) replays test/unit/test_vectors.json (Fp) and
  test/unit/test_vectors_fr.json (Fr) of the reference.
* The 9 x 29-bit form of the bucket loops (field29.hpp) is checked against
  Python big ints at the edges of its documented input forms: N (limbs <
  2^29, value < 2p), L (a + 4p - b before normalisation: limbs < 3*2^29,
  value < 6p), normalized values up to 6p, sums of up to four products, and
  values up to 16p for the reduction; the unnormalized negations 2p - y and
  4p - y as product operands, the G1 mixed addition's Y3 sum, and the Fp2
  product and T*A - B*C of the G2 loops. Every result must be congruent to
  the definition and inside its documented output form.
"""
import numpy as np
import pytest

import oracle as O
import zyklop_amd as zk
from tests import pyref

pytestmark = pytest.mark.gpu
P, R = pyref.P, pyref.R
RP = 1 << 261                 # Montgomery radix of the 29-bit form
M29 = (1 << 29) - 1
P4 = [0x21f3f51c, 0x241182da, 0x31ca8d3b, 0x2b548b42, 0x361765df, 0x2b6d0301, 0x229b8503, 0x397098cf, 0x00c19138]


@pytest.fixture(scope="module")
def ctx():
    c = zk.Context(0)
    yield c
    c.close()


def vec(entries, key):
    return np.array([[int(w) for w in e[key]] for e in entries], dtype=np.uint64)


def test_fp_golden_device(ctx, ref_vectors):
    d = ref_vectors("test_vectors.json")
    for op, key, code in (("fp_add", "b", 0), ("fp_sub", "b", 1), ("fp_neg", None, 2), ("fp_mul", "b", 3),
                          ("fp_sqr", None, 4), ("fp_mul", "b", 8)):
        a, b = vec(d[op], "a"), vec(d[op], key) if key else None
        assert np.array_equal(ctx.field_op(0, code, a, b), vec(d[op], "r")), (op, code)
    # fp_inv vectors are canonical in and out; the device inverse is Montgomery in and out
    a = vec(d["fp_inv"], "a")
    got = O.field_op(O.FP, O.FROM_MONT, ctx.field_op(0, 5, O.field_op(O.FP, O.TO_MONT, a)))
    assert np.array_equal(got, vec(d["fp_inv"], "r"))
    # conversions
    x = vec(d["fp_mul"], "a")
    assert np.array_equal(ctx.field_op(0, 6, x), O.field_op(O.FP, O.TO_MONT, x))
    assert np.array_equal(ctx.field_op(0, 7, x), O.field_op(O.FP, O.FROM_MONT, x))


def test_fr_golden_device(ctx, ref_vectors):
    d = ref_vectors("test_vectors_fr.json")
    by = {}
    for v in d["vectors"]:
        by.setdefault(v["op"], []).append(v)
    for op, key, code in (("frAdd", "b", 0), ("frSub", "b", 1), ("frNeg", None, 2), ("frMul", "b", 3),
                          ("frSqr", None, 4), ("frMul", "b", 8)):
        a, b = vec(by[op], "a"), vec(by[op], key) if key else None
        assert np.array_equal(ctx.field_op(1, code, a, b), vec(by[op], "r")), (op, code)
    vs = by["frInv"]                                   # Montgomery in and out
    assert np.array_equal(ctx.field_op(1, 5, vec(vs, "a")), vec(vs, "r"))
    vs = by["frNTT_CT"]                                # a + w*b, a - w*b
    t = ctx.field_op(1, 3, vec(vs, "b"), vec(vs, "omega"))
    assert np.array_equal(ctx.field_op(1, 0, vec(vs, "a"), t), vec(vs, "a_out"))
    assert np.array_equal(ctx.field_op(1, 1, vec(vs, "a"), t), vec(vs, "b_out"))
    vs = by["frRoundTrip"]
    assert np.array_equal(ctx.field_op(1, 6, vec(vs, "x")), vec(vs, "x_enc"))
    assert np.array_equal(ctx.field_op(1, 7, vec(vs, "x_enc")), vec(vs, "x_dec"))


# ---------------------------------------------------------------- 29-bit form
def limbs29(v):
    return [(v >> (29 * i)) & M29 for i in range(8)] + [v >> 232]


def value29(ls):
    return sum(int(x) << (29 * i) for i, x in enumerate(ls))


def arr29(rows):
    return np.array(rows, dtype=np.uint32).reshape(-1, 9)


def check_n(out_row, expect_mod_p, bound=2 * P):
    v = value29(out_row)
    assert all(int(x) <= M29 for x in out_row[:8]), "limbs must be normalized"
    assert v < bound, "value above the documented bound"
    assert v % P == expect_mod_p % P


def edge_values(rng, k, hi):
    """values in [0, hi): both ends, values just below multiples of p, random"""
    vals = [0, 1, P - 1, P, P + 1, 2 * P - 1, hi - 1, hi - 2]
    vals += [m * P - 1 for m in range(1, hi // P + 1) if m * P - 1 < hi]
    vals += [int.from_bytes(rng.bytes(33), "little") % hi for _ in range(k)]
    return [v for v in vals if 0 <= v < hi]


def l_form(u, v):
    """limbs of u + 4p - v as f29_sub4 leaves them (not normalized): limbs < 3 * 2^29"""
    return [a + b - c for a, b, c in zip(limbs29(u), P4, limbs29(v))]


def test_f29_products_at_bounds(ctx):
    rng = np.random.default_rng(29)
    rinv = pow(RP, -1, P)
    # N x N with values up to 2p, and normalized values up to 6p on both sides
    for hi in (2 * P, 6 * P):
        xs = edge_values(rng, 300, hi)
        ys = list(reversed(xs))
        out = ctx.field29_op(0, arr29([limbs29(x) for x in xs]), arr29([limbs29(y) for y in ys]))
        for row, x, y in zip(out, xs, ys):
            check_n(row, x * y * rinv)
        sq = ctx.field29_op(1, arr29([limbs29(x) for x in xs]))
        for row, x in zip(sq, xs):
            check_n(row, x * x * rinv)
    # L x N: a difference of two N values before normalisation, times an N value
    us = edge_values(rng, 200, 2 * P)
    vs = list(reversed(us))
    ws = edge_values(rng, 200, 2 * P)[:len(us)]
    while len(ws) < len(us):
        ws.append(P - 1)
    a = [l_form(u, v) for u, v in zip(us, vs)]
    assert max(max(r) for r in a) < 3 << 29
    out = ctx.field29_op(0, arr29(a), arr29([limbs29(w) for w in ws]))
    for row, u, v, w in zip(out, us, vs, ws):
        check_n(row, (u + 4 * P - v) * w * rinv)


def test_f29_sums_reduce_neg_roundtrip(ctx):
    rng = np.random.default_rng(31)
    rinv = pow(RP, -1, P)
    xs = edge_values(rng, 200, 6 * P)
    ys = edge_values(rng, 200, 6 * P)[:len(xs)]
    while len(ys) < len(xs):
        ys.append(6 * P - 1)
    A, B = arr29([limbs29(x) for x in xs]), arr29([limbs29(y) for y in ys])
    for row, x, y in zip(ctx.field29_op(2, A, B), xs, ys):          # x*y + y*x
        check_n(row, 2 * x * y * rinv)
    for row, x, y in zip(ctx.field29_op(3, A, B), xs, ys):          # (x + y)^2
        check_n(row, (x + y) * (x + y) * rinv)
    # reduction of normalized values below 16p
    zs = edge_values(rng, 300, 16 * P)
    for row, z in zip(ctx.field29_op(6, arr29([limbs29(z) for z in zs])), zs):
        check_n(row, z)
    # negation of y < p: 2p - y, value in (p, 2p]
    ns = edge_values(rng, 100, P)
    for row, y in zip(ctx.field29_op(4, arr29([limbs29(y) for y in ns])), ns):
        v = value29(row)
        assert P < v <= 2 * P and (v + y) % P == 0
    # a + 4p - b normalized: value in (2p, 6p)
    us, vs = edge_values(rng, 100, 2 * P), edge_values(rng, 100, 2 * P)
    k = min(len(us), len(vs))
    for row, u, v in zip(ctx.field29_op(5, arr29([limbs29(u) for u in us[:k]]), arr29([limbs29(v) for v in vs[:k]])),
                         us[:k], vs[:k]):
        val = value29(row)
        assert all(int(x) <= M29 for x in row[:8]) and val == u + 4 * P - v
    # to the standard form and back: canonical (< p) and congruent
    ts = edge_values(rng, 100, 2 * P)
    for row, t in zip(ctx.field29_op(7, arr29([limbs29(t) for t in ts])), ts):
        check_n(row, t, bound=P)


def n_form_edges(rng, k):
    """normalized values below 2p, with one whose limbs below the top are all 2^29 - 1"""
    vals = edge_values(rng, k, 2 * P)
    vals.append(value29([M29] * 8 + [((2 * P) >> 232) - 1]))
    return vals


def test_f29_lazy_negation_and_g1_y3_at_bounds(ctx):
    """4p - y without carries (limbs in (0, 2^30)) and the G1 mixed addition's
    Y3 = R*(Q - X3) + Y1*(4p - PPP) with the difference and the negation left
    unnormalized: a column then sums terms up to 2^60, 2^59 and 2^58."""
    rng = np.random.default_rng(37)
    rinv = pow(RP, -1, P)
    ys = n_form_edges(rng, 200)
    for row, y in zip(ctx.field29_op(8, arr29([limbs29(y) for y in ys])), ys):
        assert all(0 < int(x) < (1 << 30) for x in row), "limbs of 4p - y lie in (0, 2^30)"
        assert value29(row) == 4 * P - y
    xs = n_form_edges(rng, 300)
    zs = list(reversed(xs))
    out = ctx.field29_op(9, arr29([limbs29(x) for x in xs]), arr29([limbs29(z) for z in zs]))
    for row, x, z in zip(out, xs, zs):
        check_n(row, (x * (x + 4 * P - z) + z * (4 * P - z)) * rinv)


def test_f29q_products_at_bounds(ctx):
    """Fp2 products of the G2 loops (element pairs): a*b with a's components
    normalized up to 6p or 2p - y negations of y < p, b's in the N form; and
    T*A - B*C (here A = B = C) with T up to 6p, each component one sum of four
    products with one reduction and two unnormalized 4p - x negations."""
    rng = np.random.default_rng(41)
    rinv = pow(RP, -1, P)
    hi6 = edge_values(rng, 150, 6 * P)
    small = edge_values(rng, 100, P)
    lazy = ctx.field29_op(4, arr29([limbs29(y) for y in small]))      # 2p - y, unnormalized
    a_rows = [limbs29(v) for v in hi6] + [[int(x) for x in r] for r in lazy]
    a_vals = hi6 + [2 * P - y for y in small]
    b_vals = n_form_edges(rng, len(a_vals))
    n = min(len(a_vals), len(b_vals)) // 2 * 2
    A, B = arr29(a_rows[:n]), arr29([limbs29(v) for v in b_vals[:n]])
    out = ctx.field29_op(10, A, B)
    for i in range(n // 2):
        a0, a1, b0, b1 = a_vals[2 * i], a_vals[2 * i + 1], b_vals[2 * i], b_vals[2 * i + 1]
        check_n(out[2 * i], (a0 * b0 - a1 * b1) * rinv)
        check_n(out[2 * i + 1], (a0 * b1 + a1 * b0) * rinv)
    m = min(len(hi6), len(b_vals)) // 2 * 2
    T = arr29([limbs29(v) for v in hi6[:m]])
    out = ctx.field29_op(11, T, arr29([limbs29(v) for v in b_vals[:m]]))
    for i in range(m // 2):
        t0, t1, y0, y1 = hi6[2 * i], hi6[2 * i + 1], b_vals[2 * i], b_vals[2 * i + 1]
        check_n(out[2 * i], ((t0 * y0 - t1 * y1) - (y0 * y0 - y1 * y1)) * rinv)
        check_n(out[2 * i + 1], ((t0 * y1 + t1 * y0) - 2 * y0 * y1) * rinv)
