"""Pins the CPU oracle: every golden vector the reference's unit tests hold for
this path (test/unit/test_vectors*.json), plus closed forms that a plausible
mistake (dropped term, wrong sign, transposed operand, wrong root) would fail.
"""
import os
import random

import numpy as np
import pytest

import oracle as O
from tests import pyref

RINV_P = pow(pyref.RM, -1, pyref.P)
RINV_R = pow(pyref.RM, -1, pyref.R)


def u(x):
    return [int(v) for v in x]


def vec(entries, key):
    return np.array([[int(w) for w in e[key]] for e in entries], dtype=np.uint64)


# ---------------------------------------------------------------- Fp golden
@pytest.mark.parametrize("op,key,code", [
    ("fp_add", "b", O.ADD), ("fp_sub", "b", O.SUB), ("fp_neg", None, O.NEG),
    ("fp_mul", "b", O.MONT_MUL), ("fp_sqr", None, O.MONT_SQR), ("fp_inv", None, O.INV),
])
def test_fp_golden(ref_vectors, op, key, code):
    d = ref_vectors("test_vectors.json")[op]
    a = vec(d, "a")
    b = vec(d, key) if key else None
    r = O.field_op(O.FP, code, a, b)
    assert np.array_equal(r, vec(d, "r")), op


def test_fp_constants_golden(ref_vectors):
    meta = ref_vectors("test_vectors.json")["meta"]
    assert int(meta["p"], 16) == pyref.P
    assert int(meta["R_mod_p"], 16) == pyref.RM % pyref.P
    assert int(meta["R2"], 16) == pyref.RM * pyref.RM % pyref.P
    assert int(meta["p_prime"], 16) == (-pow(pyref.P, -1, 1 << 64)) % (1 << 64)
    one = O.field_op(O.FP, O.TO_MONT, O.arr([1]))
    assert O.from_limbs(one[0]) == pyref.RM % pyref.P


def test_curve_g1_golden(ref_vectors):
    d = ref_vectors("test_vectors.json")["curve_g1"]
    L = O.lib()
    n_add = n_mul = 0
    for e in d:
        exp = np.array([int(w) for w in e["r"][0]] + [int(w) for w in e["r"][1]], dtype=np.uint64)
        out = np.zeros(8, dtype=np.uint64)
        inf = O.ctypes.c_int(0)
        if "p1" in e:
            p1 = np.array([int(w) for w in e["p1"][0]] + [int(w) for w in e["p1"][1]], dtype=np.uint64)
            p2 = np.array([int(w) for w in e["p2"][0]] + [int(w) for w in e["p2"][1]], dtype=np.uint64)
            L.or_g1_add(O.ptr(p1), 0, O.ptr(p2), 0, O.ptr(out), O.ctypes.byref(inf))
            n_add += 1
        else:
            out, _ = O.g1_mul(O.g1_gen(), int(e["scalar"]))
            n_mul += 1
        assert np.array_equal(out, exp), e.get("note")
    assert n_add == 2 and n_mul == 32


# ---------------------------------------------------------------- Fr golden
def test_fr_golden(ref_vectors):
    d = ref_vectors("test_vectors_fr.json")
    assert int(d["r"], 16) == pyref.R
    R_fr = O.from_limbs([int(x) for x in d["R_Fr"]])
    assert R_fr == pyref.RM % pyref.R
    assert O.from_limbs([int(x) for x in d["R2_Fr"]]) == pyref.RM * pyref.RM % pyref.R
    by_op = {}
    for v in d["vectors"]:
        by_op.setdefault(v["op"], []).append(v)
    simple = {"frAdd": (O.ADD, "b"), "frSub": (O.SUB, "b"), "frNeg": (O.NEG, None),
              "frMul": (O.MONT_MUL, "b"), "frSqr": (O.MONT_SQR, None), "frInvNormal": (O.INV, None)}
    for op, (code, key) in simple.items():
        vs = by_op[op]
        r = O.field_op(O.FR, code, vec(vs, "a"), vec(vs, key) if key else None)
        assert np.array_equal(r, vec(vs, "r")), op
    # frInv: Montgomery in / Montgomery out, and a*inv == 1 (Montgomery) unless a == 0
    vs = by_op["frInv"]
    a = vec(vs, "a")
    r = O.field_op(O.FR, O.TO_MONT, O.field_op(O.FR, O.INV, O.field_op(O.FR, O.FROM_MONT, a)))
    assert np.array_equal(r, vec(vs, "r"))
    prod = O.field_op(O.FR, O.MONT_MUL, a, r)
    assert np.array_equal(prod, vec(vs, "product"))
    # encode / decode
    (enc,) = by_op["frMontEncode_1"]
    assert np.array_equal(O.field_op(O.FR, O.TO_MONT, O.arr([1]))[0], np.array([int(x) for x in enc["expected_R_Fr"]], np.uint64))
    (dec,) = by_op["frMontDecode_RFr"]
    assert O.from_limbs(O.field_op(O.FR, O.FROM_MONT, vec([dec], "R_Fr"))[0]) == 1
    vs = by_op["frRoundTrip"]
    assert np.array_equal(O.field_op(O.FR, O.TO_MONT, vec(vs, "x")), vec(vs, "x_enc"))
    assert np.array_equal(O.field_op(O.FR, O.FROM_MONT, vec(vs, "x_enc")), vec(vs, "x_dec"))
    # CT butterfly a' = a + w*b, b' = a - w*b (Montgomery)
    vs = by_op["frNTT_CT"]
    a, b, w = vec(vs, "a"), vec(vs, "b"), vec(vs, "omega")
    t = O.field_op(O.FR, O.MONT_MUL, b, w)
    assert np.array_equal(O.field_op(O.FR, O.ADD, a, t), vec(vs, "a_out"))
    assert np.array_equal(O.field_op(O.FR, O.SUB, a, t), vec(vs, "b_out"))
    # pow (Montgomery base, u64 exponent)
    vs = by_op["frPow"]
    base = vec(vs, "base")
    ex = np.array([[int(v["exp"]), 0, 0, 0] for v in vs], dtype=np.uint64)
    assert np.array_equal(O.field_op(O.FR, O.MONT_POW, base, ex), vec(vs, "r"))


def test_field_closed_forms():
    rnd = random.Random(7)
    for field, m, rinv in ((O.FP, pyref.P, RINV_P), (O.FR, pyref.R, RINV_R)):
        xs = [rnd.randrange(m) for _ in range(200)] + [0, 1, m - 1]
        ys = [rnd.randrange(m) for _ in range(len(xs))]
        a, b = O.arr(xs), O.arr(ys)
        chk = {
            O.ADD: lambda x, y: (x + y) % m, O.SUB: lambda x, y: (x - y) % m,
            O.MUL: lambda x, y: x * y % m, O.MONT_MUL: lambda x, y: x * y * rinv % m,
            O.NEG: lambda x, y: (-x) % m, O.INV: lambda x, y: pow(x, -1, m) if x else 0,
            O.TO_MONT: lambda x, y: x * pyref.RM % m, O.FROM_MONT: lambda x, y: x * rinv % m,
        }
        for code, f in chk.items():
            r = O.field_op(field, code, a, b)
            assert [O.from_limbs(x) for x in r] == [f(x, y) for x, y in zip(xs, ys)], (field, code)


# ---------------------------------------------------------------- MSM golden
def test_msm_g1_golden(ref_vectors):
    for v in ref_vectors("test_vectors_msm.json"):
        px = np.array(v["Px"], dtype=np.uint64)
        py = np.array(v["Py"], dtype=np.uint64)
        sc = np.array(v["scalars"], dtype=np.uint64)
        exp = np.array(v["out_x"] + v["out_y"], dtype=np.uint64)
        for naive in (False, True):
            out, inf = O.msm_g1(px, py, sc, naive=naive)
            assert not inf
            assert np.array_equal(out, exp), (v["label"], naive)


def test_msm_g2_golden(ref_vectors):
    for v in ref_vectors("test_vectors_msm_g2.json"):
        px = np.array([p["a0"] + p["a1"] for p in v["Px"]], dtype=np.uint64)
        py = np.array([p["a0"] + p["a1"] for p in v["Py"]], dtype=np.uint64)
        sc = np.array(v["scalars"], dtype=np.uint64)
        out, inf = O.msm_g2(px, py, sc)
        o = v["out"]
        assert inf == o["infinity"], v["label"]
        if not inf:
            exp = np.array(o["x"]["a0"] + o["x"]["a1"] + o["y"]["a0"] + o["y"]["a1"], dtype=np.uint64)
            assert np.array_equal(out, exp), v["label"]


def _mont_g1(pt):
    x, y = pt
    return [x * pyref.RM % pyref.P, y * pyref.RM % pyref.P]


def test_msm_g1_discrete_log_closed_form():
    """Points k_i*G with known k_i: MSM = (sum s_i k_i mod r) * G. Includes
    infinity points, repeated points, P and -P, and zero / r-1 scalars."""
    rnd = random.Random(11)
    n = 300
    ks = [rnd.randrange(1, pyref.R) for _ in range(n)]
    ks[5] = ks[4]                    # repeated point
    ks[7] = pyref.R - ks[6]          # P and -P
    ks[9] = 0                        # infinity (0,0)
    sc = [rnd.randrange(pyref.R) for _ in range(n)]
    sc[0], sc[1], sc[2] = 0, 1, pyref.R - 1
    g = O.g1_gen()
    px = np.zeros((n, 4), np.uint64)
    py = np.zeros((n, 4), np.uint64)
    for i, k in enumerate(ks):
        if k == 0:
            continue
        p, _ = O.g1_mul(g, k)
        mx = O.field_op(O.FP, O.TO_MONT, p.reshape(2, 4))
        px[i], py[i] = mx[0], mx[1]
    total = sum(s * k for s, k in zip(sc, ks)) % pyref.R
    exp, _ = O.g1_mul(g, total)
    expm = O.field_op(O.FP, O.TO_MONT, exp.reshape(2, 4)).reshape(8)
    out, inf = O.msm_g1(px, py, O.arr(sc))
    assert not inf and np.array_equal(out, expm)


def test_g2_closed_forms():
    """G2 generator on the twist, k*G2 against affine big-int arithmetic."""
    g2 = O.g2_gen()
    assert O.lib().or_g2_on_curve(O.ptr(g2)) == 1
    assert O.from_limbs(g2[0:4]) == pyref.G2[0][0] and O.from_limbs(g2[4:8]) == pyref.G2[0][1]
    for k in (1, 2, 3, 12345, pyref.R - 1):
        got, inf = O.g2_mul(g2, k)
        exp = pyref.ec_mul(pyref.G2, k, pyref.Fp2)
        assert not inf
        assert [O.from_limbs(got[i:i + 4]) for i in (0, 4, 8, 12)] == [exp[0][0], exp[0][1], exp[1][0], exp[1][1]]
    _, inf = O.g2_mul(g2, pyref.R)
    assert inf   # r * G2 = infinity (G2 has order r)


# ---------------------------------------------------------------- NTT
def test_ntt_golden(ref_vectors):
    for v in ref_vectors("test_vectors_ntt.json"):
        x = np.array(v["input"], dtype=np.uint64)
        exp = np.array(v["expected"], dtype=np.uint64)
        assert np.array_equal(O.ntt(x, v["log_n"], 0, mont=False), exp), v["desc"]
        assert np.array_equal(O.ntt(exp, v["log_n"], 1, mont=False), x), v["desc"]


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 7])
def test_ntt_closed_forms(log_n):
    rnd = random.Random(log_n)
    n = 1 << log_n
    xs = [rnd.randrange(pyref.R) for _ in range(n)]
    a = O.arr(xs)
    w = pyref.root(log_n)
    wl = np.zeros(4, np.uint64)
    O.lib().or_root_of_unity(log_n, O.ptr(wl))
    assert O.from_limbs(wl) == w
    fwd = [O.from_limbs(r) for r in O.ntt(a, log_n, 0, mont=False)]
    assert fwd == pyref.dft(xs, w)
    naive = [O.from_limbs(r) for r in O.ntt(a, log_n, 4, mont=False)]
    assert naive == fwd
    inv = [O.from_limbs(r) for r in O.ntt(a, log_n, 1, mont=False)]
    ninv = pow(n, -1, pyref.R)
    assert inv == [v * ninv % pyref.R for v in pyref.dft(xs, pow(w, -1, pyref.R))]
    # coset: evaluations of the polynomial with coefficients xs at w_2n^(2j+1)
    w2 = pyref.root(log_n + 1)
    cos = [O.from_limbs(r) for r in O.ntt(a, log_n, 2, mont=False)]
    for j in range(n):
        pt = pow(w2, 2 * j + 1, pyref.R)
        assert cos[j] == sum(c * pow(pt, i, pyref.R) for i, c in enumerate(xs)) % pyref.R
    back = O.ntt(O.ntt(a, log_n, 2, mont=False), log_n, 3, mont=False)
    assert np.array_equal(back, a)
    # Montgomery I/O path agrees with the canonical one
    am = O.field_op(O.FR, O.TO_MONT, a)
    fm = O.field_op(O.FR, O.FROM_MONT, O.ntt(am, log_n, 0, mont=True))
    assert [O.from_limbs(r) for r in fm] == fwd


def test_ntt_invariants():
    log_n = 10
    n = 1 << log_n
    delta = np.zeros((n, 4), np.uint64)
    delta[0, 0] = 1
    out = O.ntt(delta, log_n, 0, mont=False)
    assert all(O.from_limbs(r) == 1 for r in out)          # NTT(delta_0) = 1
    ones = O.arr([1] * n)
    out = O.ntt(ones, log_n, 0, mont=False)
    assert O.from_limbs(out[0]) == n and not out[1:].any()  # NTT(1) = n delta_0


# ---------------------------------------------------------------- SpMV + H
def test_spmv_closed_form():
    rnd = random.Random(5)
    n_rows, n_vars = 40, 25
    row_ptr, cols, coefs = [0], [], []
    for _ in range(n_rows):
        k = rnd.randrange(0, 5)
        for _ in range(k):
            cols.append(rnd.randrange(n_vars))
            coefs.append(rnd.randrange(pyref.R))
        row_ptr.append(len(cols))
    w = [rnd.randrange(pyref.R) for _ in range(n_vars)]
    raw = O.arr([c * pyref.RM * pyref.RM % pyref.R for c in coefs])
    out = np.zeros((n_rows, 4), np.uint64)
    rp = np.array(row_ptr, np.uint32)
    cc = np.array(cols, np.uint32)
    wa = O.arr(w)
    assert O.lib().or_spmv(O.ptr(rp), O.ptr(cc), O.ptr(raw), n_rows, O.ptr(wa), n_vars, O.ptr(out)) == 0
    for i in range(n_rows):
        s = sum(coefs[k] * w[cols[k]] for k in range(row_ptr[i], row_ptr[i + 1])) % pyref.R
        assert O.from_limbs(out[i]) == s * pyref.RM % pyref.R


@pytest.mark.parametrize("log_n", [2, 4, 6])
def test_hpoly_closed_form(log_n):
    """h_j = A(x_j)B(x_j) - C(x_j) at x_j = w_2n^(2j+1), where A, B, C are the
    degree < n interpolants of Aw, Bw and Aw o Bw (src/groth16.cu:1037-1090)."""
    rnd = random.Random(log_n)
    n = 1 << log_n
    aw = [rnd.randrange(pyref.R) for _ in range(n)]
    bw = [rnd.randrange(pyref.R) for _ in range(n)]
    cw = [x * y % pyref.R for x, y in zip(aw, bw)]
    am = O.arr([x * pyref.RM % pyref.R for x in aw])
    bm = O.arr([x * pyref.RM % pyref.R for x in bw])
    h = np.zeros((n, 4), np.uint64)
    assert O.lib().or_hpoly(O.ptr(am), O.ptr(bm), log_n, O.ptr(h)) == 0
    w2 = pyref.root(log_n + 1)
    for j in range(n):
        x = pow(w2, 2 * j + 1, pyref.R)
        e = (pyref.interp_eval(aw, x) * pyref.interp_eval(bw, x) - pyref.interp_eval(cw, x)) % pyref.R
        assert O.from_limbs(h[j]) == e


# ---------------------------------------------------------------- full proof
@pytest.mark.parametrize("m,n_pub,seed", [(37, 3, 1), (100, 8, 2), (250, 5, 3)])
def test_prove_matches_closed_form(m, n_pub, seed):
    c = O.Circuit(seed, m, n_pub, with_crs=True)
    r, s = c.blinding()
    proof = c.prove(r, s)
    assert np.array_equal(proof, c.prove_trapdoor(r, s))
    assert c.verify(proof, r, s)
    z = np.zeros(4, np.uint64)
    p0 = c.prove(z, z)                           # reference mode (no blinding)
    assert np.array_equal(p0, c.prove_trapdoor(z, z))
    assert c.verify(p0, z, z)
    assert not np.array_equal(p0, proof)
    bad = proof.copy()
    bad[24] ^= 1                                  # corrupt pi_C.x
    assert not c.verify(bad, r, s)


def test_unsatisfied_witness_fails_verification():
    c = O.Circuit(9, 60, 4, with_crs=False)
    w = c.witness()
    w[-1, 0] ^= 1                                 # break the last constraint
    r, s = c.blinding()
    proof = c.prove_trapdoor(r, s, w=w)
    assert not c.verify(proof, r, s, w=w)


def test_zkey_writer_layout(tmp_path):
    c = O.Circuit(4, 20, 2, with_crs=True)
    zk, wt = str(tmp_path / "c.zkey"), str(tmp_path / "c.wtns")
    c.write(zk, wt)
    b = open(zk, "rb").read()
    assert b[:4] == b"zkey" and int.from_bytes(b[4:8], "little") == 1
    n_sec = int.from_bytes(b[8:12], "little")
    pos, secs = 12, {}
    for _ in range(n_sec):
        t = int.from_bytes(b[pos:pos + 4], "little")
        sz = int.from_bytes(b[pos + 4:pos + 12], "little")
        secs[t] = (pos + 12, sz)
        pos += 12 + sz
    assert pos == len(b)
    assert secs[5][1] == 64 * c.n_vars and secs[7][1] == 128 * c.n_vars
    assert secs[8][1] == 64 * (c.n_vars - c.n_pub - 1) and secs[9][1] == 64 * c.n
    st = secs[4][0]
    ncoef = int.from_bytes(b[st:st + 4], "little")
    assert ncoef == c.nnz_a + c.nnz_b and secs[4][1] == 4 + 44 * ncoef
    w = open(wt, "rb").read()
    assert w[:4] == b"wtns" and len(w) == 12 + 12 + 40 + 12 + 32 * c.n_vars


def test_full_size_helpers_pinned():
    """The oracle helpers behind the full-size GPU checks, against textbook
    definitions: known-log points vs affine double-and-add (including block
    boundaries of the batched construction), Horner evaluation vs the DFT
    definition, and the Fr dot product vs Python big ints."""
    rng = np.random.default_rng(17)
    k0, d = int.from_bytes(rng.bytes(32), "little") % pyref.R, int.from_bytes(rng.bytes(32), "little") % pyref.R
    n = 9000
    pts = O.g1_progression(k0, d, n)
    for i in (0, 1, 2, 4095, 4096, 4097, 8191, 8192, n - 1):
        x, y = pyref.ec_mul(pyref.G1, (k0 + i * d) % pyref.R, pyref.Fp1)
        assert [O.from_limbs(pts[i, :4]), O.from_limbs(pts[i, 4:])] == [x * pyref.RM % pyref.P, y * pyref.RM % pyref.P], i
    # exponent 0 at i = 0 gives the all-zero infinity encoding
    z = O.g1_progression(0, 5, 3)
    assert not z[0].any() and O.from_limbs(z[2, :4]) == pyref.ec_mul(pyref.G1, 10, pyref.Fp1)[0] * pyref.RM % pyref.P
    # Horner evaluation == DFT rows
    log_n = 6
    xs = [int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(1 << log_n)]
    w = pyref.root(log_n)
    dft = pyref.dft(xs, w)
    ks = [0, 1, 5, 63]
    assert O.poly_eval(O.arr(xs), [pow(w, k, pyref.R) for k in ks]) == [dft[k] for k in ks]
    a = [int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(40000)]
    b = [int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(40000)]
    assert O.fr_dot(O.arr(a), O.arr(b)) == sum(x * y for x, y in zip(a, b)) % pyref.R
