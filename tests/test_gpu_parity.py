"""GPU parity: every kernel-level entry point and the full prover, through the
C ABI on the MI355X, against the reference's golden vectors and the oracle.
All arithmetic is exact: every comparison is bit-for-bit."""
import os

import numpy as np
import pytest

import oracle as O
import zyklop_amd as zk
from tests import pyref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = zk.Context(0)
    yield c
    c.close()


def mont_pts_g1(ks):
    """AoS Montgomery affine points k*G (k = 0 gives the (0,0) infinity)."""
    g = O.g1_gen()
    out = np.zeros((len(ks), 8), np.uint64)
    cache = {}
    for i, k in enumerate(ks):
        if k % pyref.R == 0:
            continue
        if k not in cache:
            p, _ = O.g1_mul(g, k)
            cache[k] = O.field_op(O.FP, O.TO_MONT, p.reshape(2, 4)).reshape(8)
        out[i] = cache[k]
    return out


def mont_pts_g2(ks):
    g = O.g2_gen()
    out = np.zeros((len(ks), 16), np.uint64)
    for i, k in enumerate(ks):
        if k % pyref.R == 0:
            continue
        p, _ = O.g2_mul(g, k)
        out[i] = O.field_op(O.FP, O.TO_MONT, p.reshape(4, 4)).reshape(16)
    return out


def oracle_msm_g1(pts, sc):
    return O.msm_g1(np.ascontiguousarray(pts[:, :4]), np.ascontiguousarray(pts[:, 4:]), sc)


def oracle_msm_g2(pts, sc):
    return O.msm_g2(np.ascontiguousarray(pts[:, :8]), np.ascontiguousarray(pts[:, 8:]), sc)


# ------------------------------------------------------------------ MSM
def test_msm_g1_golden(ctx, ref_vectors):
    for v in ref_vectors("test_vectors_msm.json"):
        pts = np.concatenate([np.array(v["Px"], np.uint64), np.array(v["Py"], np.uint64)], axis=1)
        sc = np.array(v["scalars"], np.uint64)
        exp = np.array(v["out_x"] + v["out_y"], np.uint64)
        for window in (0, 4, 7):
            out, inf = ctx.msm_g1(pts, sc, window)
            assert not inf and np.array_equal(out, exp), (v["label"], window)


def test_msm_g2_golden(ctx, ref_vectors):
    for v in ref_vectors("test_vectors_msm_g2.json"):
        px = np.array([p["a0"] + p["a1"] for p in v["Px"]], np.uint64)
        py = np.array([p["a0"] + p["a1"] for p in v["Py"]], np.uint64)
        pts = np.concatenate([px, py], axis=1)
        sc = np.array(v["scalars"], np.uint64)
        out, inf = ctx.msm_g2(pts, sc)
        o = v["out"]
        assert inf == o["infinity"], v["label"]
        if not inf:
            exp = np.array(o["x"]["a0"] + o["x"]["a1"] + o["y"]["a0"] + o["y"]["a1"], np.uint64)
            assert np.array_equal(out, exp), v["label"]


@pytest.mark.parametrize("n,window", [(1, 0), (2, 4), (100, 0), (1000, 8), (3000, 0), (3000, 13), (2000, 9), (2000, 19),
                                      (700, 22)])
def test_msm_g1_random_vs_oracle(ctx, n, window):
    rng = np.random.default_rng(n + window)
    base = [int(x) for x in rng.integers(1, 2**62, size=min(n, 700))]
    ks = [base[i % len(base)] for i in range(n)]
    if n >= 100:
        ks[3] = 0                                  # infinity point
        ks[7] = pyref.R - ks[6]                    # P and -P
    pts = mont_pts_g1(ks)
    sc = O.arr([int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(n)])
    if n >= 100:
        sc[0] = 0
        sc[1] = O.to_limbs(pyref.R - 1)
    out, inf = ctx.msm_g1(pts, sc, window)
    exp, einf = oracle_msm_g1(pts, sc)
    assert inf == einf and np.array_equal(out, exp)


@pytest.mark.parametrize("window", [0, 13])
def test_msm_g1_equal_points_in_bucket(ctx, window):
    """Buckets that receive a point equal to their running sum (doubling) or its
    negative (the sum becomes infinity, then restarts): two-entry buckets make
    the order inside the bucket irrelevant. Checked against the closed form."""
    a, b = 123456789, 987654321
    cases = [
        ([a, a], [1, 1]),                      # P + P: doubling branch
        ([a, pyref.R - a], [1, 1]),            # P + (-P): infinity
        ([a, pyref.R - a, b], [5, 5, 5]),      # cancellation, then another point
        ([a, a, a], [7, 7, 7]),                # doubling, then a plain add
        ([b, b], [pyref.R - 1, pyref.R - 1]),  # negative digits on equal points
    ]
    for ks, ss in cases:
        pts = mont_pts_g1(ks)
        sc = O.arr(ss)
        out, inf = ctx.msm_g1(pts, sc, window)
        tot = sum(s * k for s, k in zip(ss, ks)) % pyref.R
        if tot == 0:
            assert inf, (ks, ss)
            continue
        e, _ = O.g1_mul(O.g1_gen(), tot)
        exp = O.field_op(O.FP, O.TO_MONT, e.reshape(2, 4)).reshape(8)
        assert not inf and np.array_equal(out, exp), (ks, ss, window)


@pytest.mark.parametrize("window", [0, 8, 13])
def test_msm_equal_sums_in_reduction(ctx, window):
    """Consecutive buckets holding the same point (or its negative): the
    reduction's running sums and tree nodes then meet equal and opposite
    operands (doubling and infinity branches of the general addition).
    Scalars 1..m put one point in each of the lowest buckets of window 0."""
    a = 192837465
    m = 300
    sc = O.arr(list(range(1, m + 1)))
    for ks in ([a] * m, [a if i % 2 else pyref.R - a for i in range(m)],
               [a if (i // 3) % 2 else pyref.R - a for i in range(m)]):
        tot = sum((i + 1) * k for i, k in enumerate(ks)) % pyref.R
        out, inf = ctx.msm_g1(mont_pts_g1(ks), sc, window)
        e, _ = O.g1_mul(O.g1_gen(), tot)
        exp = O.field_op(O.FP, O.TO_MONT, e.reshape(2, 4)).reshape(8)
        assert not inf and np.array_equal(out, exp), (window, ks[:4])
    # G2: buckets 0..29 get P and -P (infinity), buckets 30..59 P alone
    pts = mont_pts_g2([a, pyref.R - a])
    pts = np.concatenate([np.repeat(pts[:1], 60, 0), np.repeat(pts[1:], 30, 0)])
    sc2 = O.arr(list(range(1, 61)) + list(range(1, 31)))
    out, inf = ctx.msm_g2(pts, sc2, window)
    exp, einf = oracle_msm_g2(pts, sc2)
    assert not einf and not inf and np.array_equal(out, exp), window


def test_msm_g1_skewed_large_buckets(ctx):
    """Most scalars equal 1 or 2: a few buckets get thousands of entries and take
    the chunked (oversized-bucket) path; the rest stay small."""
    rng = np.random.default_rng(5)
    n = 20000
    base = [int(x) for x in rng.integers(1, 2**62, size=500)]
    ks = [base[i % 500] for i in range(n)]
    pts = mont_pts_g1(ks)
    vals = rng.choice([1, 2, 3], size=n, p=[0.6, 0.3, 0.1])
    sc = np.zeros((n, 4), np.uint64)
    sc[:, 0] = vals
    sc[::97] = O.arr([int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(len(sc[::97]))])
    out, inf = ctx.msm_g1(pts, sc, 10)
    # closed form: sum s_i k_i
    tot = sum(O.from_limbs(sc[i]) * ks[i] for i in range(n)) % pyref.R
    e, _ = O.g1_mul(O.g1_gen(), tot)
    exp = O.field_op(O.FP, O.TO_MONT, e.reshape(2, 4)).reshape(8)
    assert not inf and np.array_equal(out, exp)


def test_msm_g1_all_zero_and_empty(ctx):
    pts = mont_pts_g1([5, 6, 7])
    out, inf = ctx.msm_g1(pts, np.zeros((3, 4), np.uint64))
    assert inf and not out.any()
    out, inf = ctx.msm_g1(np.zeros((0, 8), np.uint64), np.zeros((0, 4), np.uint64))
    assert inf


def test_msm_rejects_noncanonical_scalar(ctx):
    pts = mont_pts_g1([5])
    with pytest.raises(zk.ZkError) as e:
        ctx.msm_g1(pts, O.arr([pyref.R]))
    assert e.value.code == 1


@pytest.mark.parametrize("n,window", [(1, 0), (40, 0), (400, 0), (400, 9), (400, 13)])
def test_msm_g2_random_vs_oracle(ctx, n, window):
    rng = np.random.default_rng(100 + n)
    base = [int(x) for x in rng.integers(1, 2**62, size=min(n, 60))]
    ks = [base[i % len(base)] for i in range(n)]
    pts = mont_pts_g2(ks)
    sc = O.arr([int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(n)])
    out, inf = ctx.msm_g2(pts, sc, window)
    exp, einf = oracle_msm_g2(pts, sc)
    assert inf == einf and np.array_equal(out, exp)


def test_msm_g2_equal_points_in_bucket(ctx):
    """G2 buckets receiving their own running sum (doubling) or its negative."""
    a, b = 1122334455, 5544332211
    for ks, ss in [([a, a], [1, 1]), ([a, pyref.R - a], [1, 1]), ([a, pyref.R - a, b], [3, 3, 3]),
                   ([b, b], [pyref.R - 1, pyref.R - 1])]:
        pts = mont_pts_g2(ks)
        sc = O.arr(ss)
        out, inf = ctx.msm_g2(pts, sc)
        exp, einf = oracle_msm_g2(pts, sc)
        assert inf == einf and np.array_equal(out, exp), (ks, ss)


LAMBDA = 0xB3C4D79D41A917585BFC41088D8DAAA78B17EA66B99C90DD   # cube root of unity mod r (GLV)


def glv_edge_case():
    """Bases P_i = k_i*G where lambda*P_0 = P_1 and lambda*P_1 = P_2 (the
    endomorphism image of one base is another base, so the split puts equal
    points in one bucket), and scalars at the edges of the split: 0, 1,
    lambda and its powers, r - 1, (r +- 1)/2, values around 2^127 and r - 2^127."""
    R = pyref.R
    a, b = 987654321123, 55555
    ks = [a, LAMBDA * a % R, LAMBDA * LAMBDA * a % R, b, R - b, 7, 8, 9, 10, 11, 12, 13, 14, 15]
    ss = [5, 5, 5, 0, 1, LAMBDA, LAMBDA * LAMBDA % R, R - LAMBDA, R - 1, (R - 1) // 2, (R + 1) // 2,
          2**127, 2**127 - 1, R - 2**127]
    return ks, ss


def test_msm_g1_glv_edges(ctx):
    ks, ss = glv_edge_case()
    pts = mont_pts_g1(ks)
    sc = O.arr(ss)
    for window in (0, 4, 13):
        out, inf = ctx.msm_g1(pts, sc, window)
        exp, einf = oracle_msm_g1(pts, sc)
        assert inf == einf and np.array_equal(out, exp), window


def test_msm_g2_glv_edges(ctx):
    ks, ss = glv_edge_case()
    pts = mont_pts_g2(ks)
    sc = O.arr(ss)
    out, inf = ctx.msm_g2(pts, sc)
    exp, einf = oracle_msm_g2(pts, sc)
    assert inf == einf and np.array_equal(out, exp)


# ------------------------------------------------------------------ NTT
def test_ntt_golden(ctx, ref_vectors):
    for v in ref_vectors("test_vectors_ntt.json"):
        x = O.field_op(O.FR, O.TO_MONT, np.array(v["input"], np.uint64))
        exp = O.field_op(O.FR, O.TO_MONT, np.array(v["expected"], np.uint64))
        assert np.array_equal(ctx.ntt(x, v["log_n"], 0), exp), v["desc"]
        assert np.array_equal(ctx.ntt(exp, v["log_n"], 1), x), v["desc"]


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 8, 9, 11, 12, 13, 16])
def test_ntt_vs_oracle(ctx, log_n):
    rng = np.random.default_rng(log_n)
    n = 1 << log_n
    x = O.field_op(O.FR, O.TO_MONT, O.arr([int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(n)]))
    for kind in range(4):
        got = ctx.ntt(x, log_n, kind)
        exp = O.ntt(x, log_n, kind, mont=True)
        assert np.array_equal(got, exp), (log_n, kind)


@pytest.mark.parametrize("log_n", [20, 22])
def test_ntt_large_roundtrip(ctx, log_n):
    """Size-independent checks at full size: inverse(forward(x)) == x,
    coset_inverse(coset(x)) == x, and NTT(delta_j) = (w^(jk))_k at one row."""
    n = 1 << log_n
    rng = np.random.default_rng(log_n)
    x = np.zeros((n, 4), np.uint64)
    x[:, :3] = rng.integers(0, 2**63, size=(n, 3), dtype=np.uint64)
    x[:, 3] = rng.integers(0, 2**58, size=n, dtype=np.uint64)      # < r
    f = ctx.ntt(x, log_n, 0)
    assert np.array_equal(ctx.ntt(f, log_n, 1), x)
    c = ctx.ntt(x, log_n, 2)
    assert np.array_equal(ctx.ntt(c, log_n, 3), x)
    # delta at j = 3: X_k = w^(3k); check a few k against the oracle's root
    d = np.zeros((n, 4), np.uint64)
    d[3] = O.field_op(O.FR, O.TO_MONT, O.arr([1]))[0]
    X = O.field_op(O.FR, O.FROM_MONT, ctx.ntt(d, log_n, 0))
    w = pyref.root(log_n)
    for k in (0, 1, 2, 12345, n - 1):
        assert O.from_limbs(X[k]) == pow(w, 3 * k, pyref.R)


# ------------------------------------------------------------------ SpMV / H
def test_spmv_vs_oracle(ctx):
    rng = np.random.default_rng(3)
    n_rows, n_vars = 5000, 3000
    counts = rng.integers(0, 6, size=n_rows)
    row_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    nnz = int(row_ptr[-1])
    cols = rng.integers(0, n_vars, size=nnz).astype(np.uint32)
    vals = O.arr([int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(nnz)])
    w = O.arr([int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(n_vars)])
    got = ctx.spmv(row_ptr, cols, vals, w)
    exp = np.zeros((n_rows, 4), np.uint64)
    assert O.lib().or_spmv(O.ptr(row_ptr), O.ptr(cols), O.ptr(vals), n_rows, O.ptr(w), n_vars, O.ptr(exp)) == 0
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("log_n", [1, 4, 10, 14])
def test_hpoly_vs_oracle(ctx, log_n):
    rng = np.random.default_rng(log_n)
    n = 1 << log_n
    aw = O.arr([int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(n)])
    bw = O.arr([int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(n)])
    got = ctx.hpoly(aw, bw, log_n)
    exp = np.zeros((n, 4), np.uint64)
    assert O.lib().or_hpoly(O.ptr(aw), O.ptr(bw), log_n, O.ptr(exp)) == 0
    assert np.array_equal(got, exp)


# ------------------------------------------------------------------ full proof
@pytest.mark.parametrize("seed,m,n_pub", [(1, 37, 3), (2, 300, 8), (3, 2000, 0), (4, 1500, 40)])
def test_prove_synthetic_vs_oracle(ctx, seed, m, n_pub):
    circ = ctx.synth_circuit(seed, m, n_pub)
    ref = O.Circuit(seed, m, n_pub, with_crs=True)
    w = zk.synth_witness(seed, m, n_pub)
    r, s = zk.synth_blinding(seed)
    got = ctx.prove(circ, w, r, s).words()
    exp = ref.prove(r, s)
    assert np.array_equal(got, exp)
    assert ref.verify(got, r, s)
    z = np.zeros(4, np.uint64)
    got0 = ctx.prove(circ, w).words()          # reference (unblinded) mode
    assert np.array_equal(got0, ref.prove(z, z))


def test_prove_files_vs_oracle(ctx, tmp_path):
    ref = O.Circuit(21, 700, 6, with_crs=True)
    zkey, wtns = str(tmp_path / "c.zkey"), str(tmp_path / "c.wtns")
    ref.write(zkey, wtns)
    r, s = ref.blinding()
    got = ctx.prove_files(zkey, wtns, r, s).words()
    assert np.array_equal(got, ref.prove(r, s))
    got0 = zk.groth16_prove(zkey, wtns).words()
    z = np.zeros(4, np.uint64)
    assert np.array_equal(got0, ref.prove(z, z))


def test_zkey_written_by_product_equals_oracle(ctx, tmp_path):
    seed, m, l = 8, 400, 4
    circ = ctx.synth_circuit(seed, m, l)
    ref = O.Circuit(seed, m, l, with_crs=True)
    a, b = str(tmp_path / "p.zkey"), str(tmp_path / "o.zkey")
    circ.write_zkey(a)
    ref.write(b, str(tmp_path / "o.wtns"))
    assert open(a, "rb").read() == open(b, "rb").read()


def test_prove_batch_equals_single(ctx):
    circs = [ctx.synth_circuit(30 + i, 200 + 150 * i, 3) for i in range(3)]
    ws, rs, ss, exp = [], [], [], []
    for i, c in enumerate(circs):
        w = zk.synth_witness(30 + i, 200 + 150 * i, 3)
        r, s = zk.synth_blinding(30 + i)
        ws.append(w); rs.append(r); ss.append(s)
        exp.append(ctx.prove(c, w, r, s).words())
    order = [0, 1, 2, 2, 1, 0, 1]
    out = ctx.prove_batch([circs[k] for k in order], [ws[k] for k in order],
                          np.stack([rs[k] for k in order]), np.stack([ss[k] for k in order]))
    for j, k in enumerate(order):
        assert np.array_equal(out[j].words(), exp[k]), j


def test_prove_2pow16_closed_form(ctx):
    """Full-size-class proof checked against the trapdoor closed form."""
    seed, m, l = 7, 1 << 16, 1024
    circ = ctx.synth_circuit(seed, m, l)
    ref = O.Circuit(seed, m, l, with_crs=False)
    w = zk.synth_witness(seed, m, l)
    r, s = zk.synth_blinding(seed)
    got = ctx.prove(circ, w, r, s).words()
    assert np.array_equal(got, ref.prove_trapdoor(r, s))
    assert ref.verify(got, r, s)


@pytest.mark.parametrize("seed,m,n_pub", [(5, 37, 3), (6, 2000, 8), (7, 1 << 16, 1024)])
def test_device_witness_equals_host(ctx, seed, m, n_pub):
    """zk_circuit_witness_device (rows level by level on the GPU) == the host
    generator, for the seed's public inputs and a fresh stream of them."""
    circ = ctx.synth_circuit(seed, m, n_pub)
    d = ctx.dev_alloc(circ.n_vars * 32)
    try:
        for pub_seed in (seed, 4242 + seed):
            circ.witness_device(pub_seed, d)
            got = ctx.download(d, circ.n_vars * 32).reshape(-1, 4)
            assert np.array_equal(got, circ.witness(pub_seed)), pub_seed
        assert np.array_equal(got[:1 + n_pub], zk.synth_witness(seed, m, n_pub, 4242 + seed)[:1 + n_pub])
    finally:
        ctx.dev_free(d)
    circ.free()
