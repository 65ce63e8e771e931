"""GPU parity of the window tables: an MSM over a circuit's resident bases
through its window tables (table w holds 2^(offset of window w) * P_i, and the
digits of every window share one set of buckets) equals the MSM over the same
points without them, and a proof is the same with and without tables. All
comparisons are bit for bit."""
import numpy as np
import pytest

import oracle as O
import zyklop_amd as zk
from tests import pyref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = zk.Context(0)
    c.set_window_tables(1)   # whatever ZK_WINDOW_TABLES says for the other suites
    yield c
    c.close()


def rand_scalars(rng, n):
    return O.arr([int.from_bytes(rng.bytes(32), "little") % pyref.R for _ in range(n)])


def plain_msm(ctx, circ, which, d_sc, window=0):
    """The same MSM through the allocating entry point over the standard-form points."""
    p, n = circ.bases(which)
    fn = ctx.msm_g2_device if which == 2 else ctx.msm_g1_device
    return fn(p, d_sc, n, window)


def check_all_arrays(ctx, circ, make_scalars, window=0):
    for which in range(5):
        _, n = circ.bases(which)
        sc = make_scalars(which, n)
        d_sc = ctx.dev_alloc(sc.nbytes)
        try:
            ctx.upload(d_sc, sc)
            got, ginf = ctx.circuit_msm_device(circ, which, d_sc, window)
            exp, einf = plain_msm(ctx, circ, which, d_sc, window)
            assert ginf == einf and np.array_equal(got, exp), (which, window)
        finally:
            ctx.dev_free(d_sc)


def test_tables_built_in_auto_mode(ctx):
    try:
        ctx.set_window_tables(-1)
        circ = ctx.synth_circuit(41, 3000, 8)
        cw, ch = ctx.circuit_window_tables(circ)
        assert cw >= 11 and ch >= 11
    finally:
        ctx.set_window_tables(1)


@pytest.mark.parametrize("m", [300, 5000])
def test_circuit_msm_over_tables_equals_plain(ctx, m):
    circ = ctx.synth_circuit(42, m, 8)
    assert ctx.circuit_window_tables(circ) != (0, 0)
    rng = np.random.default_rng(m)

    def scalars(which, n):
        sc = rand_scalars(rng, n)
        sc[0] = 0
        sc[1] = O.to_limbs(pyref.R - 1)
        sc[2] = O.to_limbs(1)
        return sc

    check_all_arrays(ctx, circ, scalars)


def test_circuit_msm_over_tables_vs_oracle(ctx):
    """A' (G1) and B2' (G2) of a small circuit: the tables' MSM against the
    oracle's MSM of the downloaded points."""
    circ = ctx.synth_circuit(43, 700, 4)
    rng = np.random.default_rng(43)
    for which, words in ((0, 8), (2, 16)):
        p, n = circ.bases(which)
        pts = ctx.download(p, n * words * 8).reshape(n, words)
        sc = rand_scalars(rng, n)
        d_sc = ctx.dev_alloc(sc.nbytes)
        try:
            ctx.upload(d_sc, sc)
            got, ginf = ctx.circuit_msm_device(circ, which, d_sc)
        finally:
            ctx.dev_free(d_sc)
        h = words // 2
        exp, einf = (O.msm_g2 if which == 2 else O.msm_g1)(np.ascontiguousarray(pts[:, :h]),
                                                          np.ascontiguousarray(pts[:, h:]), sc)
        assert ginf == einf and np.array_equal(got, exp), which


def test_circuit_msm_over_tables_skewed(ctx):
    """Scalars mostly 1, 2 or 3: the shared buckets of small digits get
    thousands of entries and take the chunked path, which reads the window
    tables too."""
    circ = ctx.synth_circuit(44, 20000, 8)
    rng = np.random.default_rng(44)

    def scalars(which, n):
        sc = np.zeros((n, 4), np.uint64)
        sc[:, 0] = rng.choice([1, 2, 3], size=n, p=[0.6, 0.3, 0.1])
        k = len(sc[::97])
        sc[::97] = rand_scalars(rng, k)
        return sc

    check_all_arrays(ctx, circ, scalars)


def test_wire_classes_leave_out_infinite_bases(ctx):
    """A wire that no row of A (of B) uses has an infinite A' (B1', B2') base;
    a quarter of a synthetic circuit's wires are such. Over the same scalars
    the A and B MSMs read fewer entries than C, which reads them all, B1 and
    B2 the same ones, and every result equals the MSM without classes."""
    circ = ctx.synth_circuit(48, 6000, 8)
    rng = np.random.default_rng(48)
    _, n = circ.bases(0)
    sc = rand_scalars(rng, n)
    d_sc = ctx.dev_alloc(sc.nbytes)
    try:
        ctx.upload(d_sc, sc)
        read = []
        for which in range(4):
            got, ginf = ctx.circuit_msm_device(circ, which, d_sc)
            read.append(int(ctx.last_timing()[3]))
            exp, einf = plain_msm(ctx, circ, which, d_sc)
            assert ginf == einf and np.array_equal(got, exp), which
    finally:
        ctx.dev_free(d_sc)
    a, b1, b2, c = read
    assert b1 == b2
    assert 0.5 * c < a < 0.95 * c and 0.5 * c < b1 < 0.95 * c, read


def test_circuit_msm_window_without_tables(ctx):
    """A window whose window count differs from the tables' runs over table 0
    (plain buckets per window); one with the same count uses the tables."""
    circ = ctx.synth_circuit(45, 4000, 8)
    cw, _ = ctx.circuit_window_tables(circ)
    rng = np.random.default_rng(45)
    for window in (9, 13, cw):
        check_all_arrays(ctx, circ, lambda which, n: rand_scalars(rng, n), window)


def test_proofs_equal_with_and_without_tables(ctx):
    seed, m, l = 46, 3000, 8
    circ = ctx.synth_circuit(seed, m, l)
    ref = O.Circuit(seed, m, l, with_crs=True)
    w = zk.synth_witness(seed, m, l)
    r, s = zk.synth_blinding(seed)
    on = ctx.prove(circ, w, r, s).words()
    assert np.array_equal(on, ref.prove(r, s))
    assert ctx.circuit_window_tables(circ, 0) == (0, 0)
    off = ctx.prove(circ, w, r, s).words()
    assert np.array_equal(off, on)
    tabs = ctx.circuit_window_tables(circ, 1)
    assert tabs[0] and tabs[1]
    again = ctx.prove(circ, w, r, s).words()
    assert np.array_equal(again, on)
    # forced windows: the tables' window count keeps them, another does not
    try:
        for fw, fh in ((tabs[0], tabs[1]), (12, 14)):
            ctx.set_windows(fw, fh)
            assert np.array_equal(ctx.prove(circ, w, r, s).words(), on), (fw, fh)
    finally:
        ctx.set_windows(0, 0)


def test_tables_mode_for_new_circuits(ctx):
    try:
        ctx.set_window_tables(0)
        circ = ctx.synth_circuit(47, 500, 3)
        assert ctx.circuit_window_tables(circ) == (0, 0)
        ctx.set_window_tables(1)
        circ1 = ctx.synth_circuit(47, 500, 3)
        assert ctx.circuit_window_tables(circ1) != (0, 0)
        w = zk.synth_witness(47, 500, 3)
        r, s = zk.synth_blinding(47)
        assert np.array_equal(ctx.prove(circ, w, r, s).words(), ctx.prove(circ1, w, r, s).words())
    finally:
        ctx.set_window_tables(1)
