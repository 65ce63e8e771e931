"""GPU parity at the benchmark workloads' own sizes, through the C ABI.

Each test names the workload it covers:
  * G1 MSM of 2^22 points over uniform 254-bit scalars, and forward / inverse /
    coset / coset-inverse NTTs of 2^22 uniform elements;
  * the full prove at m = 2^20 constraints with 4096 public inputs;
  * the 64-proof mixed batch (m log-uniform in [2^14, 2^20], half over shared
    circuits), stream-batched;
  * the served mix over resident 2^18 and 2^20 circuits, and the 2^24 circuit
    (a size-2^25 coset NTT) served on prepared slots.

What the oracle can afford at these sizes decides the check: MSM results
against the discrete-log closed form (bases with known logs, result
(sum s_i k_i)*G); NTT rows against direct polynomial evaluation at the row's
point; proofs bit for bit against the trapdoor closed form of the oracle's
circuit, plus the Groth16 equation in the exponent. Proofs also pass the
pairing check (zk_verify), which at 2^24 stands in for the closed form.
Every comparison is exact.
"""
import time

import numpy as np
import pytest

import oracle as O
import zyklop_amd as zk
from tests import pyref

pytestmark = pytest.mark.gpu

RR = pyref.R


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = zk.Context(0)
    yield c
    c.close()


def uniform_fr(rng, n):
    """n uniform-ish canonical Fr values (top word below r's top word, so < r)."""
    x = rng.integers(0, 2**64, size=(n, 4), dtype=np.uint64)
    x[:, 3] = rng.integers(0, 0x30644E72E131A029, size=n, dtype=np.uint64)
    return x


def fr_scalar(rng):
    return int.from_bytes(rng.bytes(32), "little") % RR


def mont_g1(k):
    p, inf = O.g1_mul(O.g1_gen(), k % RR)
    assert not inf
    return O.field_op(O.FP, O.TO_MONT, p.reshape(2, 4)).reshape(8)


# ------------------------------------------------------------------ G1 MSM, 2^22 points
def test_msm_g1_2p22_closed_form(ctx):
    """G1 MSM of 2^22 points by uniform 254-bit scalars at the default window
    and at c = 16: result == (sum_i s_i k_i) * G for bases k_i * G."""
    n = 1 << 22
    rng = np.random.default_rng(2222)
    k0, d = fr_scalar(rng), fr_scalar(rng)
    t0 = time.time()
    pts = O.g1_progression(k0, d, n)
    sc = uniform_fr(rng, n)
    sc[5] = 0
    sc[6] = O.to_limbs(RR - 1)
    # sum_i s_i (k0 + i d) = k0 * sum s_i + d * sum i s_i
    ones = np.zeros((n, 4), np.uint64)
    ones[:, 0] = 1
    idx = np.zeros((n, 4), np.uint64)
    idx[:, 0] = np.arange(n, dtype=np.uint64)
    tot = (k0 * O.fr_dot(sc, ones) + d * O.fr_dot(sc, idx)) % RR
    exp = mont_g1(tot)
    log(f"[msm 2^22] oracle inputs and closed form in {time.time() - t0:.1f} s")
    for window in (0, 16):
        out, inf = ctx.msm_g1(pts, sc, window)
        assert not inf and np.array_equal(out, exp), window
        log(f"[msm 2^22] window {window}: ok")


# ------------------------------------------------------------------ NTT, 2^22 elements
def test_ntt_2p22_rows_by_definition(ctx):
    """All four NTT kinds at 2^22 on uniform data, 8 rows each, against the
    polynomial evaluated at the row's point:
      forward      X_k = x(w^k)
      coset        X_k = x(g^(2k+1)),              g = w_2n
      inverse      x_j = n^-1 * X(w^-j)
      coset inv.   x_j = n^-1 * g^-j * X(w^-j)."""
    log_n = 22
    n = 1 << log_n
    rng = np.random.default_rng(4242)
    x = uniform_fr(rng, n)
    rows = [0, 1, 2, 3, n // 2, n - 1] + [int(v) for v in rng.integers(4, n - 1, size=2)]
    w = pyref.root(log_n)
    g = pyref.root(log_n + 1)
    winv = pow(w, RR - 2, RR)
    ginv = pow(g, RR - 2, RR)
    ninv = pow(n, RR - 2, RR)
    t0 = time.time()
    f = ctx.ntt(x, log_n, 0)
    for k, v in zip(rows, O.poly_eval(x, [pow(w, k, RR) for k in rows])):
        assert O.from_limbs(f[k]) == v, ("forward", k)
    c = ctx.ntt(x, log_n, 2)
    for k, v in zip(rows, O.poly_eval(x, [pow(g, 2 * k + 1, RR) for k in rows])):
        assert O.from_limbs(c[k]) == v, ("coset", k)
    i = ctx.ntt(x, log_n, 1)
    ev = O.poly_eval(x, [pow(winv, j, RR) for j in rows])
    for j, v in zip(rows, ev):
        assert O.from_limbs(i[j]) == v * ninv % RR, ("inverse", j)
    ci = ctx.ntt(x, log_n, 3)
    for j, v in zip(rows, ev):
        assert O.from_limbs(ci[j]) == v * ninv * pow(ginv, j, RR) % RR, ("coset inverse", j)
    log(f"[ntt 2^22] 4 kinds x {len(rows)} rows checked in {time.time() - t0:.1f} s")


# ------------------------------------------------------------------ full prove, m = 2^20, 4096 inputs
SEED2, M2, L2 = 20202, 1 << 20, 4096


@pytest.fixture(scope="module")
def circ2(ctx):
    t0 = time.time()
    circ = ctx.synth_circuit(SEED2, M2, L2)
    log(f"[prove 2^20] device circuit in {time.time() - t0:.1f} s")
    t0 = time.time()
    ref = O.Circuit(SEED2, M2, L2, with_crs=False)
    log(f"[prove 2^20] oracle circuit in {time.time() - t0:.1f} s")
    yield circ, ref
    circ.free()


def fresh_rs(seed):
    rng = np.random.default_rng(seed)
    return O.arr([fr_scalar(rng)])[0], O.arr([fr_scalar(rng)])[0]


def test_prove_2p20_device_vs_trapdoor(ctx, circ2):
    """Full prove at m = 2^20, 4096 inputs (n = 2^21), witness resident in HBM,
    fresh r and s: bit-exact against the oracle's closed form, and verified."""
    circ, ref = circ2
    w = zk.synth_witness(SEED2, M2, L2)
    assert np.array_equal(w, ref.witness())
    r, s = fresh_rs(1)
    d_w = ctx.dev_alloc(w.nbytes)
    try:
        ctx.upload(d_w, w)
        got = ctx.prove_device(circ, d_w, r, s).words()
    finally:
        ctx.dev_free(d_w)
    t0 = time.time()
    assert np.array_equal(got, ref.prove_trapdoor(r, s))
    assert ref.verify(got, r, s)
    log(f"[prove 2^20] device proof == closed form, verified ({time.time() - t0:.1f} s oracle)")
    t0 = time.time()
    assert zk.verify(circ, w[1:L2 + 1], got)
    log(f"[prove 2^20] pairing check passes ({time.time() - t0:.2f} s)")


def test_prove_2p20_batch_vs_trapdoor(ctx, circ2):
    """Stream-batched proofs at m = 2^20: two witnesses (the seed's and one with
    fresh public inputs) in a batch of three with distinct r, s; each proof
    against the closed form for its own witness."""
    circ, ref = circ2
    w1 = zk.synth_witness(SEED2, M2, L2)
    w2 = circ.witness(777)
    assert not np.array_equal(w1[1:L2 + 1], w2[1:L2 + 1])
    ws = [w1, w2, w1]
    rs = [fresh_rs(10 + k) for k in range(3)]
    d = [ctx.dev_alloc(w.nbytes) for w in ws]
    try:
        for p, w in zip(d, ws):
            ctx.upload(p, w)
        out = ctx.prove_batch_device([circ] * 3, d, np.stack([x[0] for x in rs]), np.stack([x[1] for x in rs]))
    finally:
        for p in d:
            ctx.dev_free(p)
    for k in range(3):
        r, s = rs[k]
        got = out[k].words()
        assert np.array_equal(got, ref.prove_trapdoor(r, s, w=ws[k])), k
        assert ref.verify(got, r, s, w=ws[k]), k
    log("[prove 2^20] batched proofs == closed form, verified")


# ------------------------------------------------------------------ mixed batch of 64
def test_mixed_batch_64(ctx):
    """64 proofs, m log-uniform in [2^14, 2^20], 1024 inputs: 32 over 4 shared
    circuits, 32 over circuits of their own, stream-batched on 4 slots. Every
    proof equals the same proof made alone; one proof of each domain size is
    checked against the oracle's closed form and verified."""
    rng = np.random.default_rng(6464)
    n_pub = 1024
    shared_m = [int(round(2 ** rng.uniform(14, 20))) for _ in range(4)]
    distinct_m = [int(round(2 ** rng.uniform(14, 20))) for _ in range(32)]
    specs = [(70000 + i, m) for i, m in enumerate(shared_m)] + [(80000 + j, m) for j, m in enumerate(distinct_m)]
    t0 = time.time()
    circs = [ctx.synth_circuit(sd, m, n_pub) for sd, m in specs]
    plan = [int(rng.integers(0, 4)) for _ in range(32)] + list(range(4, 36))
    plan = [plan[i] for i in rng.permutation(64)]
    ws = [circs[ci].witness(90000 + k) for k, ci in enumerate(plan)]
    log(f"[mixed] 36 circuits, 64 witnesses in {time.time() - t0:.1f} s")
    rs = [fresh_rs(100 + k) for k in range(64)]
    R = np.stack([x[0] for x in rs])
    S = np.stack([x[1] for x in rs])
    d = [ctx.dev_alloc(w.nbytes) for w in ws]
    try:
        for p, w in zip(d, ws):
            ctx.upload(p, w)
        ctx.set_batch_slots(4)
        out = ctx.prove_batch_device([circs[ci] for ci in plan], d, R, S)
        alone = [ctx.prove_device(circs[ci], d[k], R[k], S[k]).words() for k, ci in enumerate(plan)]
    finally:
        ctx.set_batch_slots(3)
        for p in d:
            ctx.dev_free(p)
    for k in range(64):
        assert np.array_equal(out[k].words(), alone[k]), k
    # one proof per domain size against the oracle
    by_size = {}
    for k, ci in enumerate(plan):
        by_size.setdefault(circs[ci].log_n, k)
    t0 = time.time()
    for log_n, k in sorted(by_size.items()):
        sd, m = specs[plan[k]]
        ref = O.Circuit(sd, m, n_pub, with_crs=False)
        assert np.array_equal(alone[k], ref.prove_trapdoor(R[k], S[k], w=ws[k])), (log_n, k)
        assert ref.verify(alone[k], R[k], S[k], w=ws[k]), (log_n, k)
    log(f"[mixed] 64 batched == alone; sizes {sorted(by_size)} checked against the oracle in {time.time() - t0:.1f} s")
    for c in circs:
        c.free()


# ------------------------------------------------------------------ served mix
def test_served_2p18_2p20(ctx, circ2):
    """The server over resident 2^18 (1024 inputs) and 2^20 (4096 inputs)
    circuits: 12 requests with fresh witnesses and blinding submitted at once.
    Each result equals the same proof made alone; one per circuit is checked
    against the oracle's closed form and verified."""
    c20, ref20 = circ2
    seed18, m18 = 181818, 1 << 18
    c18 = ctx.synth_circuit(seed18, m18, 1024)
    reqs = []
    for k in range(12):
        c = c18 if k % 2 == 0 else c20
        reqs.append((c, c.witness(5000 + k), *fresh_rs(500 + k)))
    srv = zk.Server(ctx, max_slots=3)
    try:
        tickets = [srv.submit(c, w, r, s) for c, w, r, s in reqs]
        served = [srv.wait(t)[0].words() for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert st["completed"] == 12 and st["failed"] == 0
    for k, (c, w, r, s) in enumerate(reqs):
        assert np.array_equal(served[k], ctx.prove(c, w, r, s).words()), k
    ref18 = O.Circuit(seed18, m18, 1024, with_crs=False)
    for k, ref in ((0, ref18), (1, ref20)):
        c, w, r, s = reqs[k]
        assert np.array_equal(served[k], ref.prove_trapdoor(r, s, w=w)), k
        assert ref.verify(served[k], r, s, w=w), k
    for k, (c, w, r, s) in enumerate(reqs):
        assert zk.verify(c, w[1:c.n_public + 1], served[k]), k
    log(f"[served] 12 requests == alone, peak in flight {st['peak_in_flight']:.0f}; 2^18 and 2^20 verified")
    c18.free()


def test_served_2p24_pairing_verified(ctx):
    """The largest served circuit, m = 2^24 with 1024 inputs (n = 2^25, a
    size-2^25 coset NTT): the server, with its slots prepared for the circuit,
    proves two requests (fresh blinding each); each equals the proof made
    alone and passes the pairing check against the circuit's verification key.
    The oracle's closed form is not run at this size (minutes of host work);
    the pairing check does not need the trapdoor."""
    t0 = time.time()
    c = ctx.synth_circuit(242424, 1 << 24, 1024)
    w = c.witness(4242)
    log(f"[served 2^24] circuit and witness in {time.time() - t0:.1f} s, {c.device_bytes / 2**30:.1f} GiB resident")
    rs = [fresh_rs(2400 + k) for k in range(2)]
    srv = zk.Server(ctx, max_slots=2)
    try:
        slots = srv.prepare([c])
        t0 = time.time()
        tickets = [srv.submit(c, w, r, s) for r, s in rs]
        served = [srv.wait(t)[0].words() for t in tickets]
        log(f"[served 2^24] 2 requests on {slots} prepared slots in {time.time() - t0:.2f} s")
    finally:
        srv.close()
    for k, (r, s) in enumerate(rs):
        assert np.array_equal(served[k], ctx.prove(c, w, r, s).words()), k
        assert zk.verify(c, w[1:1025], served[k]), k
    assert not np.array_equal(served[0], served[1])
    log("[served 2^24] both == alone, pairing check passes")
    c.free()
