"""Serving loop (zk_server_*) on the MI355X: every result is the same alone and
in a mix of concurrent requests over several resident circuits (tolerance 0:
the arithmetic is exact), failures stay per request, cancellation and HBM
admission behave."""
import random

import numpy as np
import pytest

import oracle as O
import zyklop_amd as zk

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    ctx = zk.Context(0)
    specs = [(71, 300, 3), (72, 1200, 16), (73, 5000, 64)]
    circs = [ctx.synth_circuit(seed, m, l) for seed, m, l in specs]
    yield ctx, specs, circs
    for c in circs:
        c.free()
    ctx.close()


def test_mix_equals_alone(setup):
    ctx, specs, circs = setup
    rng = random.Random(1)
    reqs = []
    for k in range(18):
        ci = rng.randrange(len(circs))
        seed = specs[ci][0]
        w = circs[ci].witness(seed * 100 + k)
        r, s = zk.synth_blinding(1000 + k)
        reqs.append((ci, w, r, s))
    alone = [ctx.prove(circs[ci], w, r, s).words() for ci, w, r, s in reqs]
    srv = zk.Server(ctx, max_slots=4, reserve_bytes=1 << 30)
    tickets = [srv.submit(circs[ci], w, r, s) for ci, w, r, s in reqs]
    for k, t in enumerate(tickets):
        pr, times = srv.wait(t)
        assert np.array_equal(pr.words(), alone[k]), k
        assert times[2] >= times[1] >= 0
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_in_flight"] >= 2          # requests really overlapped
    srv.close()
    # one of them against the oracle (trapdoor closed form, verified)
    ci, w, r, s = reqs[0]
    seed, m, l = specs[ci]
    ref = O.Circuit(seed, m, l, with_crs=False)
    assert np.array_equal(alone[0], ref.prove_trapdoor(r, s, w=w))
    assert ref.verify(alone[0], r, s, w=w)


def test_device_witnesses_in_mix(setup):
    ctx, specs, circs = setup
    reqs = []
    for k in range(8):
        ci = k % len(circs)
        w = circs[ci].witness(specs[ci][0] * 7 + k)
        r, s = zk.synth_blinding(2000 + k)
        reqs.append((ci, w, r, s))
    d_ws = []
    for _, w, _, _ in reqs:
        p = ctx.dev_alloc(w.nbytes)
        ctx.upload(p, w)
        d_ws.append(p)
    srv = zk.Server(ctx, max_slots=3, reserve_bytes=1 << 30)
    tickets = []
    for k, (ci, w, r, s) in enumerate(reqs):     # device and host witnesses interleaved
        if k % 2:
            tickets.append(srv.submit(circs[ci], w, r, s))
        else:
            tickets.append(srv.submit_device(circs[ci], d_ws[k], r, s))
    for k, t in enumerate(tickets):
        pr, _ = srv.wait(t)
        ci, w, r, s = reqs[k]
        assert np.array_equal(pr.words(), ctx.prove_device(circs[ci], d_ws[k], r, s).words()), k
    srv.close()
    for p in d_ws:
        ctx.dev_free(p)


def test_bad_request_is_isolated(setup):
    ctx, specs, circs = setup
    srv = zk.Server(ctx, max_slots=2, reserve_bytes=1 << 30)
    w = circs[0].witness(5)
    good = srv.submit(circs[0], w)
    bad = w.copy()
    bad[3] = np.array([0xFFFFFFFFFFFFFFFF] * 4, np.uint64)   # >= r
    with pytest.raises(zk.ZkError) as e:
        srv.submit(circs[0], bad)
    assert e.value.code == 1
    with pytest.raises(zk.ZkError):
        srv.submit(circs[1], w)                 # too short for circuit 1
    pr, _ = srv.wait(good)
    assert np.array_equal(pr.words(), ctx.prove(circs[0], w).words())
    srv.close()


def test_cancel_queued(setup):
    ctx, specs, circs = setup
    srv = zk.Server(ctx, max_slots=1, reserve_bytes=1 << 30)
    w = circs[2].witness(9)
    ts = [srv.submit(circs[2], w) for _ in range(6)]
    assert srv.cancel(ts[-1])
    with pytest.raises(zk.ZkError) as e:
        srv.wait(ts[-1])
    assert e.value.code == 6                    # ZK_E_STATE: cancelled
    for t in ts[:-1]:
        srv.wait(t)
    st = srv.stats()
    assert st["cancelled"] == 1 and st["completed"] == 5 and st["slots"] == 1
    srv.close()


def test_admission_refuses_impossible_request(setup):
    ctx, specs, circs = setup
    srv = zk.Server(ctx, max_slots=2, reserve_bytes=1 << 50)   # reserve beyond any GPU
    with pytest.raises(zk.ZkError) as e:
        srv.submit(circs[0], circs[0].witness(1))
    assert e.value.code == 5                    # ZK_E_NOMEM
    assert srv.stats()["rejected"] == 1
    srv.close()
