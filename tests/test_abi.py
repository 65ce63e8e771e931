"""CPU-side checks of the C ABI library: it loads, exports every symbol the
header declares, and its host-only entry points (file formats, serialization,
sparse MVM, synthetic generator, error paths) agree with the oracle. No GPU
call is made here."""
import json
import os
import re

import numpy as np
import pytest

import oracle as O
import zyklop_amd as zk
from tests import pyref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "zk_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zk_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = header_functions()
    assert len(names) >= 30
    L = zk.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(zk.ABI), "Python binding out of sync with zk_abi.h"


def test_version_string():
    assert b"gfx950" in zk.lib().zk_version()


@pytest.fixture(scope="module")
def small_circuit(tmp_path_factory):
    d = tmp_path_factory.mktemp("zk")
    c = O.Circuit(77, 40, 3, with_crs=True)
    zkey, wtns = str(d / "c.zkey"), str(d / "c.wtns")
    c.write(zkey, wtns)
    return c, zkey, wtns


def test_zkey_header_and_wtns(small_circuit):
    c, zkey, wtns = small_circuit
    h = zk.read_zkey_header(zkey)
    assert h["n_vars"] == c.n_vars and h["n_public"] == c.n_pub
    assert h["domain_size"] == c.n and h["log_domain"] == c.log_n and h["prime"] == pyref.R
    w = zk.read_wtns(wtns)
    assert np.array_equal(w, c.witness())


def test_public_json(small_circuit):
    c, zkey, wtns = small_circuit
    txt = zk.public_json(wtns, c.n_pub)
    vals = json.loads(txt)
    w = c.witness()
    assert vals == [str(O.from_limbs(w[i])) for i in range(1, c.n_pub + 1)]
    # byte layout of src/main.cu:72-108
    assert txt.startswith("[\n \"") and txt.endswith("\"\n]\n")


def test_proof_json_and_calldata(small_circuit):
    c, _, _ = small_circuit
    r, s = c.blinding()
    words = c.prove(r, s)
    pr = zk.Proof()
    for i in range(8):
        pr.pi_a[i] = int(words[i])
        pr.pi_c[i] = int(words[24 + i])
    for i in range(16):
        pr.pi_b[i] = int(words[8 + i])
    txt = zk.proof_to_snarkjs_json(pr)
    j = json.loads(txt)
    canon = O.field_op(O.FP, O.FROM_MONT, words.reshape(8, 4))
    dec = [str(O.from_limbs(x)) for x in canon]
    assert j["pi_a"] == [dec[0], dec[1], "1"]
    assert j["pi_b"] == [[dec[2], dec[3]], [dec[4], dec[5]], ["1", "0"]]
    assert j["pi_c"] == [dec[6], dec[7], "1"]
    assert j["protocol"] == "groth16" and j["curve"] == "bn128"
    assert txt.startswith('{\n "pi_a": ["')   # src/groth16.cu:1490-1501 layout
    cd = zk.proof_to_calldata(pr)
    order = [0, 1, 3, 2, 5, 4, 6, 7]           # G2 as (a1, a0), big-endian
    for k, idx in enumerate(order):
        assert int.from_bytes(cd[32 * k:32 * k + 32], "big") == int(dec[idx])


def test_sparse_mvm_matches_oracle(small_circuit):
    c, zkey, wtns = small_circuit
    data = open(zkey, "rb").read()
    # locate section 4
    pos, n_sec = 12, int.from_bytes(data[8:12], "little")
    sec4 = None
    for _ in range(n_sec):
        t = int.from_bytes(data[pos:pos + 4], "little")
        sz = int.from_bytes(data[pos + 4:pos + 12], "little")
        if t == 4:
            sec4 = data[pos + 12:pos + 12 + sz]
        pos += 12 + sz
    ncoef = int.from_bytes(sec4[:4], "little")
    coefs = sec4[4:4 + 44 * ncoef]
    w = c.witness()
    for sel in (0, 1):
        out = zk.sparse_mvm(coefs, sel, w, c.n_vars, c.n)
        # closed form: sum c*w over the selected matrix, canonical
        acc = [0] * c.n
        rinv2 = pow(pyref.RM * pyref.RM, -1, pyref.R)
        wi = [O.from_limbs(x) for x in w]
        for k in range(ncoef):
            e = coefs[44 * k:44 * k + 44]
            mat = int.from_bytes(e[0:4], "little")
            if mat != sel:
                continue
            row = int.from_bytes(e[4:8], "little")
            wire = int.from_bytes(e[8:12], "little")
            raw = int.from_bytes(e[12:44], "little")
            acc[row] = (acc[row] + raw * rinv2 * wi[wire]) % pyref.R
        assert [O.from_limbs(x) for x in out] == acc


def test_synth_generator_matches_oracle():
    for seed, m, l in ((3, 50, 4), (9, 200, 0), (12345, 513, 17)):
        w = zk.synth_witness(seed, m, l)
        ref = O.Circuit(seed, m, l, with_crs=False)
        assert np.array_equal(w, ref.witness())
        r, s = zk.synth_blinding(seed)
        rr, ss = ref.blinding()
        assert np.array_equal(r, rr) and np.array_equal(s, ss)
    # fresh public inputs keep the R1CS satisfied
    w2 = zk.synth_witness(3, 50, 4, pub_seed=99)
    ref = O.Circuit(3, 50, 4, with_crs=False)
    rr, ss = ref.blinding()
    assert not np.array_equal(w2, ref.witness())
    assert ref.verify(ref.prove_trapdoor(rr, ss, w=w2), rr, ss, w=w2)


def _bad_file(tmp_path, name, data):
    p = tmp_path / name
    p.write_bytes(data)
    return str(p)


def test_error_paths(tmp_path, small_circuit):
    _, zkey, wtns = small_circuit
    good = open(zkey, "rb").read()
    cases = {
        "magic": b"zkez" + good[4:],
        "version": good[:4] + (7).to_bytes(4, "little") + good[8:],
        "truncated": good[: len(good) // 2],
        "short": b"zk",
    }
    for name, data in cases.items():
        with pytest.raises(zk.ZkError) as e:
            zk.read_zkey_header(_bad_file(tmp_path, name + ".zkey", data))
        assert e.value.code == 4, name   # ZK_E_FORMAT
    with pytest.raises(zk.ZkError) as e:
        zk.read_zkey_header(str(tmp_path / "missing.zkey"))
    assert e.value.code == 3                 # ZK_E_IO
    # protocol != groth16: patch section 1 payload
    pos = 12
    bad = bytearray(good)
    t = int.from_bytes(bad[pos:pos + 4], "little")
    assert t == 1
    bad[pos + 12:pos + 16] = (2).to_bytes(4, "little")
    with pytest.raises(zk.ZkError):
        zk.read_zkey_header(_bad_file(tmp_path, "proto.zkey", bytes(bad)))
    # wtns with the wrong magic
    w = open(wtns, "rb").read()
    with pytest.raises(zk.ZkError):
        zk.read_wtns(_bad_file(tmp_path, "bad.wtns", b"wtnz" + w[4:]))
    # buffer too small for calldata / JSON
    pr = zk.Proof()
    buf = (zk.ctypes.c_uint8 * 100)()
    assert zk.lib().zk_proof_to_calldata(zk.ctypes.byref(pr), buf, 100) == 1


def test_wtns_roundtrip(tmp_path):
    w = zk.synth_uniform(5, 33)
    p = str(tmp_path / "x.wtns")
    zk.write_wtns(p, w)
    assert np.array_equal(zk.read_wtns(p), w)


def test_groth16_header_compiles(tmp_path):
    """include/groth16.cuh (drop-in C++ API) compiles and links against libzkp.so."""
    import subprocess
    src = tmp_path / "use.cpp"
    src.write_text('#include "groth16.cuh"\n'
                   'int main(int c, char** v) { Groth16Proof p; if (c > 3) groth16Prove(v[1], v[2], p);\n'
                   '  char j[2048]; groth16ProofToSnarkjsJson(p, j, sizeof j); return 0; }\n')
    out = tmp_path / "use"
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(out),
                        "-L", os.path.join(ROOT, "zyklop_amd"), "-lzkp"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_glv_split_identity_and_bounds():
    """The MSM's scalar split, host build of the device source: for edge and
    random scalars, k1 + k2*lambda = k (mod r) with |k1|, |k2| < 2^127, where
    lambda is a cube root of unity mod r (a root of x^2 + x + 1)."""
    R = pyref.R
    lam = 0xB3C4D79D41A917585BFC41088D8DAAA78B17EA66B99C90DD
    assert (lam * lam + lam + 1) % R == 0
    rng = np.random.default_rng(9)
    ks = [0, 1, 2, R - 1, R - 2, lam, lam * lam % R, R - lam, (R - 1) // 2, (R + 1) // 2,
          2**127, 2**127 - 1, 2**128, R - 2**127, 2**253]
    ks += [int.from_bytes(rng.bytes(32), "little") % R for _ in range(3000)]
    out = zk.glv_split(O.arr(ks))
    for k, o in zip(ks, out):
        v1 = int(o[0]) | int(o[1]) << 64
        v2 = int(o[2]) | int(o[3]) << 64
        k1 = -(v1 & (2**127 - 1)) if v1 >> 127 else v1
        k2 = -(v2 & (2**127 - 1)) if v2 >> 127 else v2
        assert (k1 + k2 * lam - k) % R == 0, k
        assert abs(k1) < 2**127 and abs(k2) < 2**127, k
        assert not (v1 >> 127 and k1 == 0) and not (v2 >> 127 and k2 == 0), k   # no negative zero
    with pytest.raises(zk.ZkError) as e:
        zk.glv_split(O.arr([R]))
    assert e.value.code == 1


def _sec4(zkey):
    data = open(zkey, "rb").read()
    pos, n_sec = 12, int.from_bytes(data[8:12], "little")
    for _ in range(n_sec):
        t = int.from_bytes(data[pos:pos + 4], "little")
        sz = int.from_bytes(data[pos + 4:pos + 12], "little")
        if t == 4:
            sec4 = data[pos + 12:pos + 12 + sz]
            n = int.from_bytes(sec4[:4], "little")
            return sec4[4:4 + 44 * n], n
        pos += 12 + sz
    raise AssertionError("no section 4")


def test_sparse_mvm_canonical_contract(small_circuit, tmp_path):
    """groth16SparseMvm's documented contract (canonical coefficient values,
    include/groth16.cuh of the reference, :93-98 and :183-196), through the C
    ABI and through this repository's C++ header compiled by g++."""
    import subprocess
    c, zkey, wtns = small_circuit
    coefs, ncoef = _sec4(zkey)
    rinv2 = pow(pyref.RM * pyref.RM, -1, pyref.R)
    entries = []
    for k in range(ncoef):
        e = coefs[44 * k:44 * k + 44]
        mat, row, wire = (int.from_bytes(e[4 * i:4 * i + 4], "little") for i in range(3))
        entries.append((mat, row, wire, int.from_bytes(e[12:44], "little") * rinv2 % pyref.R))
    canon = b"".join(m.to_bytes(4, "little") + r.to_bytes(4, "little") + w.to_bytes(4, "little") + v.to_bytes(32, "little")
                     for m, r, w, v in entries)
    w = c.witness()
    wi = [O.from_limbs(x) for x in w]
    src = tmp_path / "mvm.cpp"
    src.write_text(r'''
#include <cstdio>
#include <vector>
#include "groth16.cuh"
int main() {
    size_t n; unsigned nv, dom, sel;
    if (scanf("%zu %u %u %u", &n, &nv, &dom, &sel) != 4) return 2;
    std::vector<Groth16CoefEntry> e(n);
    for (auto& x : e) {
        unsigned m;
        if (scanf("%u %u %u %lu %lu %lu %lu", &m, &x.constraint, &x.signal, &x.value[0], &x.value[1], &x.value[2],
                  &x.value[3]) != 7) return 2;
        x.matrix = (uint8_t)m;
    }
    std::vector<uint64_t> w(4ull * nv), out(4ull * dom);
    for (auto& x : w) if (scanf("%lu", &x) != 1) return 2;
    groth16SparseMvm(e.data(), n, (uint8_t)sel, w.data(), nv, out.data(), dom);
    for (auto x : out) printf("%lu\n", x);
}
''')
    exe = tmp_path / "mvm"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "include"), str(src), "-o", str(exe),
                           "-L", os.path.join(root, "zyklop_amd"), "-lzkp",
                           "-Wl,-rpath," + os.path.join(root, "zyklop_amd")])
    for sel in (0, 1):
        acc = [0] * c.n
        for m, r, wire, v in entries:
            if m == sel:
                acc[r] = (acc[r] + v * wi[wire]) % pyref.R
        out = zk.sparse_mvm(canon, sel, w, c.n_vars, c.n, canonical=True)
        assert [O.from_limbs(x) for x in out] == acc
        lines = [f"{len(entries)} {c.n_vars} {c.n} {sel}"]
        lines += [f"{m} {r} {wire} " + " ".join(str(x) for x in O.to_limbs(v)) for m, r, wire, v in entries]
        lines += [" ".join(str(int(x)) for x in row) for row in w]
        res = subprocess.run([str(exe)], input="\n".join(lines), capture_output=True, text=True, timeout=60)
        assert res.returncode == 0, res.stderr
        words = [int(x) for x in res.stdout.split()]
        assert [O.from_limbs(words[4 * i:4 * i + 4]) for i in range(c.n)] == acc
    # a value not below r is an argument error, as is the raw encoding's check
    bad = canon[:12] + (pyref.R + 1).to_bytes(32, "little") + canon[44:]
    with pytest.raises(zk.ZkError) as e:
        zk.sparse_mvm(bad, 0, w, c.n_vars, c.n, canonical=True)
    assert e.value.code == 1
