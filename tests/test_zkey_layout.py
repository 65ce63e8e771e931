"""The .zkey / .wtns layout, read the way the reference reads it, against the
Groth16 CRS computed here from its definition. No GPU.

The oracle writes a synthetic circuit's zkey (byte-identical to the product's
writer, checked on the GPU). This test parses it with the reference reader's
offsets -- the container (src/binfile_utils.cpp:32-80), §1 protocol and the
§2 field order (src/zkey_utils.cpp:19-53), the §4 entries {u32 matrix, u32
row, u32 wire, c*R^2 mod r} (src/groth16.cu:737-805) -- and checks every field
against values fixed independently: the R1CS regenerated from its seed by a
Python restatement of the generator, and the CRS points as multiples of the
generators by exponents computed from the toxic waste with Python big ints
(snarkjs's Groth16 setup, with the odd-coset Lagrange basis in §9).
"""
import struct

import numpy as np

import oracle as O
from tests import pyref

P, R = pyref.P, pyref.R
RM = pyref.RM
M64 = (1 << 64) - 1

SEED, M, L = 777, 60, 3


# ---- the synthetic generator, restated (SplitMix64 -> xoshiro256**, rejection-sampled Fr)
class Xo:
    def __init__(self, seed):
        x = seed
        self.s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & M64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
            self.s.append(z ^ (z >> 31))

    def next(self):
        s = self.s
        rot = lambda v, k: ((v << k) | (v >> (64 - k))) & M64
        res = (rot((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]
        s[2] ^= t
        s[3] = rot(s[3], 45)
        return res

    def fr(self):
        while True:
            w = [self.next() for _ in range(4)]
            w[3] &= 0x3FFFFFFFFFFFFFFF
            v = sum(x << (64 * i) for i, x in enumerate(w))
            if v < R:
                return v

    def below(self, b):
        return (self.next() * b) >> 64


TAG_R1CS = 0x5231435352314353


def r1cs_rows(seed, m, l):
    g = Xo(seed ^ TAG_R1CS)
    rows = []
    for i in range(m):
        bound = 1 + l + i
        a = [(g.below(bound), g.fr()) for _ in range(3)]
        b = [(g.below(bound), g.fr()) for _ in range(3)]
        rows.append((a, b))
    return rows


# ---- container reader at the reference's offsets
def read_sections(blob, magic):
    assert blob[:4] == magic
    version, nsec = struct.unpack_from("<II", blob, 4)
    off = 12
    secs = {}
    for _ in range(nsec):
        typ, size = struct.unpack_from("<IQ", blob, off)
        off += 12
        secs.setdefault(typ, (off, size))
        off += size
    assert off == len(blob)
    return version, secs


def le(b):
    return int.from_bytes(b, "little")


def g1_words(b):
    return [le(b[0:32]), le(b[32:64])]


def mont1(k):
    p, inf = O.g1_mul(O.g1_gen(), k % R)
    return [0, 0] if inf else [O.from_limbs(p[:4]) * RM % P, O.from_limbs(p[4:]) * RM % P]


def mont2(k):
    p, inf = O.g2_mul(O.g2_gen(), k % R)
    return [0] * 4 if inf else [O.from_limbs(p[4 * i:4 * i + 4]) * RM % P for i in range(4)]


def g2_words(b):
    return [le(b[32 * i:32 * i + 32]) for i in range(4)]


def lagrange(tau, n):
    """L_k(tau), k < n, on the size-n domain of w_n = 5^((r-1)/n)."""
    w = pyref.root(n.bit_length() - 1)
    zf = (pow(tau, n, R) - 1) * pow(n, R - 2, R) % R
    out, wk = [], 1
    for _ in range(n):
        out.append(zf * wk * pow((tau - wk) % R, R - 2, R) % R)
        wk = wk * w % R
    return out


def test_zkey_layout_against_definition(tmp_path):
    ref = O.Circuit(SEED, M, L, with_crs=True)
    zkey, wtns = str(tmp_path / "c.zkey"), str(tmp_path / "c.wtns")
    ref.write(zkey, wtns)
    blob = open(zkey, "rb").read()
    version, secs = read_sections(blob, b"zkey")
    assert version == 1 and sorted(secs) == list(range(1, 10))
    o1, s1 = secs[1]
    assert s1 == 4 and struct.unpack_from("<I", blob, o1)[0] == 1          # groth16
    # §2 in the reference reader's order
    o, _ = secs[2]
    n8q = struct.unpack_from("<I", blob, o)[0]; o += 4
    q = le(blob[o:o + n8q]); o += n8q
    n8r = struct.unpack_from("<I", blob, o)[0]; o += 4
    r = le(blob[o:o + n8r]); o += n8r
    n_vars, n_pub, dom = struct.unpack_from("<III", blob, o); o += 12
    alpha1 = g1_words(blob[o:o + 64]); o += 64
    beta1 = g1_words(blob[o:o + 64]); o += 64
    beta2 = g2_words(blob[o:o + 128]); o += 128
    gamma2 = g2_words(blob[o:o + 128]); o += 128
    delta1 = g1_words(blob[o:o + 64]); o += 64
    delta2 = g2_words(blob[o:o + 128]); o += 128
    assert o == secs[2][0] + secs[2][1]
    n = 1
    while n < M + L + 1:
        n *= 2
    assert (n8q, q, n8r, r) == (32, P, 32, R)
    assert (n_vars, n_pub, dom) == (1 + L + M, L, n)
    tau, alpha, beta, gamma, delta = ref.toxic()
    assert alpha1 == mont1(alpha) and beta1 == mont1(beta) and delta1 == mont1(delta)
    assert beta2 == mont2(beta) and gamma2 == mont2(gamma) and delta2 == mont2(delta)

    # §4: u32 nCoefs, then {matrix, row, wire, c*R^2 mod r}
    o, size = secs[4]
    n_coefs = struct.unpack_from("<I", blob, o)[0]
    assert size == 4 + 44 * n_coefs and size // (12 + n8r) == n_coefs   # zkey_utils.cpp:51
    got = set()
    for k in range(n_coefs):
        e = o + 4 + 44 * k
        mat, row, wire = struct.unpack_from("<III", blob, e)
        got.add((mat, row, wire, le(blob[e + 12:e + 44])))
    rows = r1cs_rows(SEED, M, L)
    R2 = RM * RM % R
    want = set()
    for i, (a, b) in enumerate(rows):
        for wire, c in a:
            want.add((0, i, wire, c * R2 % R))
        for wire, c in b:
            want.add((1, i, wire, c * R2 % R))
    for kk in range(L + 1):
        want.add((0, M + kk, kk, R2))                                     # public rows A[m+k][k] = 1
    assert got == want and n_coefs == len(want)

    # CRS points from their definitions (QAP polynomials at tau)
    lag = lagrange(tau, n)
    u, v, ww = [0] * n_vars, [0] * n_vars, [0] * n_vars
    for i, (a, b) in enumerate(rows):
        for wire, c in a:
            u[wire] = (u[wire] + c * lag[i]) % R
        for wire, c in b:
            v[wire] = (v[wire] + c * lag[i]) % R
        ww[1 + L + i] = lag[i]
    for kk in range(L + 1):
        u[kk] = (u[kk] + lag[M + kk]) % R
    gi, di = pow(gamma, R - 2, R), pow(delta, R - 2, R)
    pts = lambda sec, sz: [blob[secs[sec][0] + sz * i: secs[sec][0] + sz * (i + 1)] for i in range(secs[sec][1] // sz)]
    A, B1, B2, C, H, IC = pts(5, 64), pts(6, 64), pts(7, 128), pts(8, 64), pts(9, 64), pts(3, 64)
    assert (len(A), len(B1), len(B2), len(C), len(H), len(IC)) == (n_vars, n_vars, n_vars, n_vars - L - 1, n, L + 1)
    for i in (0, 1, L, L + 1, n_vars // 2, n_vars - 1):
        assert g1_words(A[i]) == mont1(u[i]), ("A", i)
        assert g1_words(B1[i]) == mont1(v[i]), ("B1", i)
        assert g2_words(B2[i]) == mont2(v[i]), ("B2", i)
        e = (beta * u[i] + alpha * v[i] + ww[i]) % R
        if i <= L:
            assert g1_words(IC[i]) == mont1(e * gi), ("IC", i)
        else:
            assert g1_words(C[i - L - 1]) == mont1(e * di), ("C", i)
    lag2 = lagrange(tau, 2 * n)
    for j in (0, 1, n // 2, n - 1):
        assert g1_words(H[j]) == mont1(lag2[2 * j + 1] * di), ("H", j)   # odd-coset Lagrange basis

    # .wtns: §1 {n8, r, nVars}, §2 the witness (src/wtns_utils.cpp:13-26)
    wb = open(wtns, "rb").read()
    wv, wsecs = read_sections(wb, b"wtns")
    assert wv == 2 and sorted(wsecs) == [1, 2]
    o, _ = wsecs[1]
    n8, prime, nv = struct.unpack_from("<I", wb, o)[0], le(wb[o + 4:o + 36]), struct.unpack_from("<I", wb, o + 36)[0]
    assert (n8, prime, nv) == (32, R, n_vars)
    o, size = wsecs[2]
    wit = [le(wb[o + 32 * i:o + 32 * i + 32]) for i in range(nv)]
    assert wit[0] == 1
    for i, (a, b) in enumerate(rows):                                     # aux_i = (A_i w)(B_i w)
        sa = sum(c * wit[wire] for wire, c in a) % R
        sb = sum(c * wit[wire] for wire, c in b) % R
        assert wit[1 + L + i] == sa * sb % R
