"""ctypes binding of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
where it is the checker. The product package (zyklop_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FP, FR = 0, 1
ADD, SUB, NEG, MONT_MUL, MONT_SQR, INV, TO_MONT, FROM_MONT, MUL, MONT_POW = range(10)

P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def build():
    """Compile liboracle.so from oracle.c (needs gcc + GMP)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_field_op.argtypes = [ctypes.c_int, ctypes.c_int, _u64p, _u64p, _u64p, ctypes.c_size_t]
        L.or_g1_add.argtypes = [_u64p, ctypes.c_int, _u64p, ctypes.c_int, _u64p, ctypes.POINTER(ctypes.c_int)]
        L.or_g1_mul.argtypes = [_u64p, ctypes.c_int, _u64p, _u64p, ctypes.POINTER(ctypes.c_int)]
        L.or_g2_mul.argtypes = [_u64p, ctypes.c_int, _u64p, _u64p, ctypes.POINTER(ctypes.c_int)]
        L.or_g1_on_curve.argtypes = [_u64p]
        L.or_g2_on_curve.argtypes = [_u64p]
        L.or_g1_generator.argtypes = [_u64p]
        L.or_g2_generator.argtypes = [_u64p]
        for fn in (L.or_msm_g1, L.or_msm_g1_naive, L.or_msm_g2):
            fn.argtypes = [_u64p, _u64p, _u64p, ctypes.c_size_t, _u64p, ctypes.POINTER(ctypes.c_int)]
        L.or_ntt.argtypes = [_u64p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_root_of_unity.argtypes = [ctypes.c_int, _u64p]
        L.or_spmv.argtypes = [_u32p, _u32p, _u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _u64p]
        L.or_hpoly.argtypes = [_u64p, _u64p, ctypes.c_int, _u64p]
        L.or_circuit_create.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.or_circuit_create.restype = ctypes.c_void_p
        L.or_circuit_free.argtypes = [ctypes.c_void_p]
        L.or_circuit_sizes.argtypes = [ctypes.c_void_p, _u64p]
        L.or_circuit_witness.argtypes = [ctypes.c_void_p, _u64p]
        L.or_circuit_blinding.argtypes = [ctypes.c_void_p, _u64p, _u64p]
        L.or_circuit_toxic.argtypes = [ctypes.c_void_p, _u64p]
        L.or_circuit_write_zkey.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.or_circuit_write_wtns.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.or_prove.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p]
        L.or_prove_trapdoor.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p]
        L.or_prove_trapdoor_w.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p, _u64p]
        L.or_verify_trapdoor.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p, _u64p]
        L.or_to_dec.argtypes = [_u64p, ctypes.c_char_p, ctypes.c_size_t]
        L.or_g1_progression.argtypes = [_u64p, _u64p, ctypes.c_size_t, _u64p]
        L.or_fr_poly_eval.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _u64p]
        L.or_fr_dot.argtypes = [_u64p, _u64p, ctypes.c_size_t, _u64p]
        L.or_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def ptr(a):
    """Pointer to a C-contiguous numpy uint64/uint32 array."""
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.uint64:
        return a.ctypes.data_as(_u64p)
    if a.dtype == np.uint32:
        return a.ctypes.data_as(_u32p)
    raise TypeError(a.dtype)


def to_limbs(x, n=4):
    return [(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(n)]


def from_limbs(limbs):
    v = 0
    for i, w in enumerate(limbs):
        v |= int(w) << (64 * i)
    return v


def arr(values):
    """list of python ints -> (n,4) uint64 array"""
    return np.array([to_limbs(v) for v in values], dtype=np.uint64).reshape(-1, 4)


def field_op(field, op, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros_like(a)
    bp = None
    if b is not None:
        b = np.ascontiguousarray(b, dtype=np.uint64).reshape(-1, 4)
        bp = ptr(b)
    rc = lib().or_field_op(field, op, ptr(a), bp, ptr(out), a.shape[0])
    assert rc == 0
    return out


def msm_g1(px, py, sc, naive=False):
    px = np.ascontiguousarray(px, dtype=np.uint64)
    py = np.ascontiguousarray(py, dtype=np.uint64)
    sc = np.ascontiguousarray(sc, dtype=np.uint64)
    out = np.zeros(8, dtype=np.uint64)
    inf = ctypes.c_int(0)
    fn = lib().or_msm_g1_naive if naive else lib().or_msm_g1
    assert fn(ptr(px), ptr(py), ptr(sc), sc.size // 4, ptr(out), ctypes.byref(inf)) == 0
    return out, bool(inf.value)


def msm_g2(px, py, sc):
    px = np.ascontiguousarray(px, dtype=np.uint64)
    py = np.ascontiguousarray(py, dtype=np.uint64)
    sc = np.ascontiguousarray(sc, dtype=np.uint64)
    out = np.zeros(16, dtype=np.uint64)
    inf = ctypes.c_int(0)
    assert lib().or_msm_g2(ptr(px), ptr(py), ptr(sc), sc.size // 4, ptr(out), ctypes.byref(inf)) == 0
    return out, bool(inf.value)


def ntt(data, log_n, kind, mont=True):
    d = np.ascontiguousarray(data, dtype=np.uint64).copy()
    assert lib().or_ntt(ptr(d), log_n, kind, 1 if mont else 0) == 0
    return d


def g1_mul(p, k, p_inf=False):
    p = np.ascontiguousarray(p, dtype=np.uint64)
    kk = np.array(to_limbs(k), dtype=np.uint64)
    out = np.zeros(8, dtype=np.uint64)
    inf = ctypes.c_int(0)
    lib().or_g1_mul(ptr(p), int(p_inf), ptr(kk), ptr(out), ctypes.byref(inf))
    return out, bool(inf.value)


def g2_mul(p, k, p_inf=False):
    p = np.ascontiguousarray(p, dtype=np.uint64)
    kk = np.array(to_limbs(k), dtype=np.uint64)
    out = np.zeros(16, dtype=np.uint64)
    inf = ctypes.c_int(0)
    lib().or_g2_mul(ptr(p), int(p_inf), ptr(kk), ptr(out), ctypes.byref(inf))
    return out, bool(inf.value)


def g1_gen():
    g = np.zeros(8, dtype=np.uint64)
    lib().or_g1_generator(ptr(g))
    return g


def g2_gen():
    g = np.zeros(16, dtype=np.uint64)
    lib().or_g2_generator(ptr(g))
    return g


def set_threads(n=0):
    """Threads of the oracle's parallel loops (n <= 0 keeps the setting); returns the count in effect."""
    return lib().or_set_threads(n)


def g1_progression(k0, d, n):
    """(n, 8) Montgomery affine AoS points (k0 + i*d)*G: bases with known logs."""
    out = np.zeros((n, 8), np.uint64)
    k = np.array(to_limbs(k0 % R), np.uint64)
    dd = np.array(to_limbs(d % R), np.uint64)
    assert lib().or_g1_progression(ptr(k), ptr(dd), n, ptr(out)) == 0
    return out


def poly_eval(coeffs, zs):
    """sum_j coeffs[j] * z^j mod r for each python-int z (coefficient form carries through)."""
    c = np.ascontiguousarray(coeffs, np.uint64).reshape(-1, 4)
    z = arr([int(v) % R for v in zs])
    out = np.zeros_like(z)
    assert lib().or_fr_poly_eval(ptr(c), c.shape[0], ptr(z), z.shape[0], ptr(out)) == 0
    return [from_limbs(o) for o in out]


def fr_dot(a, b):
    """sum_i a_i * b_i mod r (canonical (n, 4) arrays) as a python int."""
    a = np.ascontiguousarray(a, np.uint64).reshape(-1, 4)
    b = np.ascontiguousarray(b, np.uint64).reshape(-1, 4)
    assert a.shape == b.shape
    out = np.zeros(4, np.uint64)
    assert lib().or_fr_dot(ptr(a), ptr(b), a.shape[0], ptr(out)) == 0
    return from_limbs(out)


class Circuit:
    """Seeded synthetic R1CS + witness (+ CRS when with_crs) in the oracle."""

    def __init__(self, seed, m, n_pub, with_crs=True):
        self.h = lib().or_circuit_create(seed, m, n_pub, 1 if with_crs else 0)
        s = np.zeros(7, dtype=np.uint64)
        lib().or_circuit_sizes(self.h, ptr(s))
        self.m, self.n_pub, self.n_vars, self.n, self.log_n, self.nnz_a, self.nnz_b = (int(x) for x in s)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_circuit_free(self.h)
            self.h = None

    def witness(self):
        w = np.zeros((self.n_vars, 4), dtype=np.uint64)
        lib().or_circuit_witness(self.h, ptr(w))
        return w

    def blinding(self):
        r = np.zeros(4, dtype=np.uint64)
        s = np.zeros(4, dtype=np.uint64)
        lib().or_circuit_blinding(self.h, ptr(r), ptr(s))
        return r, s

    def toxic(self):
        t = np.zeros(20, dtype=np.uint64)
        lib().or_circuit_toxic(self.h, ptr(t))
        return [from_limbs(t[4 * i:4 * i + 4]) for i in range(5)]

    def write(self, zkey_path, wtns_path):
        assert lib().or_circuit_write_zkey(self.h, zkey_path.encode()) == 0
        assert lib().or_circuit_write_wtns(self.h, wtns_path.encode()) == 0

    def prove(self, r, s):
        out = np.zeros(32, dtype=np.uint64)
        assert lib().or_prove(self.h, ptr(np.asarray(r, np.uint64)), ptr(np.asarray(s, np.uint64)), ptr(out)) == 0
        return out

    def prove_trapdoor(self, r, s, w=None):
        out = np.zeros(32, dtype=np.uint64)
        r = np.ascontiguousarray(r, np.uint64)
        s = np.ascontiguousarray(s, np.uint64)
        if w is None:
            assert lib().or_prove_trapdoor(self.h, ptr(r), ptr(s), ptr(out)) == 0
        else:
            w = np.ascontiguousarray(w, np.uint64)
            assert lib().or_prove_trapdoor_w(self.h, ptr(w), ptr(r), ptr(s), ptr(out)) == 0
        return out

    def verify(self, proof, r, s, w=None):
        if w is None:
            w = self.witness()
        w = np.ascontiguousarray(w, np.uint64)
        return lib().or_verify_trapdoor(self.h, ptr(w), ptr(np.ascontiguousarray(r, np.uint64)),
                                        ptr(np.ascontiguousarray(s, np.uint64)),
                                        ptr(np.ascontiguousarray(proof, np.uint64))) == 1
