"""zyklop_amd: MI355X-native Groth16/BN254 prover (Python binding of libzkp.so).

The compute path is the HIP library built from zyklop_amd/csrc for gfx950;
this module only binds its C ABI (include/zk_abi.h) with ctypes. There is no
CPU fallback: if libzkp.so is missing or no GPU is present the calls fail.

Python mirror of the reference API (include/groth16.cuh):
    groth16_prove(zkey, wtns)            -> Proof         (groth16Prove)
    proof_to_snarkjs_json(proof)         -> str           (groth16ProofToSnarkjsJson)
    proof_to_calldata(proof)             -> bytes         (groth16ProofToCalldata)
    read_zkey_header(zkey)               -> dict          (groth16ReadZkeyHeader)
    sparse_mvm(coefs, sel, w, n_vars, n) -> ndarray       (groth16SparseMvm)
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzkp.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p

P = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001


class ZkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[zk status {code}] {msg}")
        self.code = code


class Proof(ctypes.Structure):
    _fields_ = [("pi_a", ctypes.c_uint64 * 8), ("pi_b", ctypes.c_uint64 * 16), ("pi_c", ctypes.c_uint64 * 8)]

    def words(self):
        return np.array(list(self.pi_a) + list(self.pi_b) + list(self.pi_c), dtype=np.uint64)


class ZkeyHeader(ctypes.Structure):
    _fields_ = [("n_constraints", ctypes.c_uint32), ("n_vars", ctypes.c_uint32), ("n_public", ctypes.c_uint32),
                ("domain_size", ctypes.c_uint32), ("log_domain", ctypes.c_int32), ("prime", ctypes.c_uint64 * 4)]


# name -> (restype, argtypes); every symbol include/zk_abi.h declares
ABI = {
    "zk_last_error": (ctypes.c_char_p, []),
    "zk_version": (ctypes.c_char_p, []),
    "zk_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "zk_ctx_destroy": (None, [_vp]),
    "zk_prove_files": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_char_p, _u64p, _u64p, ctypes.POINTER(Proof)]),
    "zk_circuit_load": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "zk_circuit_free": (None, [_vp]),
    "zk_circuit_info": (ctypes.c_int, [_vp, _u64p]),
    "zk_circuit_write_zkey": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "zk_prove": (ctypes.c_int, [_vp, _vp, _u64p, ctypes.c_uint32, _u64p, _u64p, ctypes.POINTER(Proof)]),
    "zk_prove_device": (ctypes.c_int, [_vp, _vp, _vp, _u64p, _u64p, ctypes.POINTER(Proof)]),
    "zk_set_windows": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "zk_set_window_tables": (ctypes.c_int, [_vp, ctypes.c_int]),
    "zk_circuit_window_tables": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "zk_autotune": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "zk_tune_get": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "zk_tune_load": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    "zk_tune_save": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "zk_batch_graph_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int,
                                             ctypes.POINTER(_vp)]),
    "zk_batch_graph_run": (ctypes.c_int, [_vp, _u64p, _u64p, ctypes.POINTER(Proof)]),
    "zk_batch_graph_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int)]),
    "zk_batch_graph_destroy": (None, [_vp]),
    "zk_prove_batch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_u64p), _u64p, _u64p,
                                      ctypes.POINTER(Proof)]),
    "zk_prove_batch_device": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _u64p, _u64p,
                                             ctypes.POINTER(Proof)]),
    "zk_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "zk_set_batch_slots": (ctypes.c_int, [_vp, ctypes.c_int]),
    "zk_prove_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "zk_server_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "zk_server_destroy": (None, [_vp]),
    "zk_server_prepare": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "zk_server_submit": (ctypes.c_int, [_vp, _vp, _u64p, ctypes.c_uint32, _u64p, _u64p, ctypes.POINTER(ctypes.c_uint64)]),
    "zk_server_submit_device": (ctypes.c_int, [_vp, _vp, _vp, _u64p, _u64p, ctypes.POINTER(ctypes.c_uint64)]),
    "zk_server_poll": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(Proof),
                                      ctypes.POINTER(ctypes.c_double)]),
    "zk_server_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(Proof), ctypes.POINTER(ctypes.c_double)]),
    "zk_server_cancel": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "zk_server_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "zk_proof_to_json": (ctypes.c_int, [ctypes.POINTER(Proof), ctypes.c_char_p, ctypes.c_size_t]),
    "zk_proof_to_calldata": (ctypes.c_int, [ctypes.POINTER(Proof), _u8p, ctypes.c_size_t]),
    "zk_public_json": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t]),
    "zk_read_zkey_header": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ZkeyHeader)]),
    "zk_read_wtns": (ctypes.c_int, [ctypes.c_char_p, _u64p, ctypes.c_uint32, _u32p]),
    "zk_write_wtns": (ctypes.c_int, [ctypes.c_char_p, _u64p, ctypes.c_uint32]),
    "zk_sparse_mvm": (ctypes.c_int, [_u8p, ctypes.c_size_t, ctypes.c_uint8, _u64p, ctypes.c_uint32, _u64p,
                                     ctypes.c_uint32]),
    "zk_sparse_mvm_canonical": (ctypes.c_int, [_u8p, ctypes.c_size_t, ctypes.c_uint8, _u64p, ctypes.c_uint32, _u64p,
                                               ctypes.c_uint32]),
    "zk_ntt": (ctypes.c_int, [_vp, _u64p, ctypes.c_int, ctypes.c_int]),
    "zk_msm_g1": (ctypes.c_int, [_vp, _u64p, _u64p, ctypes.c_size_t, ctypes.c_int, _u64p, ctypes.POINTER(ctypes.c_int)]),
    "zk_msm_g2": (ctypes.c_int, [_vp, _u64p, _u64p, ctypes.c_size_t, ctypes.c_int, _u64p, ctypes.POINTER(ctypes.c_int)]),
    "zk_spmv": (ctypes.c_int, [_vp, _u32p, _u32p, _u64p, ctypes.c_uint32, _u64p, ctypes.c_uint32, _u64p]),
    "zk_hpoly": (ctypes.c_int, [_vp, _u64p, _u64p, ctypes.c_int, _u64p]),
    "zk_dev_alloc": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "zk_dev_free": (ctypes.c_int, [_vp, _vp]),
    "zk_dev_upload": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "zk_dev_download": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "zk_dev_copy": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "zk_ntt_device": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int]),
    "zk_msm_g1_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int, _u64p,
                                        ctypes.POINTER(ctypes.c_int)]),
    "zk_msm_g2_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int, _u64p,
                                        ctypes.POINTER(ctypes.c_int)]),
    "zk_msm_g1_device_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int, _u64p,
                                                ctypes.POINTER(ctypes.c_int)]),
    "zk_last_timing": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "zk_circuit_bases": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_uint64)]),
    "zk_circuit_msm_device": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _u64p,
                                             ctypes.POINTER(ctypes.c_int)]),
    "zk_synth_circuit": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "zk_synth_witness": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, _u64p]),
    "zk_circuit_witness": (ctypes.c_int, [_vp, ctypes.c_uint64, _u64p]),
    "zk_circuit_witness_device": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, _vp]),
    "zk_synth_blinding": (ctypes.c_int, [ctypes.c_uint64, _u64p, _u64p]),
    "zk_synth_uniform": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_size_t, _u64p]),
    "zk_glv_split": (ctypes.c_int, [_u64p, ctypes.c_size_t, _u64p]),
    "zk_synth_sizes": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]),
    "zk_verify": (ctypes.c_int, [_vp, _u64p, ctypes.c_uint32, ctypes.POINTER(Proof), ctypes.POINTER(ctypes.c_int)]),
    "zk_verify_zkey": (ctypes.c_int, [ctypes.c_char_p, _u64p, ctypes.c_uint32, ctypes.POINTER(Proof),
                                      ctypes.POINTER(ctypes.c_int)]),
    "zk_verify_json": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    "zk_pairing": (ctypes.c_int, [_u64p, _u64p, _u64p]),
    "zk_field_op_device": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _u64p, _u64p, _u64p, ctypes.c_size_t]),
    "zk_field29_op_device": (ctypes.c_int, [_vp, ctypes.c_int, _u32p, _u32p, _u32p, ctypes.c_size_t]),
    "zk_probe_valu": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "zk_circuit_spmv_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "zk_slot_bytes_check": (ctypes.c_int, [_vp, _vp, _u64p]),
}


def build(incremental=False):
    """Compile libzkp.so for gfx950 in-tree (hipcc cross-compiles without a GPU).
    Every object is rebuilt unless incremental (then make's header dependencies
    decide)."""
    cmd = ["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")]
    if not incremental:
        cmd.insert(1, "-B")
    subprocess.check_call(cmd)


def lib():
    """Load libzkp.so; raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ZkError(-1, f"{LIB_PATH} is missing: run zyklop_amd.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in ABI.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise ZkError(rc, lib().zk_last_error().decode(errors="replace"))


def u64p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def u32p(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def _opt4(v):
    if v is None:
        return None
    a = np.ascontiguousarray(v, dtype=np.uint64).reshape(4)
    return a


class Context:
    """One device context (streams, NTT tables, workspace)."""

    def __init__(self, device=0):
        h = _vp()
        _check(lib().zk_ctx_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().zk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- kernel-level ----
    def ntt(self, data, log_n, kind):
        d = np.ascontiguousarray(data, dtype=np.uint64).copy()
        _check(lib().zk_ntt(self.h, u64p(d), log_n, kind))
        return d

    def msm_g1(self, points, scalars, window=0):
        pts = np.ascontiguousarray(points, dtype=np.uint64)
        sc = np.ascontiguousarray(scalars, dtype=np.uint64)
        out = np.zeros(8, np.uint64)
        inf = ctypes.c_int(0)
        _check(lib().zk_msm_g1(self.h, u64p(pts), u64p(sc), sc.size // 4, window, u64p(out), ctypes.byref(inf)))
        return out, bool(inf.value)

    def msm_g2(self, points, scalars, window=0):
        pts = np.ascontiguousarray(points, dtype=np.uint64)
        sc = np.ascontiguousarray(scalars, dtype=np.uint64)
        out = np.zeros(16, np.uint64)
        inf = ctypes.c_int(0)
        _check(lib().zk_msm_g2(self.h, u64p(pts), u64p(sc), sc.size // 4, window, u64p(out), ctypes.byref(inf)))
        return out, bool(inf.value)

    def spmv(self, row_ptr, cols, vals, witness):
        rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        cc = np.ascontiguousarray(cols, dtype=np.uint32)
        vv = np.ascontiguousarray(vals, dtype=np.uint64)
        w = np.ascontiguousarray(witness, dtype=np.uint64)
        n_rows = rp.size - 1
        out = np.zeros((n_rows, 4), np.uint64)
        _check(lib().zk_spmv(self.h, u32p(rp), u32p(cc), u64p(vv), n_rows, u64p(w), w.size // 4, u64p(out)))
        return out

    def hpoly(self, aw, bw, log_n):
        a = np.ascontiguousarray(aw, dtype=np.uint64)
        b = np.ascontiguousarray(bw, dtype=np.uint64)
        h = np.zeros_like(a)
        _check(lib().zk_hpoly(self.h, u64p(a), u64p(b), log_n, u64p(h)))
        return h

    def last_timing(self):
        """(total_ms, sort_ms, accumulation_kernel_ms, entries, reduction_ms) of the last kernel-level call."""
        ms = (ctypes.c_float * 5)()
        _check(lib().zk_last_timing(self.h, ms))
        return tuple(float(v) for v in ms)

    def set_windows(self, w_witness=0, w_h=0):
        _check(lib().zk_set_windows(self.h, w_witness, w_h))

    def set_window_tables(self, mode):
        """Window tables for circuits loaded from now on: 1 on, 0 off, -1 auto (default)."""
        _check(lib().zk_set_window_tables(self.h, mode))

    def circuit_window_tables(self, circuit, mode=-2):
        """Build (1), free (0), decide (-1) or only report (-2) circuit's window tables;
        returns their (witness, h) windows, 0 for none."""
        w = (ctypes.c_int * 2)()
        _check(lib().zk_circuit_window_tables(self.h, circuit.h, mode, w))
        return int(w[0]), int(w[1])

    # ---- device buffers ----
    def dev_alloc(self, nbytes):
        p = _vp()
        _check(lib().zk_dev_alloc(self.h, nbytes, ctypes.byref(p)))
        return p

    def dev_free(self, p):
        _check(lib().zk_dev_free(self.h, p))

    def upload(self, dst, arr):
        a = np.ascontiguousarray(arr)
        _check(lib().zk_dev_upload(self.h, dst, a.ctypes.data_as(_vp), a.nbytes))

    def download(self, src, nbytes, dtype=np.uint64):
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        _check(lib().zk_dev_download(self.h, out.ctypes.data_as(_vp), src, nbytes))
        return out

    def dev_copy(self, dst, src, nbytes):
        _check(lib().zk_dev_copy(self.h, dst, src, nbytes))

    def ntt_device(self, d_in, d_out, log_n, kind):
        n = len(d_in)
        ins = (_vp * n)(*d_in)
        outs = (_vp * n)(*d_out)
        _check(lib().zk_ntt_device(self.h, ins, outs, n, log_n, kind))

    def msm_g1_device(self, d_points, d_scalars, n, window=0):
        out = np.zeros(8, np.uint64)
        inf = ctypes.c_int(0)
        _check(lib().zk_msm_g1_device(self.h, d_points, d_scalars, n, window, u64p(out), ctypes.byref(inf)))
        return out, bool(inf.value)

    def circuit_msm_device(self, circuit, which, d_scalars, window=0):
        """MSM over a circuit's resident base array (0 A', 1 B1', 2 B2' (G2), 3 C', 4 H) through
        the arrays the prover reads (window tables when present); (words, is_inf)."""
        out = np.zeros(16, np.uint64)
        inf = ctypes.c_int(0)
        _check(lib().zk_circuit_msm_device(self.h, circuit.h, which, d_scalars, window, u64p(out),
                                           ctypes.byref(inf)))
        return (out if which == 2 else out[:8].copy()), bool(inf.value)

    def msm_g1_device_noalloc(self, d_points, d_scalars, n, window=0):
        """G1 MSM on device buffers with the library's fixed scratch (no allocation)."""
        out = np.zeros(8, np.uint64)
        inf = ctypes.c_int(0)
        _check(lib().zk_msm_g1_device_noalloc(self.h, d_points, d_scalars, n, window, u64p(out), ctypes.byref(inf)))
        return out, bool(inf.value)

    def field_op(self, field, op, a, b=None):
        """Element-wise device Fp (field 0) / Fr (1) op on (n, 4) words (zk_field_op_device)."""
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 4)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.uint64).reshape(-1, 4)
        out = np.zeros_like(a)
        _check(lib().zk_field_op_device(self.h, field, op, u64p(a), u64p(bb) if bb is not None else None, u64p(out),
                                        a.shape[0]))
        return out

    def field29_op(self, op, a, b=None):
        """Element-wise device op of the 29-bit Fp form on (n, 9) raw limbs (zk_field29_op_device)."""
        a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, 9)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.uint32).reshape(-1, 9)
        out = np.zeros_like(a)
        _check(lib().zk_field29_op_device(self.h, op, u32p(a), u32p(bb) if bb is not None else None, u32p(out),
                                          a.shape[0]))
        return out

    def probe_valu(self, kind=0):
        """Lane-ops/s over the GPU of kind & 15: 0 v_mad_u64_u32 (8 chains), 1 v_add_u32,
        2 v_fma_f64 (8 chains), 3 one dependent v_mad_u64_u32 chain; kind >> 4 waves per SIMD (0: 8)."""
        r = ctypes.c_double(0)
        _check(lib().zk_probe_valu(self.h, kind, ctypes.byref(r)))
        return r.value

    def slot_bytes_check(self, circuit):
        """(counted device bytes of a fresh slot for a proof of circuit, bytes the slot
        holds once reserved, bytes still counted after that)."""
        out = np.zeros(3, np.uint64)
        _check(lib().zk_slot_bytes_check(self.h, circuit.h, u64p(out)))
        return int(out[0]), int(out[1]), int(out[2])

    def spmv_device(self, circuit, d_witness, d_aw, d_bw, d_cw):
        """The circuit's fused SpMV alone on the device; returns its device ms."""
        _check(lib().zk_circuit_spmv_device(self.h, circuit.h, d_witness, d_aw, d_bw, d_cw))
        return self.last_timing()[0]

    def msm_g2_device(self, d_points, d_scalars, n, window=0):
        out = np.zeros(16, np.uint64)
        inf = ctypes.c_int(0)
        _check(lib().zk_msm_g2_device(self.h, d_points, d_scalars, n, window, u64p(out), ctypes.byref(inf)))
        return out, bool(inf.value)

    # ---- circuits and proofs ----
    def load_circuit(self, zkey_path):
        h = _vp()
        _check(lib().zk_circuit_load(self.h, zkey_path.encode(), ctypes.byref(h)))
        return Circuit(self, h)

    def synth_circuit(self, seed, m, n_pub):
        h = _vp()
        _check(lib().zk_synth_circuit(self.h, seed, m, n_pub, ctypes.byref(h)))
        return Circuit(self, h)

    def prove(self, circuit, witness, r=None, s=None):
        w = np.ascontiguousarray(witness, dtype=np.uint64)
        pr = Proof()
        r4, s4 = _opt4(r), _opt4(s)
        _check(lib().zk_prove(self.h, circuit.h, u64p(w), w.size // 4, u64p(r4) if r4 is not None else None,
                              u64p(s4) if s4 is not None else None, ctypes.byref(pr)))
        return pr

    def prove_device(self, circuit, d_witness, r=None, s=None):
        pr = Proof()
        r4, s4 = _opt4(r), _opt4(s)
        _check(lib().zk_prove_device(self.h, circuit.h, d_witness, u64p(r4) if r4 is not None else None,
                                     u64p(s4) if s4 is not None else None, ctypes.byref(pr)))
        return pr

    def prove_files(self, zkey, wtns, r=None, s=None):
        pr = Proof()
        r4, s4 = _opt4(r), _opt4(s)
        _check(lib().zk_prove_files(self.h, zkey.encode(), wtns.encode(), u64p(r4) if r4 is not None else None,
                                    u64p(s4) if s4 is not None else None, ctypes.byref(pr)))
        return pr

    def prove_batch(self, circuits, witnesses, r=None, s=None):
        n = len(circuits)
        ws = [np.ascontiguousarray(w, dtype=np.uint64) for w in witnesses]
        ch = (_vp * n)(*[c.h for c in circuits])
        wp = (_u64p * n)(*[u64p(w) for w in ws])
        rr = np.ascontiguousarray(r, dtype=np.uint64).reshape(-1) if r is not None else None
        ss = np.ascontiguousarray(s, dtype=np.uint64).reshape(-1) if s is not None else None
        out = (Proof * n)()
        _check(lib().zk_prove_batch(self.h, n, ch, wp, u64p(rr) if rr is not None else None,
                                    u64p(ss) if ss is not None else None, out))
        return list(out)


    def prove_batch_device(self, circuits, d_witnesses, r=None, s=None):
        n = len(circuits)
        ch = (_vp * n)(*[c.h for c in circuits])
        wp = (_vp * n)(*d_witnesses)
        rr = np.ascontiguousarray(r, dtype=np.uint64).reshape(-1) if r is not None else None
        ss = np.ascontiguousarray(s, dtype=np.uint64).reshape(-1) if s is not None else None
        out = (Proof * n)()
        _check(lib().zk_prove_batch_device(self.h, n, ch, wp, u64p(rr) if rr is not None else None,
                                           u64p(ss) if ss is not None else None, out))
        return list(out)

    def set_batch_slots(self, n):
        _check(lib().zk_set_batch_slots(self.h, n))

    def autotune(self, circuit, force=False):
        """Measure and record the fastest MSM windows for circuit's sizes; returns (c_witness, c_h)."""
        w = (ctypes.c_int * 2)()
        _check(lib().zk_autotune(self.h, circuit.h, 1 if force else 0, w))
        return int(w[0]), int(w[1])

    def tuned_window(self, kind, n):
        """Window a proof would use for an MSM of kind (0 witness, 1 h) and size n."""
        c = ctypes.c_int(0)
        _check(lib().zk_tune_get(self.h, kind, n, ctypes.byref(c)))
        return c.value

    def tune_load(self, path):
        k = ctypes.c_int(0)
        _check(lib().zk_tune_load(self.h, str(path).encode(), ctypes.byref(k)))
        return k.value

    def tune_save(self, path):
        _check(lib().zk_tune_save(self.h, str(path).encode()))

    def batch_graph(self, circuits, d_witnesses, slots=3):
        """Record a fixed batch of proofs as one HIP graph (BatchGraph)."""
        return BatchGraph(self, circuits, d_witnesses, slots)

    def set_profiling(self, on=True):
        _check(lib().zk_set_profiling(self.h, 1 if on else 0))

    STAT_NAMES = ["spmv_ms", "intt3_ms", "ntt3_ms", "h_sort_ms", "sort_w_ms",
                  "acc_A_ms", "acc_B1_ms", "acc_C_ms", "acc_B2_ms", "acc_H_ms",
                  "msm_A_ms", "msm_B1_ms", "msm_C_ms", "msm_B2_ms", "msm_H_ms",
                  "entries_w", "entries_h", "c_w", "c_h", "windows_w", "windows_h",
                  "stream0_ms", "stream1_ms", "host_tail_ms"]

    def prove_stats(self):
        buf = (ctypes.c_double * 24)()
        _check(lib().zk_prove_stats(self.h, buf, 24))
        return dict(zip(self.STAT_NAMES, [float(x) for x in buf]))


class BatchGraph:
    """A fixed sequence of proofs captured once as a HIP graph (zk_batch_graph_*)."""

    def __init__(self, ctx, circuits, d_witnesses, slots=3):
        n = len(circuits)
        ch = (_vp * n)(*[c.h for c in circuits])
        wp = (_vp * n)(*d_witnesses)
        h = _vp()
        _check(lib().zk_batch_graph_create(ctx.h, n, ch, wp, slots, ctypes.byref(h)))
        self.h, self.ctx, self.count = h, ctx, n
        self._keep = (list(circuits), list(d_witnesses))
        nodes, ns = ctypes.c_uint64(0), ctypes.c_int(0)
        _check(lib().zk_batch_graph_info(h, ctypes.byref(nodes), ctypes.byref(ns)))
        self.n_nodes, self.slots = nodes.value, ns.value

    def run(self, r=None, s=None):
        rr = np.ascontiguousarray(r, dtype=np.uint64).reshape(-1) if r is not None else None
        ss = np.ascontiguousarray(s, dtype=np.uint64).reshape(-1) if s is not None else None
        out = (Proof * self.count)()
        _check(lib().zk_batch_graph_run(self.h, u64p(rr) if rr is not None else None,
                                        u64p(ss) if ss is not None else None, out))
        return list(out)

    def close(self):
        if getattr(self, "h", None):
            lib().zk_batch_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Server:
    """Continuous-batching proof server (zk_server_*) on a context."""

    STAT_NAMES = ["completed", "failed", "rejected", "cancelled", "queued", "in_flight", "slots", "peak_in_flight"]

    def __init__(self, ctx, max_slots=3, reserve_bytes=8 << 30):
        h = _vp()
        _check(lib().zk_server_create(ctx.h, max_slots, reserve_bytes, ctypes.byref(h)))
        self.h = h
        self.ctx = ctx

    def close(self):
        if getattr(self, "h", None):
            lib().zk_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def prepare(self, circuits):
        """Size every slot for every circuit up front (idle server); returns the slot count."""
        n = len(circuits)
        ch = (_vp * n)(*[c.h for c in circuits])
        k = ctypes.c_int(0)
        _check(lib().zk_server_prepare(self.h, ch, n, ctypes.byref(k)))
        return k.value

    def submit(self, circuit, witness, r=None, s=None):
        w = np.ascontiguousarray(witness, dtype=np.uint64)
        r4, s4 = _opt4(r), _opt4(s)
        t = ctypes.c_uint64(0)
        _check(lib().zk_server_submit(self.h, circuit.h, u64p(w), w.size // 4, u64p(r4) if r4 is not None else None,
                                      u64p(s4) if s4 is not None else None, ctypes.byref(t)))
        return t.value

    def submit_device(self, circuit, d_witness, r=None, s=None):
        """Queue a proof whose witness is already in device memory (kept valid until done)."""
        r4, s4 = _opt4(r), _opt4(s)
        t = ctypes.c_uint64(0)
        _check(lib().zk_server_submit_device(self.h, circuit.h, d_witness, u64p(r4) if r4 is not None else None,
                                             u64p(s4) if s4 is not None else None, ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        """None while pending, else (Proof, (queued_ms, service_ms, total_ms))."""
        done = ctypes.c_int(0)
        pr = Proof()
        tm = (ctypes.c_double * 3)()
        _check(lib().zk_server_poll(self.h, ticket, ctypes.byref(done), ctypes.byref(pr), tm))
        return (pr, tuple(tm)) if done.value else None

    def wait(self, ticket):
        pr = Proof()
        tm = (ctypes.c_double * 3)()
        _check(lib().zk_server_wait(self.h, ticket, ctypes.byref(pr), tm))
        return pr, tuple(tm)

    def cancel(self, ticket):
        c = ctypes.c_int(0)
        _check(lib().zk_server_cancel(self.h, ticket, ctypes.byref(c)))
        return bool(c.value)

    def stats(self):
        buf = (ctypes.c_double * 8)()
        _check(lib().zk_server_stats(self.h, buf))
        return dict(zip(self.STAT_NAMES, [float(x) for x in buf]))


class Circuit:
    """A circuit resident in HBM (CSR of A/B + all CRS bases)."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h
        info = np.zeros(7, np.uint64)
        _check(lib().zk_circuit_info(h, u64p(info)))
        (self.n_vars, self.n_public, self.n, self.log_n, self.nnz_a, self.nnz_b,
         self.device_bytes) = (int(x) for x in info)

    def free(self):
        if getattr(self, "h", None):
            lib().zk_circuit_free(self.h)
            self.h = None

    def __del__(self):
        self.free()

    def write_zkey(self, path):
        _check(lib().zk_circuit_write_zkey(self.h, path.encode()))

    def bases(self, which):
        """(device pointer, point count) of a resident base array: 0 A', 1 B1', 2 B2', 3 C', 4 H."""
        p = _vp()
        n = ctypes.c_uint64(0)
        _check(lib().zk_circuit_bases(self.h, which, ctypes.byref(p), ctypes.byref(n)))
        return p, n.value

    def witness(self, pub_seed):
        """Satisfying witness of a synthetic circuit with public inputs from pub_seed."""
        w = np.zeros((self.n_vars, 4), np.uint64)
        _check(lib().zk_circuit_witness(self.h, pub_seed, u64p(w)))
        return w

    def witness_device(self, pub_seed, d_w):
        """The same witness computed on the device into d_w (n_vars x 32 bytes)."""
        _check(lib().zk_circuit_witness_device(self.ctx.h, self.h, pub_seed, d_w))


# ---- reference-API mirror (include/groth16.cuh) ----
_default_ctx = None


def _ctx():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def groth16_prove(zkey_path, wtns_path, r=None, s=None):
    """groth16Prove (include/groth16.cuh:119-123); r = s = None is the reference's unblinded proof."""
    return _ctx().prove_files(zkey_path, wtns_path, r, s)


def proof_to_snarkjs_json(proof):
    buf = ctypes.create_string_buffer(4096)
    _check(lib().zk_proof_to_json(ctypes.byref(proof), buf, 4096))
    return buf.value.decode()


def proof_to_calldata(proof):
    buf = (ctypes.c_uint8 * 256)()
    _check(lib().zk_proof_to_calldata(ctypes.byref(proof), buf, 256))
    return bytes(buf)


def public_json(wtns_path, n_public):
    size = 96 * (n_public + 2) + 16
    buf = ctypes.create_string_buffer(size)
    _check(lib().zk_public_json(wtns_path.encode(), n_public, buf, size))
    return buf.value.decode()


def read_zkey_header(zkey_path):
    h = ZkeyHeader()
    _check(lib().zk_read_zkey_header(zkey_path.encode(), ctypes.byref(h)))
    return {"n_constraints": h.n_constraints, "n_vars": h.n_vars, "n_public": h.n_public,
            "domain_size": h.domain_size, "log_domain": h.log_domain,
            "prime": sum(int(w) << (64 * i) for i, w in enumerate(h.prime))}


def read_wtns(path):
    n = ctypes.c_uint32(0)
    _check(lib().zk_read_wtns(path.encode(), None, 0, ctypes.byref(n)))
    w = np.zeros((n.value, 4), np.uint64)
    _check(lib().zk_read_wtns(path.encode(), u64p(w), n.value, ctypes.byref(n)))
    return w


def write_wtns(path, w):
    a = np.ascontiguousarray(w, dtype=np.uint64)
    _check(lib().zk_write_wtns(path.encode(), u64p(a), a.size // 4))


def sparse_mvm(coefs44, matrix_sel, witness, n_vars, domain_size, canonical=False):
    """groth16SparseMvm (include/groth16.cuh:183-205): 44-byte entries whose values
    are the zkey's c*R^2 encoding, or canonical c (the header's contract) with
    canonical=True; canonical witness in, canonical result out."""
    c = np.frombuffer(coefs44, dtype=np.uint8)
    w = np.ascontiguousarray(witness, dtype=np.uint64)
    out = np.zeros((domain_size, 4), np.uint64)
    fn = lib().zk_sparse_mvm_canonical if canonical else lib().zk_sparse_mvm
    _check(fn(c.ctypes.data_as(_u8p), c.size // 44, matrix_sel, u64p(w), n_vars, u64p(out), domain_size))
    return out


def synth_witness(seed, m, n_pub, pub_seed=None):
    nv = ctypes.c_uint32(0)
    dom = ctypes.c_uint32(0)
    _check(lib().zk_synth_sizes(m, n_pub, ctypes.byref(nv), ctypes.byref(dom)))
    w = np.zeros((nv.value, 4), np.uint64)
    _check(lib().zk_synth_witness(seed, m, n_pub, seed if pub_seed is None else pub_seed, u64p(w)))
    return w


def synth_blinding(seed):
    r = np.zeros(4, np.uint64)
    s = np.zeros(4, np.uint64)
    _check(lib().zk_synth_blinding(seed, u64p(r), u64p(s)))
    return r, s


def synth_uniform(seed, count):
    out = np.zeros((count, 4), np.uint64)
    _check(lib().zk_synth_uniform(seed, count, u64p(out)))
    return out


def glv_split(scalars):
    """Host run of the MSM's GLV split: (n,4) canonical scalars -> (n,4) words,
    [k1 lo, k1 hi, k2 lo, k2 hi], each half a magnitude < 2^127 with its sign
    in bit 127."""
    k = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros_like(k)
    _check(lib().zk_glv_split(u64p(k), k.shape[0], u64p(out)))
    return out


def _proof_arg(proof):
    if isinstance(proof, Proof):
        return proof
    w = np.ascontiguousarray(proof, dtype=np.uint64).reshape(32)
    pr = Proof()
    ctypes.memmove(ctypes.byref(pr), w.ctypes.data, 256)
    return pr


def _publics_arg(publics):
    p = np.ascontiguousarray(publics, dtype=np.uint64).reshape(-1, 4)
    return p, p.shape[0]


def verify(circuit, publics, proof):
    """Pairing check of a proof against a resident circuit's verification key."""
    p, n = _publics_arg(publics)
    ok = ctypes.c_int(0)
    _check(lib().zk_verify(circuit.h, u64p(p), n, ctypes.byref(_proof_arg(proof)), ctypes.byref(ok)))
    return bool(ok.value)


def verify_zkey(zkey_path, publics, proof):
    """Pairing check against the verification key of a .zkey file (sections 2 and 3)."""
    p, n = _publics_arg(publics)
    ok = ctypes.c_int(0)
    _check(lib().zk_verify_zkey(zkey_path.encode(), u64p(p), n, ctypes.byref(_proof_arg(proof)), ctypes.byref(ok)))
    return bool(ok.value)


def verify_json(zkey_path, public_json, proof_json):
    """Pairing check from public.json and proof.json texts (replaces snarkjs groth16 verify)."""
    ok = ctypes.c_int(0)
    _check(lib().zk_verify_json(zkey_path.encode(), public_json.encode(), proof_json.encode(), ctypes.byref(ok)))
    return bool(ok.value)


def pairing(p, q):
    """e(P, Q) as 48 words (Fp12, Montgomery) for G1 P (8 words) and G2 Q (16 words)."""
    pp = np.ascontiguousarray(p, dtype=np.uint64).reshape(8)
    qq = np.ascontiguousarray(q, dtype=np.uint64).reshape(16)
    out = np.zeros(48, np.uint64)
    _check(lib().zk_pairing(u64p(pp), u64p(qq), u64p(out)))
    return out
