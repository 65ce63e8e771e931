"""The zkprover CLI (reference contract: prover <zkey> <wtns> <proof.json>
<public.json>, src/main.cu:115-171) on the MI355X: proof.json and public.json
bytes against the oracle's proof of the same circuit."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import zyklop_amd as zk

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "zyklop_amd", "zkprover")


def dec(limbs):
    return str(O.from_limbs(limbs))


def test_cli_matches_oracle(tmp_path):
    c = O.Circuit(61, 500, 7, with_crs=True)
    zkey, wtns = str(tmp_path / "c.zkey"), str(tmp_path / "c.wtns")
    c.write(zkey, wtns)
    pj, pub = str(tmp_path / "proof.json"), str(tmp_path / "public.json")
    r = subprocess.run([CLI, zkey, wtns, pj, pub], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    z = np.zeros(4, np.uint64)
    words = c.prove(z, z)  # the reference's unblinded proof
    canon = O.field_op(O.FP, O.FROM_MONT, words.reshape(8, 4))
    d = [dec(x) for x in canon]
    expect = ('{\n'
              f' "pi_a": ["{d[0]}","{d[1]}","1"],\n'
              f' "pi_b": [["{d[2]}","{d[3]}"],["{d[4]}","{d[5]}"],["1","0"]],\n'
              f' "pi_c": ["{d[6]}","{d[7]}","1"],\n'
              ' "protocol": "groth16",\n'
              ' "curve": "bn128"\n'
              '}\n')
    assert open(pj).read() == expect
    w = c.witness()
    assert json.loads(open(pub).read()) == [dec(w[i]) for i in range(1, c.n_pub + 1)]
    # blinded variant verifies with the trapdoor
    r2 = subprocess.run([CLI, zkey, wtns, pj, pub, "--blind", "5"], capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    rr, ss = zk.synth_blinding(5)
    words_b = c.prove(rr, ss)
    canon_b = O.field_op(O.FP, O.FROM_MONT, words_b.reshape(8, 4))
    assert json.loads(open(pj).read())["pi_c"][:2] == [dec(canon_b[6]), dec(canon_b[7])]
    # the CLI's own pairing check accepts its outputs and rejects a changed statement
    v = subprocess.run([CLI, "verify", zkey, pub, pj], capture_output=True, text=True, timeout=120)
    assert v.returncode == 0 and v.stdout.strip() == "OK", v.stderr
    vals = json.loads(open(pub).read())
    vals[0] = str((int(vals[0]) + 1) % O.R)
    bad_pub = str(tmp_path / "bad_public.json")
    open(bad_pub, "w").write(json.dumps(vals))
    v = subprocess.run([CLI, "verify", zkey, bad_pub, pj], capture_output=True, text=True, timeout=120)
    assert v.returncode == 2 and v.stdout.strip() == "INVALID"


def test_cli_reports_bad_input(tmp_path):
    bad = tmp_path / "bad.zkey"
    bad.write_bytes(b"not a zkey")
    r = subprocess.run([CLI, str(bad), str(bad), str(tmp_path / "p"), str(tmp_path / "q")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "failed" in r.stderr
