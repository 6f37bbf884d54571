"""The plan JSON parser (proteus_amd/engine/json.cpp) under AddressSanitizer
and UBSan on the CPU: the reference's plan fixtures, their truncations and
byte mutations, deep nesting and odd literals are parsed without a memory
error, and every accepted document re-parses from its own dump."""
import os
import random
import shutil
import struct
import subprocess

import pytest

from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _inputs():
    plans = os.path.join(H.GOLDEN, "inputs", "plans")
    docs = []
    for name in sorted(os.listdir(plans)):
        if name.endswith(".json"):
            with open(os.path.join(plans, name), "rb") as f:
                docs.append(f.read())
    rs = random.Random(7)
    out = list(docs)
    for d in docs:
        for _ in range(40):
            out.append(d[:rs.randrange(len(d) + 1)])          # truncation
            b = bytearray(d)
            for _ in range(rs.randrange(1, 8)):                # byte mutations
                b[rs.randrange(len(b))] = rs.randrange(256)
            out.append(bytes(b))
    out += [b"[" * 300 + b"]" * 300, b"[" * 200 + b"]" * 200, b'{"a":' * 100 + b"1" + b"}" * 100,
            b'"\\u00e9\\ud83d"', b'"\\u12"', b"1e999999", b"-", b"-0.0e-0", b"123456789012345678901234",
            b'{"a":1,}', b"[1,2,", b"\x00", b"", b'"\\', b'"' + b"x" * 100000 + b'"',
            b"tru", b"nulll", b'{"a" 1}', b'{1:2}', b"[-]", b"0x10", b"01"]
    return out


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_parser_survives_malformed_plans(tmp_path):
    exe = str(tmp_path / "fuzz")
    src = [os.path.join(ROOT, "tests", "json_fuzz_driver.cpp"),
           os.path.join(ROOT, "proteus_amd", "engine", "json.cpp")]
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=undefined", "-o", exe] + src)
    data = tmp_path / "inputs.bin"
    with open(data, "wb") as f:
        for d in _inputs():
            f.write(struct.pack("<I", len(d)) + d)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    r = subprocess.run([exe, str(data)], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    acc, rej = (int(x) for x in r.stdout.split()[1::2])
    assert acc >= 10 and rej >= 100
