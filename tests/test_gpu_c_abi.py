"""The reference C ABI driven from a C program (tests/c_abi_roundtrip.c, the
checks of capi/test/roundtrip.c:102-222 restated): every check it makes
passes, and every stream it wrote through the reference-named entry points
(which run at Config::default, Level::Max: capi/src/lib.rs:29-34,
capi/src/typed.rs:71-73) is the oracle's stream for the same input."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "_build", "c_abi_roundtrip")

# typed-API dtype -> (numpy input type, lane type the typed ABI widens to,
# engine dtype) (capi/src/typed.rs:141-168)
TYPED = {4: (np.int64, np.int64, orc.I64), 9: (np.float32, np.float32, orc.F32),
         3: (np.int32, np.int32, orc.I32), 2: (np.int16, np.int32, orc.I32),
         1: (np.int8, np.int32, orc.I32), 6: (np.uint16, np.uint32, orc.U32),
         5: (np.uint8, np.uint32, orc.U32), 0: (np.uint8, np.uint32, orc.U32)}


def test_c_program_checks_and_streams(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/_build/c_abi_roundtrip is not built (build() compiles it)")
    print(f"case c_abi_roundtrip -> {tmp_path}", flush=True)
    p = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(p.stdout[-4000:], p.stderr[-2000:], flush=True)
    assert p.returncode == 0, f"{p.returncode} failed checks"
    cases = [ln.split() for ln in p.stdout.splitlines() if ln.startswith("CASE ")]
    assert len(cases) >= 17
    for c in cases:
        name, api, dtype, n = c[1], c[2], int(c[3]), int(c[4])
        raw = (tmp_path / f"{name}.in").read_bytes()
        got = (tmp_path / f"{name}.qs").read_bytes()
        if api == "f64":
            want = orc.compress(np.frombuffer(raw, dtype=np.float64), orc.F64, None, orc.MAX,
                                threads=16)
        elif api == "typed":
            src, lane, odt = TYPED[dtype]
            vals = np.frombuffer(raw, dtype=src).astype(lane)
            vf = tmp_path / f"{name}.valid"
            valid = np.frombuffer(vf.read_bytes(), dtype=np.uint8) if vf.exists() else None
            want = orc.compress(vals, odt, valid, orc.MAX, threads=16)
        else:
            scale, prec = int(c[5]), int(c[6])
            w = {12: 8, 13: 16, 14: 32}[dtype]
            wire, ww = (7, 32) if dtype == 14 else (6, 16)
            sig = [int.from_bytes(raw[i * w:(i + 1) * w], "little", signed=True) for i in range(n)]
            wide = b"".join(v.to_bytes(ww, "little", signed=True) for v in sig)
            want = orc.compress_decimal(wide, wire, None, scale, prec, level=orc.MAX)
        assert got == want, f"{name}: stream differs from the oracle ({len(got)} vs {len(want)} B)"
