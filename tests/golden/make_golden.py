"""Writes tests/golden/golden_vectors.json from the reference's own text.

Run in a container where the reference tree is present:

    python tests/golden/make_golden.py [/root/reference]

* move hash: the six statements of CRITICAL.cu:2-11 are parsed (operator,
  constant, shift direction and amount) and evaluated with Python integers
  modulo 2^32, so the vectors follow the reference's text, not a retyping.
* direction table: the int16 tables of UpdaterGPUScBFM_AB_Type.cu:540-542.
* bond set: the `!set_of_bondvectors` block of tests/LinearChain_l1024_n64.bfm
  (the LeMonADE bond set the reference was run with).
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
MASK = 0xFFFFFFFF


def parse_hash(critical: str):
    body = critical[critical.index("uint32_t hash"):]
    body = body[: body.index("return a;")]
    pat = re.compile(r"a\s*=\s*\(\s*a\s*([+^])\s*(0x[0-9a-fA-F]+)\s*\)\s*([+^])\s*\(\s*a\s*(<<|>>)\s*(\d+)\s*\)\s*;")
    steps = [(m.group(1), int(m.group(2), 16), m.group(3), m.group(4), int(m.group(5))) for m in pat.finditer(body)]
    if len(steps) != 6:
        raise SystemExit(f"expected 6 hash statements, found {len(steps)}")
    return steps


def eval_hash(steps, a: int) -> int:
    for op1, const, op2, sh, amount in steps:
        left = (a + const) & MASK if op1 == "+" else (a ^ const)
        right = (a << amount) & MASK if sh == "<<" else (a >> amount)
        a = (left + right) & MASK if op2 == "+" else (left ^ right)
    return a


def parse_directions(updater: str):
    out = {}
    for axis in "XYZ":
        m = re.search(r"intCUDA\s+tmp_D" + axis + r"TableIntCUDA\[6\]\s*=\s*\{([^}]*)\}", updater)
        out[axis] = [int(v) for v in m.group(1).replace(" ", "").split(",")]
    return [[out["X"][d], out["Y"][d], out["Z"][d]] for d in range(6)]


def parse_bondset(bfm: bytes):
    lines = bfm.split(b"\n")
    i = lines.index(b"!set_of_bondvectors") + 1
    vecs = []
    while lines[i].strip():
        vec, ident = lines[i].rsplit(b":", 1)
        vecs.append([int(t) for t in vec.split()] + [int(ident)])
        i += 1
    return vecs


def main() -> None:
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    critical = (ref / "src/pscbfm/CRITICAL.cu").read_text()
    updater = (ref / "src/pscbfm/UpdaterGPUScBFM_AB_Type.cu").read_text()
    steps = parse_hash(critical)
    inputs = [0, 1, 2, 3, 7, 42, 255, 256, 65535, 65536, 0x7FFFFFFF, 0x80000000, 0xDEADBEEF, 0xFFFFFFFE, 0xFFFFFFFF]
    inputs += [17 * k * k + 3 for k in range(40)]
    # the kernel's composite draw rn = H(H(j) ^ seed), direction = rn % 6
    draws = []
    for seed in (0, 42, 0x9E3779B9, 0xFFFFFFFF):
        for j in range(16):
            rn = eval_hash(steps, eval_hash(steps, j) ^ seed)
            draws.append([j, seed, rn, rn % 6])
    golden = {
        "hash": [[a, eval_hash(steps, a)] for a in inputs],
        "draws": draws,
        "directions": parse_directions(updater),
        "bondset": parse_bondset((ref / "tests/LinearChain_l1024_n64.bfm").read_bytes()),
        "source": "CRITICAL.cu:2-11, CRITICAL.cu:321-323, UpdaterGPUScBFM_AB_Type.cu:540-542, "
                  "tests/LinearChain_l1024_n64.bfm:43-151",
    }
    (HERE / "golden_vectors.json").write_text(json.dumps(golden, indent=1) + "\n")
    print(f"wrote {HERE / 'golden_vectors.json'}")


if __name__ == "__main__":
    main()
