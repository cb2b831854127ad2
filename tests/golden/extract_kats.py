"""Extract the known-answer vectors the reference's own tests hold for this
path into kats.json (data only: inputs and expected outputs).

Sources (read as text from a checkout of the reference at REF):
  crypto/ethrex-crypto/src/tests/keccak_tests.rs:59-77   Keccak-256 KATs
  crypto/crypto/src/tests/default_transcript_tests.rs:10-46  transcript challenges
  crypto/stark/src/tests/grinding_tests.rs               grinding (seed, nonce, g, valid)
  crypto/math/src/field/goldilocks.rs:536-540            two-adicity and 2^32-th root
  prover/src/tables/keccak_rc.rs:97-116, bitwise.rs:199-218,
  page.rs:416-455, cpu_timestamp_commitments.rs         static commitments of the
                                                         preprocessed columns (coset 3,
                                                         blowups 2/4/8)

Usage: python tests/golden/extract_kats.py [REF]   (REF defaults to /root/reference)
"""
from __future__ import annotations

import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kats.json")


def read(rel):
    with open(os.path.join(REF, rel)) as f:
        return f.read()


def keccak_kats():
    src = read("crypto/ethrex-crypto/src/tests/keccak_tests.rs")
    out = []
    for m in re.finditer(r'check_keccak_kat\(\s*b"((?:[^"\\]|\\.)*)",\s*"([0-9a-f]{64})"', src):
        out.append({"input_ascii": m.group(1), "digest_hex": m.group(2)})
    return out


def byte_list(s):
    return [int(x, 16) for x in re.findall(r"0x([0-9a-fA-F]{2})", s)]


def transcript_kats():
    src = read("crypto/crypto/src/tests/default_transcript_tests.rs")
    body = src[src.index("fn basic_challenge"):]
    body = body[: body.index("\n}\n")]
    points = re.findall(r"let point_[a-d]: Vec<u8> = vec!\[([^\]]*)\];", body)
    chals = re.findall(r"assert_eq!\(\s*challenge\d,\s*\[([^\]]*)\]", body)
    return {
        "appends_before_challenge1": [byte_list(points[0]), byte_list(points[1])],
        "challenge1": byte_list(chals[0]),
        "appends_before_challenge2": [byte_list(points[2]), byte_list(points[3])],
        "challenge2": byte_list(chals[1]),
    }


def grinding_kats():
    src = read("crypto/stark/src/tests/grinding_tests.rs")
    out = []
    for m in re.finditer(r"fn (test_\w+)\(\)\s*\{(.*?)\n\}", src, flags=re.S):
        body = m.group(2)
        seed = re.search(r"let seed = \[([^\]]*)\]", body)
        nonce = re.search(r"let nonce = (0x[0-9a-fA-F]+|\d+);", body)
        g = re.search(r"let grinding_factor = (\d+);", body)
        if not (seed and nonce and g):
            continue
        valid = "assert!(is_valid_nonce" in body
        out.append({
            "name": m.group(1),
            "seed": [int(x) for x in re.findall(r"\d+", seed.group(1))],
            "nonce": int(nonce.group(1), 0),
            "grinding_factor": int(g.group(1)),
            "valid": valid,
        })
    return out


def field_constants():
    src = read("crypto/math/src/field/goldilocks.rs")
    two_adicity = int(re.search(r"const TWO_ADICITY: u64 = (\d+);", src).group(1))
    root = int(re.search(r"const TWO_ADIC_PRIMITVE_ROOT_OF_UNITY: u64 = (\d+);", src).group(1))
    return {"two_adicity": two_adicity, "root_of_unity_2_32": root}


def static_commitments(rel, fn="fn static_commitment"):
    """blowup factor -> 32-byte root from a table's static-commitment match
    (`fn static_commitment` unless another function is named)."""
    src = read(rel)
    body = src[src.index(fn):]
    body = body[: body.index("\n}\n")]
    out = {}
    for m in re.finditer(r"(\d+) => Some\(\[([^\]]*)\]\)", body):
        out[m.group(1)] = byte_list(m.group(2))
    return out


def cpu_timestamp_commitments():
    """blowup factor -> {log2 height: root} of the CPU table's fixed TIMESTAMP
    column (4·(row + 1)), from the generated table cpu_timestamp_commitments.rs."""
    src = read("prover/src/tables/cpu_timestamp_commitments.rs")
    out = {}
    for blk in re.finditer(r"\(\s*(\d+),\s*&\[(.*?)\],\s*\),", src, re.S):
        out[blk.group(1)] = {m.group(1): byte_list(m.group(2))
                             for m in re.finditer(r"// 2\^(\d+)\s*\[([^\]]*)\]", blk.group(2))}
    return out


def preprocessed_commitments():
    return {"coset_offset": 3,
            "keccak_rc": static_commitments("prover/src/tables/keccak_rc.rs"),
            "bitwise": static_commitments("prover/src/tables/bitwise.rs"),
            # page.rs:416-455: OFFSET + zero INIT of a zero-initialised page, and
            # OFFSET alone (private-input pages; also GLOBAL_MEMORY's)
            "page_zero": static_commitments("prover/src/tables/page.rs", "fn static_zero_page_commitment"),
            "page_private": static_commitments("prover/src/tables/page.rs", "fn static_private_page_commitment"),
            # cpu.rs:161-165 (TIMESTAMP = 4·(row + 1)), table cpu_timestamp_commitments.rs
            "cpu_timestamp": cpu_timestamp_commitments()}


def main():
    data = {
        "keccak256": keccak_kats(),
        "transcript_basic_challenge": transcript_kats(),
        "grinding": grinding_kats(),
        "goldilocks": field_constants(),
        "preprocessed_commitments": preprocessed_commitments(),
    }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print(f"wrote {OUT}: {len(data['keccak256'])} keccak, {len(data['grinding'])} grinding vectors")


if __name__ == "__main__":
    main()
