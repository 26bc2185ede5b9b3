"""Test-side codec of the reference's wire types, written from serde's data
model and bincode 1.3 / serde_json's published encodings, independently of
the library's C++ codec (inspire-rs_amd/csrc/wire.cpp) it checks:

  bincode (default options of bincode::serialize): little-endian fixed-width
  integers, usize as u64, sequences as a u64 length then the elements,
  fixed arrays ([u8; 32]) as their elements, bool as one byte, Option as a
  0/1 byte then the value, unit enum variants as a u32 index, structs as
  their fields in order; a field with skip_serializing_if is not written at
  all when skipped.
  serde_json (compact): objects with the fields in declaration order, no
  whitespace, integers in decimal -- what json.dumps(..., separators=(",",
  ":")) writes for the same Python values.

Type layouts (field order): Poly (src/math/poly.rs:63-77), RlweCiphertext
(src/rlwe/types.rs:60-66), SeededRlweCiphertext (rlwe/types.rs:165-171),
RgswCiphertext / SeededRgswCiphertext (src/rgsw/types.rs:104-110, 225-231),
GadgetVector (rgsw/types.rs:37-45), ClientQuery / SeededClientQuery
(src/pir/query.rs:119-150), ClientPackingKeys (src/inspiring/inspiring2.rs:
603-628), ServerResponse (src/pir/respond.rs:30-37).
"""
from __future__ import annotations

import json
import struct

import numpy as np


def poly_obj(words: np.ndarray, moduli, dim: int) -> dict:
    """Poly as serde sees it; words = residue-major coefficients (K x dim)."""
    q = 1
    for m in moduli:
        q *= int(m)
    inv = pow(int(moduli[0]), -1, int(moduli[1])) if len(moduli) == 2 else 0
    return {"coeffs": [int(x) for x in np.asarray(words).reshape(-1)], "moduli": [int(m) for m in moduli], "q": q,
            "dim": dim, "crt_q0_inv_mod_q1": inv, "is_ntt": False}


def rlwe_obj(a, b, moduli, d: int, b_dim: int = None) -> dict:
    return {"a": poly_obj(a, moduli, d), "b": poly_obj(b, moduli, b_dim or d)}


def response_obj(ciphertext_a, ciphertext_b, columns, moduli, d: int, b_dim: int = None) -> dict:
    return {"ciphertext": rlwe_obj(ciphertext_a, ciphertext_b, moduli, d, b_dim),
            "column_ciphertexts": [rlwe_obj(c[0], c[1], moduli, d) for c in columns]}


def query_obj(shard_id: int, rows, moduli, d: int, gadget, seeds=None, packing_mode="inspiring", y_body=None,
              num_to_pack: int = 16, with_mode: bool = True) -> dict:
    """ClientQuery (seeds None: rows [2ell][2][K][d]) or SeededClientQuery
    (rows = b polys [2ell][K][d], seeds [2ell][32])."""
    if seeds is None:
        rws = [rlwe_obj(r[0], r[1], moduli, d) for r in rows]
    else:
        rws = [{"seed": [int(x) for x in s], "b": poly_obj(b, moduli, d)} for s, b in zip(seeds, rows)]
    o = {"shard_id": int(shard_id),
         "rgsw_ciphertext": {"rows": rws, "gadget": {"base": gadget[0], "len": gadget[1], "q": gadget[2]}}}
    if with_mode:
        o["packing_mode"] = packing_mode
    if y_body is not None:  # skip_serializing_if = "Option::is_none"
        o["inspiring_packing_keys"] = {"y_body": [poly_obj(y, moduli, d) for y in y_body], "full_key": False,
                                       "num_to_pack": num_to_pack}
    return o


def to_json(obj) -> bytes:
    return json.dumps(obj, separators=(",", ":")).encode()


# -- bincode

def _u64(v):
    return struct.pack("<Q", v)


def bin_poly(p: dict) -> bytes:
    c = np.asarray(p["coeffs"], dtype="<u8")
    return (_u64(len(c)) + c.tobytes() + _u64(len(p["moduli"])) + b"".join(_u64(m) for m in p["moduli"])
            + _u64(p["q"]) + _u64(p["dim"]) + _u64(p["crt_q0_inv_mod_q1"]) + bytes([1 if p["is_ntt"] else 0]))


def bin_rlwe(r: dict) -> bytes:
    return bin_poly(r["a"]) + bin_poly(r["b"])


def bin_response(o: dict) -> bytes:
    cols = o["column_ciphertexts"]
    return bin_rlwe(o["ciphertext"]) + _u64(len(cols)) + b"".join(bin_rlwe(c) for c in cols)


MODES = {"inspiring": 0, "tree": 1}


def bin_query(o: dict) -> bytes:
    rg = o["rgsw_ciphertext"]
    rows = rg["rows"]
    out = struct.pack("<I", o["shard_id"]) + _u64(len(rows))
    for r in rows:
        if "seed" in r:
            out += bytes(r["seed"]) + bin_poly(r["b"])
        else:
            out += bin_rlwe(r)
    g = rg["gadget"]
    out += _u64(g["base"]) + _u64(g["len"]) + _u64(g["q"])
    out += struct.pack("<I", MODES[o.get("packing_mode", "inspiring")])
    keys = o.get("inspiring_packing_keys")
    if keys is not None:
        out += b"\x01" + _u64(len(keys["y_body"])) + b"".join(bin_poly(y) for y in keys["y_body"])
        if keys.get("z_body"):
            out += _u64(len(keys["z_body"])) + b"".join(bin_poly(z) for z in keys["z_body"])
        out += bytes([1 if keys["full_key"] else 0]) + _u64(keys["num_to_pack"])
    return out


class _Reader:
    def __init__(self, b: bytes):
        self.b, self.i = b, 0

    def take(self, n):
        assert self.i + n <= len(self.b), "truncated"
        v = self.b[self.i:self.i + n]
        self.i += n
        return v

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]


def read_bin_response(b: bytes) -> dict:
    r = _Reader(b)

    def poly():
        n = r.u64()
        coeffs = np.frombuffer(r.take(8 * n), dtype="<u8").astype(np.uint64)
        moduli = [r.u64() for _ in range(r.u64())]
        q, dim, inv = r.u64(), r.u64(), r.u64()
        ntt = r.take(1)[0]
        return {"coeffs": coeffs, "moduli": moduli, "q": q, "dim": dim, "crt_q0_inv_mod_q1": inv, "is_ntt": bool(ntt)}

    ct = {"a": poly(), "b": poly()}
    cols = [{"a": poly(), "b": poly()} for _ in range(r.u64())]
    assert r.i == len(b), "trailing bytes"
    return {"ciphertext": ct, "column_ciphertexts": cols}
