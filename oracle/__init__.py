"""ctypes binding of the C restatement in oracle/oracle.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, where it is the checker (or the timed CPU
baseline).  The product in inspire-rs_amd/ never imports it.

Arrays follow the reference's `Poly.coeffs` layout (src/math/poly.rs:63-77):
a polynomial is uint64[K, d] residue-major; an RLWE ciphertext is
uint64[2, K, d] (a, b); an RGSW ciphertext is uint64[2*ell, 2, K, d].
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

Q0, Q1 = 268369921, 249561089  # DEFAULT_CRT_MODULI, src/params.rs:31
DEFAULT_Q = 1152921504606830593  # src/math/mod_q.rs:37


class OrParams(C.Structure):
    _fields_ = [
        ("d", C.c_uint32),
        ("K", C.c_uint32),
        ("moduli", C.c_uint64 * 2),
        ("q", C.c_uint64),
        ("p", C.c_uint64),
        ("gadget_base", C.c_uint64),
        ("gadget_len", C.c_uint32),
    ]


class OrChacha(C.Structure):
    _fields_ = [
        ("key", C.c_uint32 * 8),
        ("counter", C.c_uint64),
        ("buf", C.c_uint32 * 64),
        ("index", C.c_uint32),
    ]


_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)


def build():
    """Compile oracle/liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    P = C.POINTER(OrParams)
    sig = {
        "or_params_validate": (C.c_int, [P]),
        "or_delta": (C.c_uint64, [P]),
        "or_ntt_new": (C.c_void_p, [C.c_uint32, _u64p, C.c_uint32]),
        "or_ntt_free": (None, [C.c_void_p]),
        "or_ntt_psi": (C.c_uint64, [C.c_void_p, C.c_uint32]),
        "or_ntt_forward": (None, [C.c_void_p, _u64p]),
        "or_ntt_inverse": (None, [C.c_void_p, _u64p]),
        "or_ntt_twiddles": (None, [C.c_void_p, C.c_uint32, _u64p, _u64p]),
        "or_mont_mul": (C.c_uint64, [C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint64]),
        "or_mod_inverse": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "or_crt_compose_2": (C.c_uint64, [C.c_uint64] * 5),
        "or_poly_coeff": (C.c_uint64, [P, _u64p, C.c_uint32]),
        "or_poly_from_coeffs": (None, [P, _u64p, _u64p]),
        "or_poly_mul_ntt": (None, [P, C.c_void_p, _u64p, _u64p, _u64p]),
        "or_poly_negacyclic_schoolbook": (None, [P, _u64p, _u64p, _u64p]),
        "or_gadget_decompose": (None, [P, _u64p, _u64p]),
        "or_gadget_reconstruct": (None, [P, _u64p, _u64p]),
        "or_external_product": (None, [P, C.c_void_p, _u64p, _u64p, _u64p, C.c_int]),
        "or_respond": (None, [P, C.c_void_p, _u64p, C.c_uint32, _u64p, _u64p, _u64p, C.c_int]),
        "or_pack_generator": (C.c_uint32, [C.c_uint32, C.c_uint32]),
        "or_automorph_table": (None, [C.c_uint32, C.c_uint32, _u32p]),
        "or_automorph_table_bruteforce": (C.c_int, [P, C.c_void_p, C.c_uint32, C.c_uint64, _u32p]),
        "or_apply_automorphism": (None, [P, _u64p, C.c_uint32, _u64p]),
        "or_apply_automorphism_ntt": (None, [P, _u64p, _u32p, _u64p]),
        "or_generate_mask": (None, [P, _u8p, _u64p]),
        "or_pack_new": (C.c_void_p, [P, C.c_void_p, C.c_uint32, _u8p]),
        "or_pack_new_mode": (C.c_void_p, [P, C.c_void_p, C.c_uint32, _u8p, C.c_int]),
        "or_pack_free": (None, [C.c_void_p]),
        "or_pack_w_all_ntt": (_u64p, [C.c_void_p]),
        "or_packing_offline": (None, [C.c_void_p, _u64p, _u64p, _u64p]),
        "or_generate_rotations": (None, [C.c_void_p, _u64p, _u64p]),
        "or_packing_online": (None, [C.c_void_p, _u64p, _u64p, _u64p, C.c_uint32, _u64p, _u64p]),
        "or_respond_inspiring": (
            C.c_int,
            [P, C.c_void_p, C.c_void_p, _u64p, C.c_uint32, _u64p, _u64p, _u64p, C.c_uint32, _u64p],
        ),
        "or_respond_inspiring_many": (
            None,
            [P, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(_u64p), C.c_uint32, C.POINTER(_u64p),
             C.POINTER(_u64p), C.POINTER(_u64p), C.POINTER(C.c_int)],
        ),
        "or_respond_inspiring_literal": (
            C.c_int, [P, C.c_void_p, _u64p, C.c_uint32, _u64p, _u64p, _u64p, _u64p],
        ),
        "or_respond_inspiring_literal_many": (
            None,
            [P, C.c_void_p, C.c_uint32, C.POINTER(_u64p), C.c_uint32, C.POINTER(_u64p), C.POINTER(_u64p),
             C.POINTER(_u64p), C.POINTER(C.c_int)],
        ),
        "or_num_threads": (C.c_int, []),
        "or_sample_extract_coeff0": (None, [P, _u64p, _u64p, _u64p]),
        "or_homomorphic_automorph": (None, [P, C.c_void_p, _u64p, C.c_uint32, _u64p, C.c_uint32, _u64p]),
        "or_pack_lwes": (
            C.c_uint32,
            [P, C.c_void_p, C.c_uint32, _u64p, _u64p, _u64p, C.c_uint32, C.c_int, C.c_int, _u64p],
        ),
        "or_respond_one_packing": (
            C.c_uint32,
            [P, C.c_void_p, _u64p, C.c_uint32, _u64p, _u64p, C.c_uint32, C.c_int, C.c_int, _u64p],
        ),
        "or_respond_one_packing_many": (
            None,
            [P, C.c_void_p, C.c_uint32, C.POINTER(_u64p), C.c_uint32, C.POINTER(_u64p), _u64p, C.c_uint32,
             C.POINTER(_u64p)],
        ),
        "or_automorphism_ks_matrix": (
            None, [P, C.c_void_p, _u64p, C.c_uint32, C.POINTER(OrChacha), C.c_double, _u64p]),
        "or_galois_keys": (None, [P, C.c_void_p, _u64p, C.POINTER(OrChacha), C.c_double, _u64p]),
        "or_chacha20_block": (None, [_u32p, C.c_uint64, C.c_uint64, _u32p]),
        "or_chacha_seed": (None, [C.POINTER(OrChacha), _u8p]),
        "or_chacha_next_u32": (C.c_uint32, [C.POINTER(OrChacha)]),
        "or_chacha_next_u64": (C.c_uint64, [C.POINTER(OrChacha)]),
        "or_gen_range": (C.c_uint64, [C.POINTER(OrChacha), C.c_uint64]),
        "or_zone_accept": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint, C.POINTER(C.c_uint64)]),
        "or_zone_counts": (None, [C.c_uint64, C.c_uint, C.POINTER(C.c_uint64)]),
        "or_poly_from_seed": (None, [P, _u8p, _u64p]),
        "or_seeded_rgsw_expand": (None, [P, _u8p, _u64p, _u64p]),
        "or_sample_gaussian_poly": (None, [P, C.POINTER(OrChacha), C.c_double, _u64p]),
        "or_rgsw_encrypt": (None, [P, C.c_void_p, _u64p, _u64p, C.POINTER(OrChacha), C.c_double, _u64p]),
        "or_seeded_rgsw_encrypt": (
            None,
            [P, C.c_void_p, _u64p, _u64p, C.POINTER(OrChacha), C.c_double, _u8p, _u64p],
        ),
        "or_inverse_monomial": (None, [P, C.c_uint32, _u64p]),
        "or_client_packing_keys": (
            None,
            [P, C.c_void_p, C.c_void_p, _u64p, C.POINTER(OrChacha), C.c_double, _u64p, _u64p],
        ),
        "or_decrypt": (None, [P, C.c_void_p, _u64p, _u64p, _u64p]),
        "or_encode_shard": (
            None,
            [P, _u8p, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, _u64p],
        ),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _p64(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_u64p)


def _p32(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def _p8(a) -> C.POINTER(C.c_uint8):
    if isinstance(a, (bytes, bytearray)):
        a = np.frombuffer(bytes(a), dtype=np.uint8).copy()
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u8p), a


@dataclass
class Params:
    """InspireParams (src/params.rs:134-185), the fields respond reads."""

    d: int = 2048
    moduli: tuple = (Q0, Q1)
    p: int = 65536
    gadget_base: int = 1 << 20
    gadget_len: int = 3

    @property
    def K(self) -> int:
        return len(self.moduli)

    @property
    def q(self) -> int:
        q = 1
        for m in self.moduli:
            q *= m
        return q

    @property
    def delta(self) -> int:
        return self.q // self.p

    def c(self) -> OrParams:
        m = list(self.moduli) + [0] * (2 - len(self.moduli))
        return OrParams(self.d, self.K, (C.c_uint64 * 2)(*m), self.q, self.p, self.gadget_base, self.gadget_len)


def baseline_params() -> Params:
    """d=2048, CRT moduli [268369921, 249561089], p=2^16, z=2^20, ell=3."""
    return Params()


def single_modulus_params(d: int = 256) -> Params:
    """The reference tests' parameter set (tests/e2e_pir.rs:12-23)."""
    return Params(d=d, moduli=(DEFAULT_Q,), p=65536)


class Oracle:
    """Holds an NttContext for one parameter set plus helpers over numpy arrays."""

    def __init__(self, params: Params, residue_wise: bool = False):
        self.params = params
        self.residue_wise = bool(residue_wise)
        self.cp = params.c()
        m = np.array(list(params.moduli), dtype=np.uint64)
        self.ntt = lib.or_ntt_new(params.d, _p64(m), params.K)
        self._packs = {}

    def __del__(self):
        try:
            for pk in self._packs.values():
                lib.or_pack_free(pk)
            lib.or_ntt_free(self.ntt)
        except Exception:
            pass

    # -- shapes
    @property
    def P(self):
        return (self.params.K, self.params.d)

    def zeros(self, *lead):
        return np.zeros(tuple(lead) + self.P, dtype=np.uint64)

    # -- NTT / poly
    def ntt_forward(self, poly):
        out = np.ascontiguousarray(poly, dtype=np.uint64).copy()
        for x in out.reshape(-1, *self.P):
            lib.or_ntt_forward(self.ntt, _p64(x))
        return out

    def ntt_inverse(self, poly):
        out = np.ascontiguousarray(poly, dtype=np.uint64).copy()
        for x in out.reshape(-1, *self.P):
            lib.or_ntt_inverse(self.ntt, _p64(x))
        return out

    def psi(self, k):
        return lib.or_ntt_psi(self.ntt, k)

    def twiddles(self, k):
        a = np.zeros(self.params.d, dtype=np.uint64)
        b = np.zeros(self.params.d, dtype=np.uint64)
        lib.or_ntt_twiddles(self.ntt, k, _p64(a), _p64(b))
        return a, b

    def mul_ntt(self, a, b):
        out = self.zeros()
        lib.or_poly_mul_ntt(C.byref(self.cp), self.ntt, _p64(np.ascontiguousarray(a)), _p64(np.ascontiguousarray(b)), _p64(out))
        return out

    def schoolbook(self, a, b):
        out = self.zeros()
        lib.or_poly_negacyclic_schoolbook(C.byref(self.cp), _p64(np.ascontiguousarray(a)), _p64(np.ascontiguousarray(b)), _p64(out))
        return out

    def from_coeffs(self, composite):
        c = np.ascontiguousarray(composite, dtype=np.uint64)
        out = self.zeros()
        lib.or_poly_from_coeffs(C.byref(self.cp), _p64(c), _p64(out))
        return out

    def coeffs(self, poly):
        poly = np.ascontiguousarray(poly, dtype=np.uint64)
        return np.array([lib.or_poly_coeff(C.byref(self.cp), _p64(poly), i) for i in range(self.params.d)], dtype=np.uint64)

    def gadget_decompose(self, poly):
        out = self.zeros(self.params.gadget_len)
        lib.or_gadget_decompose(C.byref(self.cp), _p64(np.ascontiguousarray(poly)), _p64(out))
        return out

    def gadget_reconstruct(self, digits):
        out = self.zeros()
        lib.or_gadget_reconstruct(C.byref(self.cp), _p64(np.ascontiguousarray(digits)), _p64(out))
        return out

    def external_product(self, rlwe, rgsw, literal=False):
        out = self.zeros(2)
        lib.or_external_product(C.byref(self.cp), self.ntt, _p64(np.ascontiguousarray(rlwe)), _p64(np.ascontiguousarray(rgsw)), _p64(out), int(literal))
        return out

    # -- respond
    def respond(self, shard_polys, rgsw, literal=False):
        """InsPIRe^0 respond (respond.rs:71-115) -> (ciphertext, column_ciphertexts)."""
        shard_polys = np.ascontiguousarray(shard_polys, dtype=np.uint64)
        ncols = shard_polys.shape[0] if shard_polys.size else 0
        ct = self.zeros(2)
        cols = self.zeros(max(ncols, 1), 2)
        lib.or_respond(C.byref(self.cp), self.ntt, _p64(shard_polys) if ncols else None, ncols,
                       _p64(np.ascontiguousarray(rgsw)), _p64(ct), _p64(cols), int(literal))
        if ncols == 0:
            return ct, self.zeros(1, 2)
        return ct, cols

    def pack(self, gamma: int, w_seed: bytes):
        key = (gamma, bytes(w_seed))
        if key not in self._packs:
            ptr, keep = _p8(w_seed)
            self._packs[key] = lib.or_pack_new_mode(C.byref(self.cp), self.ntt, gamma, ptr, int(self.residue_wise))
        return self._packs[key]

    def w_all_ntt(self, gamma, w_seed):
        pk = self.pack(gamma, w_seed)
        n = (gamma - 1) * self.params.gadget_len * self.params.K * self.params.d
        ptr = lib.or_pack_w_all_ntt(pk)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy().reshape(gamma - 1, self.params.gadget_len, *self.P)

    def packing_offline(self, gamma, w_seed, a_ct):
        pk = self.pack(gamma, w_seed)
        a_hat = self.zeros()
        bold_t = self.zeros(max(gamma - 1, 1), self.params.gadget_len)
        lib.or_packing_offline(pk, _p64(np.ascontiguousarray(a_ct, dtype=np.uint64)), _p64(a_hat), _p64(bold_t))
        return a_hat, bold_t[: gamma - 1]

    def generate_rotations(self, gamma, w_seed, y_body):
        pk = self.pack(gamma, w_seed)
        y_all = self.zeros(max(gamma - 1, 1), self.params.gadget_len)
        lib.or_generate_rotations(pk, _p64(np.ascontiguousarray(y_body, dtype=np.uint64)), _p64(y_all))
        return y_all[: gamma - 1]

    def respond_inspiring(self, w_seed, shard_polys, rgsw, y_body=None, y_all=None):
        """respond_inspiring (respond.rs:284-386) -> (rc, ciphertext)."""
        shard_polys = np.ascontiguousarray(shard_polys, dtype=np.uint64)
        ncols = shard_polys.shape[0] if shard_polys.size else 0
        pk = self.pack(max(ncols, 1), w_seed)
        ct = self.zeros(2)
        yb = _p64(np.ascontiguousarray(y_body, dtype=np.uint64)) if y_body is not None else None
        if y_all is not None and len(y_all):
            ya_arr = np.ascontiguousarray(y_all, dtype=np.uint64)
            ya, ylen = _p64(ya_arr), ya_arr.shape[0]
        else:
            ya, ylen = None, 0
        rc = lib.or_respond_inspiring(C.byref(self.cp), self.ntt, pk, _p64(shard_polys) if ncols else None, ncols,
                                      _p64(np.ascontiguousarray(rgsw)), yb, ya, ylen, _p64(ct))
        return rc, ct

    def respond_inspiring_many(self, w_seed, shards, rgsws, y_bodies):
        """Many y_body respond_inspiring calls, one per OpenMP thread at a time."""
        n = len(shards)
        ncols = shards[0].shape[0]
        pk = self.pack(ncols, w_seed)
        keep = [np.ascontiguousarray(x, dtype=np.uint64) for x in list(shards) + list(rgsws) + list(y_bodies)]
        outs = [self.zeros(2) for _ in range(n)]
        arr = lambda xs: (_u64p * n)(*[_p64(x) for x in xs])
        rc = (C.c_int * n)()
        lib.or_respond_inspiring_many(C.byref(self.cp), self.ntt, pk, n, arr(keep[:n]), ncols, arr(keep[n:2 * n]),
                                      arr(keep[2 * n:]), arr(outs), rc)
        return list(rc), outs

    def respond_inspiring_literal(self, w_seed, shard_polys, rgsw, y_body):
        """respond_inspiring with nothing hoisted (per query: NTT contexts,
        literal external products, PackParams::new with every brute-force
        automorphism table, OfflinePackingKeys::generate) -> (rc, ct, work);
        work = (brute-force retries, tables built)."""
        shard_polys = np.ascontiguousarray(shard_polys, dtype=np.uint64)
        ptr, keep = _p8(w_seed)
        ct = self.zeros(2)
        work = np.zeros(2, dtype=np.uint64)
        rc = lib.or_respond_inspiring_literal(C.byref(self.cp), ptr, _p64(shard_polys), shard_polys.shape[0],
                                              _p64(np.ascontiguousarray(rgsw)),
                                              _p64(np.ascontiguousarray(y_body, dtype=np.uint64)), _p64(ct),
                                              _p64(work))
        return rc, ct, (int(work[0]), int(work[1]))

    def respond_inspiring_literal_many(self, w_seed, shards, rgsws, y_bodies):
        """Many literal respond_inspiring calls, one per OpenMP thread at a time."""
        n = len(shards)
        ncols = shards[0].shape[0]
        ptr, keep_seed = _p8(w_seed)
        keep = [np.ascontiguousarray(x, dtype=np.uint64) for x in list(shards) + list(rgsws) + list(y_bodies)]
        outs = [self.zeros(2) for _ in range(n)]
        arr = lambda xs: (_u64p * n)(*[_p64(x) for x in xs])
        rc = (C.c_int * n)()
        lib.or_respond_inspiring_literal_many(C.byref(self.cp), ptr, n, arr(keep[:n]), ncols, arr(keep[n:2 * n]),
                                              arr(keep[2 * n:]), arr(outs), rc)
        return list(rc), outs

    # -- tree packing (inspiring/automorph_pack.rs)
    @property
    def log_d(self):
        return self.params.d.bit_length() - 1

    def sample_extract(self, ct):
        """sample_extract_coeff0 (rlwe/enc.rs:195-208) -> (a [d] composite, b composite)."""
        a = np.zeros(self.params.d, dtype=np.uint64)
        b = np.zeros(1, dtype=np.uint64)
        lib.or_sample_extract_coeff0(C.byref(self.cp), _p64(np.ascontiguousarray(ct, dtype=np.uint64)), _p64(a),
                                     _p64(b))
        return a, int(b[0])

    def homomorphic_automorph(self, ct, t, ks):
        ks = np.ascontiguousarray(ks, dtype=np.uint64)
        out = self.zeros(2)
        lib.or_homomorphic_automorph(C.byref(self.cp), self.ntt, _p64(np.ascontiguousarray(ct, dtype=np.uint64)), t,
                                     _p64(ks), ks.shape[0], _p64(out))
        return out

    def _tree_out(self, words, bdim):
        K, d = self.params.K, self.params.d
        return words[:K * d].reshape(K, d).copy(), words[K * d:K * d + K * bdim].reshape(K, bdim).copy()

    def pack_lwes(self, lwe_a, lwe_b, galois_keys, literal=False, residue_wise=None):
        """pack_lwes (automorph_pack.rs:418-487) -> (a [K][d], b [K][bdim])."""
        rw = self.residue_wise if residue_wise is None else residue_wise
        K, d = self.params.K, self.params.d
        la = np.ascontiguousarray(lwe_a, dtype=np.uint64).reshape(-1, d)
        lb = np.ascontiguousarray(lwe_b, dtype=np.uint64).reshape(-1)
        gk = np.ascontiguousarray(galois_keys, dtype=np.uint64)
        words = np.zeros(3 * K * d, dtype=np.uint64)
        bdim = lib.or_pack_lwes(C.byref(self.cp), self.ntt, la.shape[0], _p64(la), _p64(lb), _p64(gk), gk.shape[0],
                                int(literal), int(rw), _p64(words))
        return self._tree_out(words, bdim)

    def respond_one_packing(self, shard_polys, rgsw, galois_keys, literal=False, residue_wise=None):
        """respond_one_packing (respond.rs:213-268) -> (a [K][d], b [K][bdim]);
        an empty shard gives the zero ciphertext (bdim = d)."""
        rw = self.residue_wise if residue_wise is None else residue_wise
        K, d = self.params.K, self.params.d
        shard_polys = np.ascontiguousarray(shard_polys, dtype=np.uint64)
        ncols = shard_polys.shape[0] if shard_polys.size else 0
        gk = np.ascontiguousarray(galois_keys, dtype=np.uint64)
        words = np.zeros(3 * K * d, dtype=np.uint64)
        bdim = lib.or_respond_one_packing(C.byref(self.cp), self.ntt, _p64(shard_polys) if ncols else None, ncols,
                                          _p64(np.ascontiguousarray(rgsw, dtype=np.uint64)), _p64(gk), gk.shape[0],
                                          int(literal), int(rw), _p64(words))
        return self._tree_out(words, bdim or d)

    def respond_one_packing_many(self, shards, rgsws, galois_keys):
        """Many respond_one_packing calls (reference b layout), one per OpenMP
        thread at a time -> list of flat ciphertext words (a then b)."""
        n = len(shards)
        K, d = self.params.K, self.params.d
        ncols = shards[0].shape[0]
        keep = [np.ascontiguousarray(x, dtype=np.uint64) for x in list(shards) + list(rgsws)]
        gk = np.ascontiguousarray(galois_keys, dtype=np.uint64)
        outs = [np.zeros(3 * K * d, dtype=np.uint64) for _ in range(n)]
        arr = lambda xs: (_u64p * n)(*[_p64(x) for x in xs])
        lib.or_respond_one_packing_many(C.byref(self.cp), self.ntt, n, arr(keep[:n]), ncols, arr(keep[n:]), _p64(gk),
                                        gk.shape[0], arr(outs))
        bdim = 2 * d if (K == 2 and not self.residue_wise and ncols >= 2) else d
        return [o[:K * d + K * bdim] for o in outs]

    def automorphism_ks_matrix(self, sk, t, r, sigma=6.4):
        out = self.zeros(self.params.gadget_len, 2)
        lib.or_automorphism_ks_matrix(C.byref(self.cp), self.ntt, _p64(np.ascontiguousarray(sk)), t, C.byref(r), sigma,
                                      _p64(out))
        return out

    def galois_keys(self, sk, r, sigma=6.4):
        """ServerCrs.galois_keys (setup.rs:167-173): [log_d][ell][2][K][d]."""
        out = self.zeros(self.log_d, self.params.gadget_len, 2)
        lib.or_galois_keys(C.byref(self.cp), self.ntt, _p64(np.ascontiguousarray(sk)), C.byref(r), sigma, _p64(out))
        return out

    # -- automorphisms
    def automorph_table(self, t):
        tab = np.zeros(self.params.d, dtype=np.uint32)
        lib.or_automorph_table(self.params.d, t, _p32(tab))
        return tab

    def automorph_table_bruteforce(self, t, seed=1):
        tab = np.zeros(self.params.d, dtype=np.uint32)
        lib.or_automorph_table_bruteforce(C.byref(self.cp), self.ntt, t, seed, _p32(tab))
        return tab

    def apply_automorphism(self, poly, g):
        out = self.zeros()
        lib.or_apply_automorphism(C.byref(self.cp), _p64(np.ascontiguousarray(poly, dtype=np.uint64)), g, _p64(out))
        return out

    def apply_automorphism_ntt(self, poly_ntt, table):
        out = self.zeros()
        lib.or_apply_automorphism_ntt(C.byref(self.cp), _p64(np.ascontiguousarray(poly_ntt, dtype=np.uint64)), _p32(np.ascontiguousarray(table, dtype=np.uint32)), _p64(out))
        return out

    def generate_mask(self, seed):
        out = self.zeros(self.params.gadget_len)
        ptr, keep = _p8(seed)
        lib.or_generate_mask(C.byref(self.cp), ptr, _p64(out))
        return out

    # -- seeds
    def poly_from_seed(self, seed):
        out = self.zeros()
        ptr, keep = _p8(seed)
        lib.or_poly_from_seed(C.byref(self.cp), ptr, _p64(out))
        return out

    def seeded_rgsw_expand(self, seeds, b):
        seeds = np.ascontiguousarray(seeds, dtype=np.uint8)
        out = self.zeros(2 * self.params.gadget_len, 2)
        lib.or_seeded_rgsw_expand(C.byref(self.cp), seeds.ctypes.data_as(_u8p), _p64(np.ascontiguousarray(b, dtype=np.uint64)), _p64(out))
        return out

    # -- client side (fixtures / decrypt checks)
    def rng(self, seed: bytes) -> OrChacha:
        r = OrChacha()
        ptr, keep = _p8(seed)
        lib.or_chacha_seed(C.byref(r), ptr)
        return r

    def gaussian_poly(self, r: OrChacha, sigma=6.4):
        out = self.zeros()
        lib.or_sample_gaussian_poly(C.byref(self.cp), C.byref(r), sigma, _p64(out))
        return out

    def inverse_monomial(self, k):
        out = self.zeros()
        lib.or_inverse_monomial(C.byref(self.cp), k, _p64(out))
        return out

    def rgsw_encrypt(self, sk, msg, r, sigma=6.4):
        out = self.zeros(2 * self.params.gadget_len, 2)
        lib.or_rgsw_encrypt(C.byref(self.cp), self.ntt, _p64(np.ascontiguousarray(sk)), _p64(np.ascontiguousarray(msg)), C.byref(r), sigma, _p64(out))
        return out

    def seeded_rgsw_encrypt(self, sk, msg, r, sigma=6.4):
        ell2 = 2 * self.params.gadget_len
        seeds = np.zeros((ell2, 32), dtype=np.uint8)
        b = self.zeros(ell2)
        lib.or_seeded_rgsw_encrypt(C.byref(self.cp), self.ntt, _p64(np.ascontiguousarray(sk)), _p64(np.ascontiguousarray(msg)), C.byref(r), sigma, seeds.ctypes.data_as(_u8p), _p64(b))
        return seeds, b

    def client_packing_keys(self, gamma, w_seed, sk, r, sigma=6.4, with_y_all=True):
        pk = self.pack(gamma, w_seed)
        ell = self.params.gadget_len
        y_body = self.zeros(ell)
        y_all = self.zeros(max(gamma - 1, 1), ell) if with_y_all else None
        lib.or_client_packing_keys(C.byref(self.cp), self.ntt, pk, _p64(np.ascontiguousarray(sk)), C.byref(r), sigma, _p64(y_body), _p64(y_all) if with_y_all else None)
        return y_body, (y_all[: gamma - 1] if with_y_all else None)

    def decrypt(self, ct, sk):
        out = np.zeros(self.params.d, dtype=np.uint64)
        lib.or_decrypt(C.byref(self.cp), self.ntt, _p64(np.ascontiguousarray(ct, dtype=np.uint64)), _p64(np.ascontiguousarray(sk)), _p64(out))
        return out

    # -- extract (src/pir/extract.rs): the client's side of a response
    def extract(self, ciphertext, column_ciphertexts, sk, entry_size: int) -> bytes:
        """extract (extract.rs:31-68): coefficient 0 of the first
        ceil(entry_size / 2) column ciphertexts; without column ciphertexts,
        coefficient 0 of the ciphertext stands for every column."""
        ncols = (entry_size * 8 + 15) // 16
        cols = [] if column_ciphertexts is None else list(column_ciphertexts)
        if cols:
            vals = [int(self.decrypt(c, sk)[0]) for c in cols[:ncols]]
        else:
            vals = [int(self.decrypt(ciphertext, sk)[0])] * ncols
        return reconstruct_entry(vals, entry_size)

    def extract_packed(self, ciphertext, sk, entry_size: int) -> bytes:
        """extract_packed (extract.rs:119-157), the OnePacking / TwoPacking arms
        of extract_with_variant (:78-92): coefficient c times d^-1 mod p, with
        1 in place of d^-1 when gcd(d, p) != 1 (mod_inverse(..).unwrap_or(1))."""
        d, p = self.params.d, self.params.p
        ncols = (entry_size * 8 + 15) // 16
        m = self.decrypt(ciphertext, sk)
        dinv = pow(d, -1, p) if math.gcd(d, p) == 1 else 1
        return reconstruct_entry([int(m[c]) * dinv % p for c in range(ncols)], entry_size)

    def extract_inspiring(self, ciphertext, sk, entry_size: int) -> bytes:
        """extract_inspiring (extract.rs:163-192): coefficient c, no scaling."""
        ncols = (entry_size * 8 + 15) // 16
        m = self.decrypt(ciphertext, sk)
        return reconstruct_entry([int(m[c]) for c in range(ncols)], entry_size)

    def encode_shard(self, db: np.ndarray, entry_size: int, shard_idx: int, entries_per_shard=None):
        db = np.ascontiguousarray(db, dtype=np.uint8)
        num_entries = db.size // entry_size
        eps = entries_per_shard or self.params.d
        ncols = (entry_size * 8 + 15) // 16
        out = self.zeros(ncols)
        lib.or_encode_shard(C.byref(self.cp), db.ctypes.data_as(_u8p), num_entries, entry_size, eps, shard_idx, _p64(out))
        return out


def chacha_block(key_words, counter, stream=0):
    k = (C.c_uint32 * 8)(*key_words)
    out = (C.c_uint32 * 16)()
    lib.or_chacha20_block(k, counter, stream, out)
    return list(out)


def reconstruct_entry(column_values, entry_size):
    """encode_db.rs:186-201."""
    e = bytearray(entry_size)
    for c, v in enumerate(column_values):
        o = 2 * c
        if o < entry_size:
            e[o] = int(v) & 0xFF
        if o + 1 < entry_size:
            e[o + 1] = (int(v) >> 8) & 0xFF
    return bytes(e)
