"""Python mirror of the reference respond API over the gfx950 C ABI.

Binds libinspire_mi355x.so (include/inspire_mi355x.h) with ctypes.  Names,
argument meaning and errors follow the Rust crate `inspire`:

    respond(ctx, db_id, query)                 src/pir/respond.rs:71-115
    respond_seeded(ctx, db_id, query)          src/pir/respond.rs:408-415
    respond_inspiring(ctx, db_id, query)       src/pir/respond.rs:284-386
    respond_seeded_inspiring(ctx, db_id, query) src/pir/respond.rs:391-398
    respond_with_variant(ctx, db_id, query, v) src/pir/respond.rs:129-183
    respond_one_packing(ctx, db_id, query)     src/pir/respond.rs:213-268
    respond_seeded_packed(ctx, db_id, query)   src/pir/respond.rs:428-435
    extract(ctx, response, sk, entry_size)     src/pir/extract.rs:31-68
    extract_with_variant(..., variant)         src/pir/extract.rs:78-92
    extract_inspiring(ctx, response, sk, n)    src/pir/extract.rs:163-192

Errors raise PirError carrying the reference's message text
(src/pir/error.rs).  There is no CPU fallback: importing this module without
the built library raises, so every call runs on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libinspire_mi355x.so")

Q0, Q1 = 268369921, 249561089  # DEFAULT_CRT_MODULI (src/params.rs:31)

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_INVALID_PARAMS = 2
ERR_UNSUPPORTED = 3
ERR_SHARD_NOT_FOUND = 4
ERR_PACKING_KEYS_MISSING = 5
ERR_PACKING_KEYS_EMPTY = 6
ERR_PACKING_KEYS_MISSING_TREE = 7
ERR_UNKNOWN_DATABASE = 8
ERR_OUT_OF_MEMORY = 9
ERR_DEVICE = 10
ERR_CAPACITY = 11
ERR_PENDING = 12
REQUEST_BORROW = 0x1

NO_PACKING, ONE_PACKING, TWO_PACKING = 0, 1, 2  # InspireVariant (src/params.rs:78-103)
EXTRACT_COLUMNS, EXTRACT_PACKED, EXTRACT_INSPIRING = 0, 1, 2  # extract / extract_packed / extract_inspiring
PACKING_INSPIRING, PACKING_TREE = 0, 1  # PackingMode (src/pir/query.rs:27-40)
FLAG_RESIDUE_WISE_AUTOMORPHISM = 0x1


class PirError(Exception):
    """PirError(String) (src/pir/error.rs:13-45) plus the status code."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code
        self.message = message


class _Params(C.Structure):
    _fields_ = [
        ("ring_dim", C.c_uint32),
        ("crt_count", C.c_uint32),
        ("crt_moduli", C.c_uint64 * 2),
        ("q", C.c_uint64),
        ("p", C.c_uint64),
        ("gadget_base", C.c_uint64),
        ("gadget_len", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class _Query(C.Structure):
    _fields_ = [
        ("db_id", C.c_uint32),
        ("shard_id", C.c_uint32),
        ("seeded", C.c_uint32),
        ("packing_mode", C.c_uint32),
        ("rgsw", C.POINTER(C.c_uint64)),
        ("rgsw_seeds", C.POINTER(C.c_uint8)),
        ("has_packing_keys", C.c_uint32),
        ("y_all_len", C.c_uint32),
        ("y_body", C.POINTER(C.c_uint64)),
        ("y_all", C.POINTER(C.c_uint64)),
    ]


class _DeviceBatch(C.Structure):
    _fields_ = [
        ("count", C.c_uint32),
        ("db_id", C.c_uint32),
        ("variant", C.c_uint32),
        ("seeded", C.c_uint32),
        ("shard_ids", C.POINTER(C.c_uint32)),
        ("rows", C.c_void_p),
        ("seeds", C.c_void_p),
        ("y_body", C.c_void_p),
        ("out", C.c_void_p),
        ("status", C.POINTER(C.c_int32)),
        ("packing_mode", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class _Response(C.Structure):
    _fields_ = [
        ("ciphertext", C.POINTER(C.c_uint64)),
        ("column_ciphertexts", C.POINTER(C.c_uint64)),
        ("column_capacity", C.c_uint32),
        ("num_column_ciphertexts", C.c_uint32),
        ("ciphertext_capacity", C.c_uint32),
        ("ciphertext_b_dim", C.c_uint32),
    ]


class _ExtractBatch(C.Structure):
    _fields_ = [
        ("count", C.c_uint32),
        ("mode", C.c_uint32),
        ("entry_size", C.c_uint32),
        ("num_column_ciphertexts", C.c_uint32),
        ("b_dim", C.c_uint32),
        ("reserved", C.c_uint32),
        ("stride", C.c_uint64),
        ("responses", C.c_void_p),
        ("secret_keys", C.c_void_p),
        ("entries", C.c_void_p),
    ]


class _ServerConfig(C.Structure):
    _fields_ = [
        ("lanes", C.c_uint32),
        ("max_batch", C.c_uint32),
        ("hbm_budget", C.c_uint64),
        ("max_wait_us", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class _Request(C.Structure):
    _fields_ = [
        ("db_id", C.c_uint32),
        ("shard_id", C.c_uint32),
        ("variant", C.c_uint32),
        ("seeded", C.c_uint32),
        ("client_id", C.c_uint32),
        ("flags", C.c_uint32),
        ("rows", C.c_void_p),
        ("seeds", C.c_void_p),
        ("arrival_ns", C.c_uint64),
    ]


class _ServerStats(C.Structure):
    _fields_ = [
        ("requests", C.c_uint64),
        ("batches", C.c_uint64),
        ("max_batch_seen", C.c_uint64),
        ("peak_inflight_bytes", C.c_uint64),
        ("peak_lanes_busy", C.c_uint64),
        ("launch_groups", C.c_uint64),
    ]


RequestStruct = _Request  # inspire_request, for callers that prepare many at once
_u64p = C.POINTER(C.c_uint64)

# (name, restype, argtypes) of every entry point in include/inspire_mi355x.h
EXPORTS = [
    ("inspire_ctx_create", C.c_int, [C.POINTER(_Params), C.POINTER(C.c_uint8), C.POINTER(C.c_void_p)]),
    ("inspire_ctx_destroy", None, [C.c_void_p]),
    ("inspire_ctx_create_lane", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("inspire_last_error", C.c_char_p, []),
    ("inspire_ctx_device_bytes", C.c_uint64, [C.c_void_p]),
    ("inspire_db_load", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32, _u64p]),
    ("inspire_db_load_entries", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint8), C.c_uint64, C.c_uint32]),
    ("inspire_db_load_shard_files", C.c_int, [C.c_void_p, C.c_uint32, C.c_char_p]),
    ("inspire_db_unload", C.c_int, [C.c_void_p, C.c_uint32]),
    ("inspire_respond", C.c_int, [C.c_void_p, C.POINTER(_Query), C.POINTER(_Response)]),
    ("inspire_respond_seeded", C.c_int, [C.c_void_p, C.POINTER(_Query), C.POINTER(_Response)]),
    ("inspire_respond_inspiring", C.c_int, [C.c_void_p, C.POINTER(_Query), C.POINTER(_Response)]),
    ("inspire_respond_seeded_inspiring", C.c_int, [C.c_void_p, C.POINTER(_Query), C.POINTER(_Response)]),
    ("inspire_respond_with_variant", C.c_int, [C.c_void_p, C.POINTER(_Query), C.c_uint32, C.POINTER(_Response)]),
    ("inspire_respond_one_packing", C.c_int, [C.c_void_p, C.POINTER(_Query), C.POINTER(_Response)]),
    ("inspire_respond_seeded_packed", C.c_int, [C.c_void_p, C.POINTER(_Query), C.POINTER(_Response)]),
    ("inspire_ctx_set_galois_keys", C.c_int, [C.c_void_p, _u64p, C.c_uint32]),
    (
        "inspire_respond_batch",
        C.c_int,
        [C.c_void_p, C.c_uint32, C.POINTER(_Query), C.POINTER(C.c_uint32), C.POINTER(_Response), C.POINTER(C.c_int32)],
    ),
    ("inspire_respond_batch_device", C.c_int, [C.c_void_p, C.POINTER(_DeviceBatch)]),
    ("inspire_respond_batch_device_noalloc", C.c_int, [C.c_void_p, C.POINTER(_DeviceBatch)]),
    ("inspire_ctx_set_graphs", C.c_int, [C.c_void_p, C.c_int]),
    ("inspire_ctx_autotune", C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_uint64, _u64p]),
    ("inspire_ctx_load_tuning", C.c_int, [C.c_void_p, C.c_char_p]),
    ("inspire_ctx_set_kernel_choice", C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    ("inspire_ctx_get_tuning", C.c_int, [C.c_void_p, C.c_char_p, C.c_uint64, _u64p]),
    ("inspire_ctx_measure_rates", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_uint32]),
    ("inspire_ctx_graph_stats", C.c_int, [C.c_void_p, _u64p, _u64p]),
    ("inspire_register_client_keys", C.c_int, [C.c_void_p, C.c_uint32, _u64p]),
    ("inspire_server_create", C.c_int, [C.c_void_p, C.POINTER(_ServerConfig), C.POINTER(C.c_void_p)]),
    ("inspire_server_wait_many", C.c_int,
     [C.c_void_p, C.c_uint32, _u64p, C.c_int32, C.POINTER(_Response), C.POINTER(C.c_int32), _u64p]),
    ("inspire_server_submit", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(_Request), _u64p]),
    (
        "inspire_server_wait",
        C.c_int,
        [C.c_void_p, C.c_uint64, C.c_int32, C.POINTER(_Response), C.POINTER(C.c_int32), _u64p],
    ),
    ("inspire_server_get_stats", C.c_int, [C.c_void_p, C.POINTER(_ServerStats)]),
    ("inspire_server_destroy", None, [C.c_void_p]),
    ("inspire_ctx_set_profiling", C.c_int, [C.c_void_p, C.c_int]),
    ("inspire_ctx_set_chain_window", C.c_int, [C.c_void_p, C.c_uint32]),
    ("inspire_ctx_chain_window", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    ("inspire_ctx_set_profiling_kernels", C.c_int, [C.c_void_p, C.c_char_p]),
    (
        "inspire_ctx_profile",
        C.c_int,
        [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.c_int],
    ),
    ("inspire_response_to_binary", C.c_int,
     [C.POINTER(_Params), C.POINTER(_Response), C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("inspire_response_to_json", C.c_int,
     [C.POINTER(_Params), C.POINTER(_Response), C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("inspire_response_from_binary", C.c_int, [C.POINTER(_Params), C.c_void_p, C.c_uint64, C.POINTER(_Response)]),
    ("inspire_query_from_json", C.c_int,
     [C.POINTER(_Params), C.c_char_p, C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(_Query)]),
    ("inspire_query_from_binary", C.c_int,
     [C.POINTER(_Params), C.c_char_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(_Query)]),
    ("inspire_wire_query_free", None, [C.c_void_p]),
    ("inspire_extract", C.c_int, [C.c_void_p, C.POINTER(_Response), _u64p, C.c_uint32, C.c_void_p]),
    ("inspire_extract_with_variant", C.c_int,
     [C.c_void_p, C.POINTER(_Response), _u64p, C.c_uint32, C.c_uint32, C.c_void_p]),
    ("inspire_extract_inspiring", C.c_int, [C.c_void_p, C.POINTER(_Response), _u64p, C.c_uint32, C.c_void_p]),
    ("inspire_extract_batch_device", C.c_int, [C.c_void_p, C.POINTER(_ExtractBatch)]),
    ("inspire_ntt_forward", C.c_int, [C.c_void_p, _u64p, C.c_uint64]),
    ("inspire_ntt_inverse", C.c_int, [C.c_void_p, _u64p, C.c_uint64]),
    ("inspire_external_product", C.c_int, [C.c_void_p, C.c_uint64, _u64p, _u64p, _u64p]),
]


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads the native library; raises when it has not been built."""
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `make -C inspire-rs_amd` (or __graft_entry__.build()); "
            "the respond path has no CPU fallback"
        )
    lib = C.CDLL(path)
    for name, res, args in EXPORTS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _raise(code: int):
    if code != OK:
        msg = lib().inspire_last_error()
        raise PirError(code, msg.decode() if msg else f"error {code}")


def _ptr64(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


@dataclass
class InspireParams:
    """InspireParams (src/params.rs:134-185): the fields respond reads."""

    ring_dim: int = 2048
    crt_moduli: Sequence[int] = (Q0, Q1)
    p: int = 65536
    gadget_base: int = 1 << 20
    gadget_len: int = 3
    flags: int = 0

    @property
    def q(self) -> int:
        q = 1
        for m in self.crt_moduli:
            q *= m
        return q

    @property
    def crt_count(self) -> int:
        return len(self.crt_moduli)

    def delta(self) -> int:  # params.rs:296-298
        return self.q // self.p

    def _c(self) -> _Params:
        m = list(self.crt_moduli) + [0] * (2 - len(self.crt_moduli))
        return _Params(self.ring_dim, len(self.crt_moduli), (C.c_uint64 * 2)(*m), self.q, self.p,
                       self.gadget_base, self.gadget_len, self.flags)


@dataclass
class ClientQuery:
    """ClientQuery / SeededClientQuery (src/pir/query.rs:119-150).

    rgsw: full rows uint64[2ell, 2, K, d], or (seeded) the b polys
    uint64[2ell, K, d] with seeds uint8[2ell, 32].
    """

    shard_id: int
    rgsw: np.ndarray
    seeds: Optional[np.ndarray] = None
    packing_mode: int = PACKING_INSPIRING
    has_packing_keys: bool = False
    y_body: Optional[np.ndarray] = None   # uint64[ell, K, d]
    y_all: Optional[np.ndarray] = None    # uint64[len, ell, K, d]

    @property
    def seeded(self) -> bool:
        return self.seeds is not None


@dataclass
class ServerResponse:
    """ServerResponse (src/pir/respond.rs:30-37).

    ciphertext is uint64[2, K, d], except for a tree-packed response whose b
    has 2d coefficients per residue (two moduli, at least two columns,
    src/inspiring/automorph_pack.rs:476-484): then it is the flat words of
    a.coeffs followed by b.coeffs.  a and b give the two parts either way."""

    ciphertext: np.ndarray
    column_ciphertexts: np.ndarray = field(default=None)  # uint64[n, 2, K, d]
    b_dim: int = 0

    @property
    def a(self) -> np.ndarray:
        K, d = self.ciphertext.shape[1:] if self.ciphertext.ndim == 3 else self._kd()
        return self.ciphertext.reshape(-1)[:K * d].reshape(K, d)

    @property
    def b(self) -> np.ndarray:
        K, d = self.ciphertext.shape[1:] if self.ciphertext.ndim == 3 else self._kd()
        return self.ciphertext.reshape(-1)[K * d:].reshape(K, -1)

    def _kd(self):
        words = self.ciphertext.size  # K d + K b_dim with b_dim = 2d
        d = self.b_dim // 2
        return words // (3 * d), d


class Context:
    """A device context: parameters, CRS seed, resident databases."""

    def __init__(self, params: InspireParams, w_seed: bytes):
        self.params = params
        self._h = C.c_void_p()
        seed = (C.c_uint8 * 32)(*bytes(w_seed))
        p = params._c()
        _raise(lib().inspire_ctx_create(C.byref(p), seed, C.byref(self._h)))
        self.K = params.crt_count
        self.d = params.ring_dim
        self.ell = params.gadget_len
        self._dbs = {}

    def lane(self) -> "Context":
        """inspire_ctx_create_lane: a handle with its own stream and
        workspaces over this context's resident state."""
        ln = Context.__new__(Context)
        ln.params, ln.K, ln.d, ln.ell, ln._dbs = self.params, self.K, self.d, self.ell, self._dbs
        ln._h = C.c_void_p()
        _raise(lib().inspire_ctx_create_lane(self._h, C.byref(ln._h)))
        ln._owner = self  # the owner outlives its lanes
        return ln

    def close(self):
        if self._h:
            lib().inspire_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def device_bytes(self) -> int:
        return lib().inspire_ctx_device_bytes(self._h)

    # -- databases (EncodedDatabase, src/pir/setup.rs:105-120)
    def load_database(self, db_id: int, shard_polys: np.ndarray, shard_ids: Optional[Sequence[int]] = None):
        """shard_polys: uint64[num_shards, polys_per_shard, K, d] (ShardData.polynomials)."""
        sp = np.ascontiguousarray(shard_polys, dtype=np.uint64)
        ns = sp.shape[0]
        ids = np.ascontiguousarray(np.arange(ns) if shard_ids is None else shard_ids, dtype=np.uint32)
        cols = sp.shape[1] if sp.ndim == 4 else 0
        _raise(lib().inspire_db_load(self._h, db_id, ns, ids.ctypes.data_as(C.POINTER(C.c_uint32)), cols,
                                     _ptr64(sp) if sp.size else None))
        self._dbs[db_id] = cols

    def load_entries(self, db_id: int, data, entry_size: int):
        """encode_database (src/pir/encode_db.rs:106-164) of raw entries, then load.
        data: bytes or a uint8 numpy array (used in place, not copied)."""
        if isinstance(data, np.ndarray):
            buf = np.ascontiguousarray(data.reshape(-1), dtype=np.uint8)
        else:
            buf = np.frombuffer(bytes(data), dtype=np.uint8)
        _raise(lib().inspire_db_load_entries(self._h, db_id, buf.ctypes.data_as(C.POINTER(C.c_uint8)), buf.size,
                                             entry_size))
        self._dbs[db_id] = (entry_size * 8 + 15) // 16 if (entry_size and len(data)) else 0

    def load_shard_files(self, db_id: int, directory: str):
        """Every shard-*.bin file of a directory in the layout of
        save_shards_binary (src/pir/mmap.rs:19-52), ingested into HBM."""
        _raise(lib().inspire_db_load_shard_files(self._h, db_id, os.fsencode(directory)))
        cols = 0
        for name in os.listdir(directory):  # every file holds the same count (checked by the library)
            if name.startswith("shard-") and name.endswith(".bin"):
                with open(os.path.join(directory, name), "rb") as f:
                    cols = int(np.frombuffer(f.read(8), dtype="<u4")[1])
                break
        self._dbs[db_id] = cols

    def set_galois_keys(self, galois_keys: np.ndarray):
        """ServerCrs.galois_keys (src/pir/setup.rs:167-173): log2(d)
        key-switching matrices, uint64[log2 d, ell, 2, K, d]."""
        gk = np.ascontiguousarray(galois_keys, dtype=np.uint64)
        _raise(lib().inspire_ctx_set_galois_keys(self._h, _ptr64(gk), gk.shape[0] if gk.ndim else 0))

    def tree_b_dim(self, cols: int) -> int:
        """b coefficients per residue of a tree-packed response."""
        rw = self.params.flags & FLAG_RESIDUE_WISE_AUTOMORPHISM
        return 2 * self.d if (self.K == 2 and not rw and cols >= 2) else self.d

    def register_client_keys(self, client_id: int, y_body: np.ndarray):
        """Resident packing keys of a client: y_body of its ClientPackingKeys
        (src/inspiring/inspiring2.rs:603-628), uint64[ell, K, d]."""
        yb = np.ascontiguousarray(y_body, dtype=np.uint64)
        _raise(lib().inspire_register_client_keys(self._h, client_id, _ptr64(yb)))

    def unload_database(self, db_id: int):
        _raise(lib().inspire_db_unload(self._h, db_id))
        self._dbs.pop(db_id, None)

    # -- query marshalling
    def _query(self, db_id: int, q: ClientQuery, keep: list) -> _Query:
        rgsw = np.ascontiguousarray(q.rgsw, dtype=np.uint64)
        keep.append(rgsw)
        cq = _Query()
        cq.db_id = db_id
        cq.shard_id = q.shard_id
        cq.seeded = 1 if q.seeded else 0
        cq.packing_mode = q.packing_mode
        cq.rgsw = _ptr64(rgsw)
        if q.seeded:
            seeds = np.ascontiguousarray(q.seeds, dtype=np.uint8)
            keep.append(seeds)
            cq.rgsw_seeds = seeds.ctypes.data_as(C.POINTER(C.c_uint8))
        cq.has_packing_keys = 1 if q.has_packing_keys else 0
        if q.y_body is not None:
            yb = np.ascontiguousarray(q.y_body, dtype=np.uint64)
            keep.append(yb)
            cq.y_body = _ptr64(yb)
        if q.y_all is not None and len(q.y_all):
            ya = np.ascontiguousarray(q.y_all, dtype=np.uint64)
            keep.append(ya)
            cq.y_all = _ptr64(ya)
            cq.y_all_len = ya.shape[0]
        return cq

    def _response(self, cols: int, keep: list):
        ct = np.zeros(3 * self.K * self.d, dtype=np.uint64)  # room for a tree-packed b of 2d coefficients
        cc = np.zeros((max(cols, 1), 2, self.K, self.d), dtype=np.uint64)
        keep += [ct, cc]
        r = _Response()
        r.ciphertext = _ptr64(ct)
        r.column_ciphertexts = _ptr64(cc)
        r.column_capacity = cc.shape[0]
        r.ciphertext_capacity = ct.size
        return r, ct, cc

    def _result(self, r: _Response, ct: np.ndarray, cc: np.ndarray) -> ServerResponse:
        bd = r.ciphertext_b_dim or self.d
        words = ct[: self.K * self.d + self.K * bd]
        ciphertext = words.reshape(2, self.K, self.d) if bd == self.d else words
        return ServerResponse(ciphertext, cc[: r.num_column_ciphertexts], bd)

    def _call(self, fn, db_id: int, q: ClientQuery, *extra) -> ServerResponse:
        keep: list = []
        cq = self._query(db_id, q, keep)
        r, ct, cc = self._response(self._dbs.get(db_id, 0), keep)
        _raise(fn(self._h, C.byref(cq), *extra, C.byref(r)))
        return self._result(r, ct, cc)

    def respond_batch(self, db_ids: Sequence[int], queries: Sequence[ClientQuery], variants: Sequence[int]):
        """Many queries in one batched launch sequence -> list of ServerResponse or PirError."""
        n = len(queries)
        keep: list = []
        qa = (_Query * n)()
        ra = (_Response * n)()
        outs = []
        # Two allocations per batch, left uninitialised: the library writes every
        # word a response returns.  Ciphertexts have room for a tree-packed b of
        # 2d coefficients; InsPIRe^0 needs its columns, the packed variants one
        # column block at most (an empty shard's zero column).
        K, d = self.K, self.d
        caps = [max(self._dbs.get(db, 0), 1) if v == NO_PACKING else 1 for db, v in zip(db_ids, variants)]
        ct_all = np.empty((n, 3 * K * d), dtype=np.uint64)
        cc_all = np.empty((sum(caps), 2, K, d), dtype=np.uint64)
        ct_base, cc_base = ct_all.ctypes.data, cc_all.ctypes.data
        u64p = _u64p
        c0 = 0
        for i, (db, q) in enumerate(zip(db_ids, queries)):
            qa[i] = self._query(db, q, keep)
            r = ra[i]
            r.ciphertext = C.cast(ct_base + i * ct_all.strides[0], u64p)
            r.column_ciphertexts = C.cast(cc_base + c0 * cc_all.strides[0], u64p)
            r.column_capacity = caps[i]
            r.ciphertext_capacity = ct_all.shape[1]
            outs.append((ct_all[i], cc_all[c0:c0 + caps[i]]))
            c0 += caps[i]
        va = (C.c_uint32 * n)(*variants)
        st = (C.c_int32 * n)()
        _raise(lib().inspire_respond_batch(self._h, n, qa, va, ra, st))
        res = []
        for i in range(n):
            if st[i] != OK:
                res.append(PirError(st[i], f"query {i} failed with status {st[i]}"))
            else:
                ct, cc = outs[i]
                res.append(self._result(ra[i], ct, cc))
        return res

    # -- device-resident batches and profiling
    def respond_batch_device(self, db_id: int, variant: int, shard_ids, rows_ptr: int, out_ptr: int,
                             seeds_ptr: int = 0, y_body_ptr: int = 0, seeded: bool = True,
                             packing_mode: int = PACKING_INSPIRING, noalloc: bool = False):
        """Runs one batch whose inputs live in device memory (pointers from e.g.
        torch tensors; layouts in include/inspire_mi355x.h).  Returns the
        per-query status array.  noalloc: through
        inspire_respond_batch_device_noalloc (library-held scratch, no
        allocation in the call)."""
        ids = np.ascontiguousarray(shard_ids, dtype=np.uint32)
        st = np.zeros(ids.size, dtype=np.int32)
        b = _DeviceBatch(ids.size, db_id, variant, 1 if seeded else 0, ids.ctypes.data_as(C.POINTER(C.c_uint32)),
                         rows_ptr, seeds_ptr or None, y_body_ptr or None, out_ptr,
                         st.ctypes.data_as(C.POINTER(C.c_int32)), packing_mode, 0)
        fn = lib().inspire_respond_batch_device_noalloc if noalloc else lib().inspire_respond_batch_device
        _raise(fn(self._h, C.byref(b)))
        return st

    def extract_batch_device(self, count: int, mode: int, entry_size: int, responses_ptr: int, secret_keys_ptr: int,
                             entries_ptr: int, num_column_ciphertexts: int = 0, b_dim: int = 0, stride: int = 0):
        """Client-side extract of `count` device-resident responses (layouts in
        include/inspire_mi355x.h, inspire_extract_batch) into entries_ptr."""
        b = _ExtractBatch(count, mode, entry_size, num_column_ciphertexts, b_dim, 0, stride, responses_ptr,
                          secret_keys_ptr, entries_ptr)
        _raise(lib().inspire_extract_batch_device(self._h, C.byref(b)))

    def autotune(self, path: Optional[str] = None) -> dict:
        """Times the kernel variants (inspire_ctx_autotune), keeps the faster
        of each and returns the report; with path, also writes it there."""
        import json

        n = C.c_uint64()
        buf = C.create_string_buffer(4096)
        _raise(lib().inspire_ctx_autotune(self._h, path.encode() if path else None, buf, 4096, C.byref(n)))
        return json.loads(buf.value.decode())

    def load_tuning(self, path: str):
        _raise(lib().inspire_ctx_load_tuning(self._h, path.encode()))

    def set_kernel_choice(self, name: str, value: int):
        _raise(lib().inspire_ctx_set_kernel_choice(self._h, name.encode(), int(value)))

    def tuning(self) -> dict:
        import json

        n = C.c_uint64()
        buf = C.create_string_buffer(1024)
        _raise(lib().inspire_ctx_get_tuning(self._h, buf, 1024, C.byref(n)))
        return json.loads(buf.value.decode())

    def measure_rates(self) -> dict:
        """Measured rates (inspire_ctx_measure_rates): T operations/s of the
        Montgomery product, its signed form and the 64-bit multiply-add, and
        the streaming read rate of HBM in TB/s."""
        t = (C.c_double * 4)()
        _raise(lib().inspire_ctx_measure_rates(self._h, t, 4))
        return {"mmul_tops": t[0], "smmul_tops": t[1], "mad64_tops": t[2], "hbm_read_tbs": t[3]}

    def set_graphs(self, on: bool):
        """Graph replay of repeated device batches (on by default)."""
        _raise(lib().inspire_ctx_set_graphs(self._h, 1 if on else 0))

    def graph_stats(self):
        """-> (captures, replays) of device-batch graphs on this handle."""
        a, b = C.c_uint64(), C.c_uint64()
        _raise(lib().inspire_ctx_graph_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_chain_window(self, window_us: int):
        """A chain launch submitted within window_us of another lane's waits for it (0: off)."""
        _raise(lib().inspire_ctx_set_chain_window(self._h, int(window_us)))

    def chain_window(self) -> int:
        """The owner's chain window in microseconds."""
        w = C.c_uint32()
        _raise(lib().inspire_ctx_chain_window(self._h, C.byref(w)))
        return int(w.value)

    def set_profiling(self, on: bool):
        _raise(lib().inspire_ctx_set_profiling(self._h, 1 if on else 0))

    def set_profiling_kernels(self, names=None):
        """Bracket only these kernels (an iterable of names); None: all."""
        _raise(lib().inspire_ctx_set_profiling_kernels(self._h, ",".join(names).encode() if names else None))

    def profile(self, kernel: str, reset: bool = False):
        """(total milliseconds, launches) of one kernel since the last reset."""
        ms = C.c_double()
        n = C.c_uint64()
        _raise(lib().inspire_ctx_profile(self._h, kernel.encode(), C.byref(ms), C.byref(n), 1 if reset else 0))
        return ms.value, n.value

    # -- math entry points
    def ntt_forward(self, polys: np.ndarray) -> np.ndarray:
        """NttContext::forward (src/math/ntt.rs:223-246), reference representation."""
        a = np.ascontiguousarray(polys, dtype=np.uint64).copy()
        _raise(lib().inspire_ntt_forward(self._h, _ptr64(a), a.size // (self.K * self.d)))
        return a

    def ntt_inverse(self, polys: np.ndarray) -> np.ndarray:
        """NttContext::inverse (src/math/ntt.rs:308-325)."""
        a = np.ascontiguousarray(polys, dtype=np.uint64).copy()
        _raise(lib().inspire_ntt_inverse(self._h, _ptr64(a), a.size // (self.K * self.d)))
        return a

    def external_product(self, rlwe: np.ndarray, rgsw: np.ndarray) -> np.ndarray:
        """external_product (src/rgsw/external_product.rs:88-150), batched over the leading axis."""
        rl = np.ascontiguousarray(rlwe, dtype=np.uint64)
        rg = np.ascontiguousarray(rgsw, dtype=np.uint64)
        count = rl.size // (2 * self.K * self.d)
        out = np.zeros((count, 2, self.K, self.d), dtype=np.uint64)
        _raise(lib().inspire_external_product(self._h, count, _ptr64(rl), _ptr64(rg), _ptr64(out)))
        return out


@dataclass
class Request:
    """A query to the serving loop (inspire_request): rows ell..2ell-1 as
    uint32 (seeded: b polys [ell, K, d] with seeds uint8[ell, 32]; full:
    [ell, 2, K, d]).  Packed variants name a registered client."""

    db_id: int
    shard_id: int
    variant: int
    rows: np.ndarray
    seeds: Optional[np.ndarray] = None
    client_id: int = 0
    arrival_ns: int = 0


class Server:
    """Continuous-batching serving loop over a Context (inspire_server_*)."""

    def __init__(self, ctx: Context, lanes: int = 2, max_batch: int = 256, hbm_budget: int = 0,
                 max_wait_us: int = 0):
        self.ctx = ctx
        cfg = _ServerConfig(lanes, max_batch, hbm_budget, max_wait_us, 0)
        self._h = C.c_void_p()
        _raise(lib().inspire_server_create(ctx._h, C.byref(cfg), C.byref(self._h)))

    def submit(self, reqs: Sequence[Request]) -> np.ndarray:
        n = len(reqs)
        arr = (_Request * n)()
        keep = []
        for i, r in enumerate(reqs):
            rows = np.ascontiguousarray(r.rows, dtype=np.uint32)
            keep.append(rows)
            seeds = None
            if r.seeds is not None:
                seeds = np.ascontiguousarray(r.seeds, dtype=np.uint8)
                keep.append(seeds)
            arr[i] = _Request(r.db_id, r.shard_id, r.variant, 1 if seeds is not None else 0, r.client_id, 0,
                              rows.ctypes.data, seeds.ctypes.data if seeds is not None else None, r.arrival_ns)
        tickets = np.zeros(n, dtype=np.uint64)
        _raise(lib().inspire_server_submit(self._h, n, arr, _ptr64(tickets)))
        return tickets

    @staticmethod
    def request_struct(r: Request, borrow: bool = False):
        """-> (inspire_request struct, arrays it points to).  With borrow, the
        arrays must stay alive and unchanged until the ticket is waited."""
        rows = np.ascontiguousarray(r.rows, dtype=np.uint32)
        seeds = np.ascontiguousarray(r.seeds, dtype=np.uint8) if r.seeds is not None else None
        st = _Request(r.db_id, r.shard_id, r.variant, 1 if seeds is not None else 0, r.client_id,
                      REQUEST_BORROW if borrow else 0, rows.ctypes.data,
                      seeds.ctypes.data if seeds is not None else None, r.arrival_ns)
        return st, (rows, seeds)

    def submit_structs(self, arr, n: int) -> np.ndarray:
        """Submits n prepared inspire_request structs (a ctypes array)."""
        tickets = np.zeros(n, dtype=np.uint64)
        _raise(lib().inspire_server_submit(self._h, n, arr, _ptr64(tickets)))
        return tickets

    def wait(self, ticket: int, cols: int = 0, timeout_ms: int = -1):
        """-> (status, ServerResponse or None, done_ns).  cols: column
        capacity for InsPIRe^0 responses."""
        K, d = self.ctx.K, self.ctx.d
        ct = np.zeros((2, K, d), dtype=np.uint64)
        cc = np.zeros((max(cols, 1), 2, K, d), dtype=np.uint64)
        r = _Response(_ptr64(ct), _ptr64(cc) if cols else None, cols, 0, ct.size, 0)
        st = C.c_int32()
        done = C.c_uint64()
        rc = lib().inspire_server_wait(self._h, int(ticket), timeout_ms, C.byref(r), C.byref(st), C.byref(done))
        if rc == ERR_PENDING:
            return ERR_PENDING, None, 0
        _raise(rc)
        if st.value != OK:
            return st.value, None, done.value
        return OK, ServerResponse(ct, cc[: r.num_column_ciphertexts]), done.value

    def wait_many(self, tickets, cts: np.ndarray, cols_for=None, timeout_ms: int = -1):
        """Waits for the tickets in order (inspire_server_wait_many).  cts:
        uint64[n, 2, K, d] receives the ciphertexts; cols_for maps a position to
        a uint64[cap, 2, K, d] buffer for that query's column ciphertexts
        (InsPIRe^0).  -> (status int32[n], done_ns uint64[n], columns uint32[n])."""
        n = len(tickets)
        t = np.ascontiguousarray(tickets, dtype=np.uint64)
        ra = (_Response * n)()
        base, stride = cts.ctypes.data, cts.strides[0]
        words = cts[0].size if n else 0
        for i in range(n):
            r = ra[i]
            r.ciphertext = C.cast(base + i * stride, _u64p)
            r.ciphertext_capacity = words
            if cols_for and i in cols_for:
                cb = cols_for[i]
                r.column_ciphertexts = C.cast(cb.ctypes.data, _u64p)
                r.column_capacity = cb.shape[0]
        st = np.zeros(n, dtype=np.int32)
        dn = np.zeros(n, dtype=np.uint64)
        rc = lib().inspire_server_wait_many(self._h, n, _ptr64(t), timeout_ms, ra,
                                            st.ctypes.data_as(C.POINTER(C.c_int32)), _ptr64(dn))
        if rc not in (OK, ERR_PENDING):
            _raise(rc)
        return st, dn, np.array([ra[i].num_column_ciphertexts for i in range(n)], dtype=np.uint32)

    def stats(self) -> dict:
        s = _ServerStats()
        _raise(lib().inspire_server_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _ServerStats._fields_}

    def close(self):
        if self._h:
            lib().inspire_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceArray:
    """A device allocation made through the HIP runtime the library uses
    (for inspire_respond_batch_device inputs and outputs).  Copies are
    synchronous null-stream copies: while a lane captures a batch into a
    graph on another thread, this runtime refuses them
    (hipErrorStreamCaptureImplicit), so copy before or after concurrent
    batches, not during them."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            lib()  # the library's runtime is the one loaded
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            cls._hip = h
        return cls._hip

    @classmethod
    def _copy(cls, dst: int, src: int, nbytes: int, kind: int, what: str):
        rc = cls.hip().hipMemcpy(dst, src, nbytes, kind)
        if rc != 0:
            raise PirError(ERR_DEVICE, f"hipMemcpy {what} failed: {rc}")

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        rc = self.hip().hipMalloc(C.byref(p), max(int(nbytes), 1))
        if rc != 0:
            raise PirError(ERR_OUT_OF_MEMORY, f"hipMalloc({nbytes}) failed: {rc}")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self._copy(self.ptr, a.ctypes.data, a.nbytes, 1, "H2D")

    def download(self, a: np.ndarray) -> np.ndarray:
        assert a.flags["C_CONTIGUOUS"] and a.nbytes <= self.nbytes
        self._copy(a.ctypes.data, self.ptr, a.nbytes, 2, "D2H")
        return a

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def respond(ctx: Context, db_id: int, query: ClientQuery) -> ServerResponse:
    return ctx._call(lib().inspire_respond, db_id, query)


def respond_seeded(ctx: Context, db_id: int, query: ClientQuery) -> ServerResponse:
    return ctx._call(lib().inspire_respond_seeded, db_id, query)


def respond_inspiring(ctx: Context, db_id: int, query: ClientQuery) -> ServerResponse:
    return ctx._call(lib().inspire_respond_inspiring, db_id, query)


def respond_seeded_inspiring(ctx: Context, db_id: int, query: ClientQuery) -> ServerResponse:
    return ctx._call(lib().inspire_respond_seeded_inspiring, db_id, query)


def respond_with_variant(ctx: Context, db_id: int, query: ClientQuery, variant: int) -> ServerResponse:
    return ctx._call(lib().inspire_respond_with_variant, db_id, query, variant)


def respond_seeded_with_variant(ctx: Context, db_id: int, query: ClientQuery, variant: int) -> ServerResponse:
    assert query.seeded
    return ctx._call(lib().inspire_respond_with_variant, db_id, query, variant)


def respond_one_packing(ctx: Context, db_id: int, query: ClientQuery) -> ServerResponse:
    """respond_one_packing (src/pir/respond.rs:213-268): tree packing with the
    CRS galois keys (Context.set_galois_keys)."""
    return ctx._call(lib().inspire_respond_one_packing, db_id, query)


def respond_seeded_packed(ctx: Context, db_id: int, query: ClientQuery) -> ServerResponse:
    """respond_seeded_packed (src/pir/respond.rs:428-435)."""
    return ctx._call(lib().inspire_respond_seeded_packed, db_id, query)


# -- client-side extract (src/pir/extract.rs)

def _extract(fn, ctx: Context, response: ServerResponse, secret_key: np.ndarray, entry_size: int, *extra) -> bytes:
    keep: list = []
    r = _response_struct(ctx.params, response, keep)
    sk = np.ascontiguousarray(secret_key, dtype=np.uint64)
    assert sk.size == ctx.K * ctx.d
    out = np.zeros(max(entry_size, 1), dtype=np.uint8)
    _raise(fn(ctx._h, C.byref(r), _ptr64(sk), entry_size, *extra, out.ctypes.data))
    return out[:entry_size].tobytes()


def extract(ctx: Context, response: ServerResponse, secret_key: np.ndarray, entry_size: int) -> bytes:
    """extract (src/pir/extract.rs:31-68); secret_key = ClientState.rlwe_secret_key, uint64[K, d]."""
    return _extract(lib().inspire_extract, ctx, response, secret_key, entry_size)


def extract_with_variant(ctx: Context, response: ServerResponse, secret_key: np.ndarray, entry_size: int,
                         variant: int) -> bytes:
    """extract_with_variant (src/pir/extract.rs:78-92)."""
    return _extract(lib().inspire_extract_with_variant, ctx, response, secret_key, entry_size, variant)


def extract_inspiring(ctx: Context, response: ServerResponse, secret_key: np.ndarray, entry_size: int) -> bytes:
    """extract_inspiring (src/pir/extract.rs:163-192)."""
    return _extract(lib().inspire_extract_inspiring, ctx, response, secret_key, entry_size)


# -- wire formats (bincode / serde_json of the reference's types)

def _response_struct(params: InspireParams, resp: ServerResponse, keep: list) -> _Response:
    K, d = params.crt_count, params.ring_dim
    ct = np.ascontiguousarray(resp.ciphertext, dtype=np.uint64).reshape(-1)
    cols = resp.column_ciphertexts if resp.column_ciphertexts is not None else np.zeros((0, 2, K, d), np.uint64)
    cc = np.ascontiguousarray(cols, dtype=np.uint64)
    keep += [ct, cc]
    r = _Response()
    r.ciphertext = _ptr64(ct)
    r.column_ciphertexts = _ptr64(cc) if cc.size else None
    r.column_capacity = cc.shape[0]
    r.num_column_ciphertexts = cc.shape[0]
    r.ciphertext_capacity = ct.size
    r.ciphertext_b_dim = resp.b_dim or (ct.size // K - d)
    return r


def _encode(fn, params: InspireParams, resp: ServerResponse) -> bytes:
    keep: list = []
    r = _response_struct(params, resp, keep)
    p = params._c()
    n = C.c_uint64()
    rc = fn(C.byref(p), C.byref(r), None, 0, C.byref(n))
    if rc != ERR_CAPACITY:
        _raise(rc)
    buf = C.create_string_buffer(n.value)
    _raise(fn(C.byref(p), C.byref(r), buf, n.value, C.byref(n)))
    return buf.raw[: n.value]


def response_to_binary(params: InspireParams, resp: ServerResponse) -> bytes:
    """ServerResponse::to_binary (src/pir/respond.rs:43-46): bincode 1.3."""
    return _encode(lib().inspire_response_to_binary, params, resp)


def response_to_json(params: InspireParams, resp: ServerResponse) -> bytes:
    """serde_json::to_vec(&ServerResponse) (the body of src/bin/server.rs:88-92)."""
    return _encode(lib().inspire_response_to_json, params, resp)


def response_from_binary(params: InspireParams, data: bytes, max_columns: int = 64) -> ServerResponse:
    """ServerResponse::from_binary (src/pir/respond.rs:48-51)."""
    K, d = params.crt_count, params.ring_dim
    ct = np.zeros(3 * K * d, dtype=np.uint64)
    cc = np.zeros((max(max_columns, 1), 2, K, d), dtype=np.uint64)
    r = _Response(_ptr64(ct), _ptr64(cc), cc.shape[0], 0, ct.size, 0)
    p = params._c()
    buf = C.create_string_buffer(bytes(data), len(data))
    _raise(lib().inspire_response_from_binary(C.byref(p), buf, len(data), C.byref(r)))
    bd = r.ciphertext_b_dim
    words = ct[: K * d + K * bd]
    return ServerResponse(words.reshape(2, K, d) if bd == d else words, cc[: r.num_column_ciphertexts], bd)


def _decoded(params: InspireParams, handle: C.c_void_p, q: _Query) -> ClientQuery:
    K, d, ell = params.crt_count, params.ring_dim, params.gadget_len
    try:
        if q.seeded:
            rgsw = np.ctypeslib.as_array(q.rgsw, shape=(2 * ell * K * d,)).reshape(2 * ell, K, d).copy()
            seeds = np.ctypeslib.as_array(q.rgsw_seeds, shape=(2 * ell * 32,)).reshape(2 * ell, 32).copy()
        else:
            rgsw = np.ctypeslib.as_array(q.rgsw, shape=(4 * ell * K * d,)).reshape(2 * ell, 2, K, d).copy()
            seeds = None
        y_body = None
        if q.y_body:
            y_body = np.ctypeslib.as_array(q.y_body, shape=(ell * K * d,)).reshape(ell, K, d).copy()
        return ClientQuery(shard_id=q.shard_id, rgsw=rgsw, seeds=seeds, packing_mode=q.packing_mode,
                           has_packing_keys=bool(q.has_packing_keys), y_body=y_body)
    finally:
        lib().inspire_wire_query_free(handle)


def query_from_json(params: InspireParams, data: bytes) -> ClientQuery:
    """A ClientQuery / SeededClientQuery as the server's handlers receive it
    (serde_json, src/bin/server.rs:120-236)."""
    p = params._c()
    h = C.c_void_p()
    q = _Query()
    _raise(lib().inspire_query_from_json(C.byref(p), bytes(data), len(data), C.byref(h), C.byref(q)))
    return _decoded(params, h, q)


def query_from_binary(params: InspireParams, data: bytes, seeded: bool) -> ClientQuery:
    """bincode::deserialize-side of a ClientQuery (seeded=False) or SeededClientQuery."""
    p = params._c()
    h = C.c_void_p()
    q = _Query()
    _raise(lib().inspire_query_from_binary(C.byref(p), bytes(data), len(data), 1 if seeded else 0, C.byref(h),
                                           C.byref(q)))
    return _decoded(params, h, q)
