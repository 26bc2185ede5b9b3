"""Seeded test inputs built with the oracle's client-side helpers.

Everything is derived from explicit seeds, so a failing case can be rebuilt
from the values printed by the test.
"""
from __future__ import annotations

import numpy as np

import oracle as O


def rng_seed(tag: int) -> bytes:
    return bytes((tag * 31 + i * 7 + 1) & 0xFF for i in range(32))


def random_db(num_entries: int, entry_size: int, seed: int, values_below_256: bool = False) -> bytes:
    g = np.random.default_rng(seed)
    db = g.integers(0, 256, num_entries * entry_size, dtype=np.uint8)
    if values_below_256:  # the reference's packed tests keep column values < 256 (tests/e2e_pir.rs:336-402)
        db = db.reshape(num_entries, entry_size)
        db[:, 1::2] = 0
        db = db.reshape(-1)
    return db.tobytes()


def encode_db(o: O.Oracle, db: bytes, entry_size: int) -> np.ndarray:
    """encode_database -> uint64[num_shards, cols, K, d] (shards of d entries)."""
    d = o.params.d
    arr = np.frombuffer(db, dtype=np.uint8)
    n = len(db) // entry_size
    nshards = (n + d - 1) // d
    return np.stack([o.encode_shard(arr, entry_size, s) for s in range(nshards)])


class Client:
    """A client key and RNG stream (client side is the oracle's, not the product's)."""

    def __init__(self, o: O.Oracle, seed_tag: int):
        self.o = o
        self.r = o.rng(rng_seed(seed_tag))
        self.sk = o.gaussian_poly(self.r)

    def query_rows(self, local_index: int) -> np.ndarray:
        """RgswCiphertext of X^(-k) (src/pir/query.rs:216-218)."""
        return self.o.rgsw_encrypt(self.sk, self.o.inverse_monomial(local_index), self.r)

    def seeded_query(self, local_index: int):
        """SeededRgswCiphertext of X^(-k): (seeds[2ell,32], b[2ell,K,d])."""
        return self.o.seeded_rgsw_encrypt(self.sk, self.o.inverse_monomial(local_index), self.r)

    def packing_keys(self, gamma: int, w_seed: bytes, with_y_all: bool = True):
        return self.o.client_packing_keys(gamma, w_seed, self.sk, self.r, with_y_all=with_y_all)


def uniform_polys(o: O.Oracle, count: int, seed: int, lead=()) -> np.ndarray:
    g = np.random.default_rng(seed)
    K, d = o.params.K, o.params.d
    out = np.empty((count,) + tuple(lead) + (K, d), dtype=np.uint64)
    for k, m in enumerate(o.params.moduli):
        out[..., k, :] = g.integers(0, m, out[..., k, :].shape, dtype=np.uint64)
    return out


def decrypt_columns(o: O.Oracle, sk, response_cols, ncols):
    return [int(o.decrypt(response_cols[c], sk)[0]) for c in range(ncols)]


def save_shards_binary(directory, ids, shard_polys, moduli, d: int):
    """Writes shard files in the layout of save_shards_binary
    (src/pir/mmap.rs:19-52): per shard `shard-{id:04}.bin` holding LE u32 id,
    u32 polynomial count, then (count > 0) u32 ring_dim, u32 crt_count, u64
    moduli, then every polynomial's residue-major u64 coefficients."""
    import os

    os.makedirs(directory, exist_ok=True)
    for sid, polys in zip(ids, shard_polys):
        polys = np.ascontiguousarray(polys, dtype=np.uint64)
        count = polys.shape[0] if polys.size else 0
        with open(os.path.join(directory, f"shard-{sid:04d}.bin"), "wb") as f:
            f.write(np.array([sid, count], dtype="<u4").tobytes())
            if count:
                f.write(np.array([d, len(moduli)], dtype="<u4").tobytes())
                f.write(np.array(moduli, dtype="<u8").tobytes())
                f.write(polys.astype("<u8").tobytes())
            else:
                f.write(np.array([0, 0], dtype="<u4").tobytes())


def rlwe_encrypt(o: O.Oracle, sk, msg_poly, r) -> np.ndarray:
    """RlweCiphertext::encrypt (rlwe/enc.rs:32-51): (a, -a*s + e + Delta*m),
    a uniform and e Gaussian from the ChaCha stream r."""
    a = o.zeros()
    for k, m in enumerate(o.params.moduli):
        a[k] = [O.lib.or_gen_range(r, m) for _ in range(o.params.d)]
    e = o.gaussian_poly(r)
    a_s = o.mul_ntt(a, sk)
    b = np.zeros_like(a)
    for k, m in enumerate(o.params.moduli):
        m = int(m)
        dm = o.params.delta % m
        b[k] = [((m - int(x)) % m + int(y) + dm * int(z)) % m for x, y, z in zip(a_s[k], e[k], msg_poly[k])]
    return np.stack([a, b])


def coeff0_message(o: O.Oracle, value: int) -> np.ndarray:
    mp = o.zeros()
    mp[:, 0] = value
    return mp


def phase_error(o: O.Oracle, sk, a, b, want) -> float:
    """max over coefficients of |a*s + b - Delta*want| (centred mod q) / Delta."""
    q, delta = o.params.q, o.params.delta
    ph = o.mul_ntt(a, sk)
    for k, m in enumerate(o.params.moduli):
        ph[k] = (ph[k] + b[k]) % m
    comp = [int(x) for x in o.coeffs(ph)]
    errs = [(c - delta * int(w)) % q for c, w in zip(comp, want)]
    return max(min(e, q - e) for e in errs) / delta
