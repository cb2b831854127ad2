"""Canonical bytes of AIR-table proofs: the fields of the reference's
`StarkProof` (crypto/stark/src/proof/stark.rs:83-134) in its declaration
order, and `MultiProof` (:143-145) as a list of them.

Encoding: integers and lengths as u64 little-endian; a field element as its
canonical u64 little-endian (an Fp3 element as its three components);
commitments as 32 raw bytes; `Option` as a one-byte tag (0 none, 1 some)
followed by the value; `Vec` as its length then its items; `Table<E>`
(crypto/stark/src/table.rs) as width, height, then the entries row by row.
The reference archives proofs with rkyv 0.8.16; its byte layout is not
reproduced here (no archived proof ships with the reference to check one
against), so these bytes are this library's canonical form, stable across
prover paths: the GPU prover and the oracle's replay serialize to the same
bytes.

Field order of one proof:
  trace_length; lde_trace_main_merkle_root; lde_trace_aux_merkle_root
  (Option); lde_trace_precomputed_merkle_root (Option: preprocessed tables);
  trace_ood_evaluations (Table: 1 row x W); trace_ood_next_evaluations
  (Table: K-1 rows x C, the C columns the AIR opens at the next rows,
  crypto/stark/src/ood.rs:94-126);
  composition_poly_root; composition_poly_parts_ood_evaluation (Vec<E>);
  fri_layers_merkle_roots (Vec); fri_final_poly_coeffs (Vec<E>); query_list
  (Vec of FriDecommitment: layers_auth_paths (Vec of Vec<Commitment>),
  layers_evaluations_sym (Vec<E>), fri/fri_decommit.rs:17-20);
  deep_poly_openings (Vec of DeepPolynomialOpening: composition_poly,
  main_trace_polys (the multiplicity columns of a preprocessed table),
  precomputed_trace_polys (Option: its fixed columns), aux_trace_polys
  (Option); each PolynomialOpenings = proof (Vec<Commitment>), evaluations,
  evaluations_sym, proof/stark.rs:46-69); nonce (Option<u64>);
  bus_public_inputs (Option<E>: the table contribution L); the query indices
  (not a reference field: the verifier re-derives them; kept so a proof
  read back carries them).
"""
from __future__ import annotations

import hashlib
import struct

import numpy as np

from .prover import AirTableProof

P = 0xFFFFFFFF00000001


class _W:
    def __init__(self):
        self.b = bytearray()

    def u64(self, v):
        self.b += struct.pack("<Q", int(v))

    def u8(self, v):
        self.b += bytes([v])

    def raw(self, data):
        self.b += bytes(data)

    def fe(self, v):
        self.u64(int(v) % P)

    def fe3(self, v):
        for x in np.asarray(v, dtype=np.uint64).reshape(3):
            self.fe(x)

    def vec_fe3(self, vals):
        vals = np.asarray(vals, dtype=np.uint64).reshape(-1, 3)
        self.u64(vals.shape[0])
        for v in vals:
            self.fe3(v)

    def vec_fe(self, vals):
        vals = np.asarray(vals, dtype=np.uint64).reshape(-1)
        self.u64(vals.size)
        for v in vals:
            self.fe(v)

    def path(self, nodes):
        nodes = np.asarray(nodes, dtype=np.uint8).reshape(-1, 32)
        self.u64(nodes.shape[0])
        for n in nodes:
            self.raw(n.tobytes())


# Frame offsets are 8-bit in the constraint program encoding, so a proof has
# fewer than 256 next-row rows.
MAX_FRAME_ROWS = 255


class _R:
    def __init__(self, data: bytes):
        self.d, self.i = memoryview(data), 0

    def u64(self):
        if self.i + 8 > len(self.d):
            raise ValueError("truncated proof")
        v = struct.unpack_from("<Q", self.d, self.i)[0]
        self.i += 8
        return v

    def u8(self):
        if self.i >= len(self.d):
            raise ValueError("truncated proof")
        v = self.d[self.i]
        self.i += 1
        return v

    def remaining(self) -> int:
        return len(self.d) - self.i

    def count(self, item_bytes: int, what: str) -> int:
        """A length prefix, refused when its items (at least item_bytes each)
        cannot fit in the bytes left."""
        k = self.u64()
        if k * item_bytes > self.remaining():
            raise ValueError(f"{what}: count {k} exceeds the proof's remaining bytes")
        return k

    def raw(self, n):
        v = bytes(self.d[self.i:self.i + n])
        if len(v) != n:
            raise ValueError("truncated proof")
        self.i += n
        return v

    def fe(self):
        v = self.u64()
        if v >= P:
            raise ValueError("non-canonical field element")
        return v

    def fe3(self):
        return [self.fe(), self.fe(), self.fe()]

    def vec_fe3(self):
        return np.array([self.fe3() for _ in range(self.count(24, "Fp3 vector"))], dtype=np.uint64).reshape(-1, 3)

    def vec_fe(self):
        return np.array([self.fe() for _ in range(self.count(8, "field vector"))], dtype=np.uint64)

    def path(self):
        k = self.count(32, "authentication path")
        return np.frombuffer(self.raw(32 * k), dtype=np.uint8).reshape(k, 32).copy()


def _write_opening(w: _W, path, rows_q, ext: bool):
    """PolynomialOpenings: proof, evaluations (row 2*iota), evaluations_sym (row 2*iota+1)."""
    w.path(path)
    for r in (rows_q[0], rows_q[1]):
        if ext:
            w.vec_fe3(r)
        else:
            w.vec_fe(r)


def _read_opening(r: _R, ext: bool):
    path = r.path()
    a = r.vec_fe3().reshape(-1) if ext else r.vec_fe()
    b = r.vec_fe3().reshape(-1) if ext else r.vec_fe()
    return path, np.stack([a, b])


def serialize_table_proof(p: AirTableProof) -> bytes:
    w = _W()
    w.u64(p.trace_rows)
    w.raw(p.main_root)
    w.u8(p.aux_root is not None)
    if p.aux_root is not None:
        w.raw(p.aux_root)
    w.u8(p.precomputed_root is not None)
    if p.precomputed_root is not None:
        w.raw(p.precomputed_root)
    b0 = np.asarray(p.ood_block0, dtype=np.uint64).reshape(-1, 3)
    W = b0.shape[0]
    w.u64(W)
    w.u64(1)
    for v in b0:
        w.fe3(v)
    b1 = np.asarray(p.ood_block1, dtype=np.uint64)
    C1 = b1.shape[0] if b1.ndim == 3 else 0  # (C, K-1, 3): next-row columns
    h1 = b1.shape[1] if C1 else 0  # a width-0 Table has height 0 (crypto/stark/src/table.rs:291-301)
    w.u64(C1)
    w.u64(h1)
    for r in range(h1):
        for j in range(C1):
            w.fe3(b1[j, r])
    w.raw(p.composition_root)
    w.vec_fe3(p.parts_ood)
    roots = [bytes(x) for x in p.fri_roots]
    w.u64(len(roots))
    for x in roots:
        w.raw(x)
    w.vec_fe3(p.fri_coeffs)
    Q = len(p.iotas)
    sym = np.asarray(p.fri_sym, dtype=np.uint64).reshape(len(roots), Q, 3) if roots else np.zeros((0, Q, 3),
                                                                                                   np.uint64)
    w.u64(Q)
    for q in range(Q):
        w.u64(len(roots))
        for li in range(len(roots)):
            w.path(p.fri_paths[li][q])
        w.vec_fe3(sym[:, q])
    w.u64(Q)
    for q in range(Q):
        _write_opening(w, p.parts_paths[q], p.parts_rows[q], ext=True)
        _write_opening(w, p.main_paths[q], p.main_rows[q], ext=False)
        w.u8(p.precomputed_rows is not None)
        if p.precomputed_rows is not None:
            _write_opening(w, p.precomputed_paths[q], p.precomputed_rows[q], ext=False)
        w.u8(p.aux_rows is not None)
        if p.aux_rows is not None:
            _write_opening(w, p.aux_paths[q], p.aux_rows[q], ext=True)
    w.u8(p.nonce is not None)
    if p.nonce is not None:
        w.u64(p.nonce)
    w.u8(p.table_contribution is not None)
    if p.table_contribution is not None:
        w.fe3(p.table_contribution)
    w.vec_fe(p.iotas)
    return bytes(w.b)


def deserialize_table_proof(data: bytes, reader: _R | None = None) -> AirTableProof:
    r = reader or _R(data)
    n = r.u64()
    main_root = r.raw(32)
    aux_root = r.raw(32) if r.u8() else None
    pre_root = r.raw(32) if r.u8() else None
    W = r.count(24, "current-row block width")
    if r.u64() != 1:
        raise ValueError("current-row block must have one row")
    b0 = np.array([r.fe3() for _ in range(W)], dtype=np.uint64).reshape(W, 3)
    # next-row block: only the columns the AIR opens at the next rows (the
    # verifier checks the width against the AIR)
    W1 = r.count(24, "next-row block width")
    h1 = r.u64()
    if h1 > MAX_FRAME_ROWS or W1 * h1 * 24 > r.remaining():
        raise ValueError("next-row block height")
    rows = [[r.fe3() for _ in range(W1)] for _ in range(h1)]
    b1 = np.array([[rows[k][j] for k in range(h1)] for j in range(W1)], dtype=np.uint64).reshape(W1, h1, 3)
    comp_root = r.raw(32)
    parts_ood = r.vec_fe3()
    nroots = r.count(32, "FRI layer roots")
    fri_roots = np.frombuffer(r.raw(32 * nroots), dtype=np.uint8).reshape(nroots, 32).copy()
    fri_coeffs = r.vec_fe3()
    Q = r.count(16, "FRI decommitments")  # each carries at least its layer count and a vector length
    fri_paths = [[] for _ in range(nroots)]
    sym_q = []
    for q in range(Q):
        if r.u64() != nroots:
            raise ValueError("FRI decommitment layer count")
        for li in range(nroots):
            fri_paths[li].append(r.path())
        s = r.vec_fe3()
        if s.shape[0] != nroots:
            raise ValueError("FRI symmetric values")
        sym_q.append(s)
    fri_sym = np.stack(sym_q, axis=1) if Q else np.zeros((nroots, 0, 3), dtype=np.uint64)
    if r.u64() != Q:
        raise ValueError("opening count")
    parts_paths, parts_rows, main_paths, main_rows, aux_paths, aux_rows = [], [], [], [], [], []
    pre_paths, pre_rows = [], []
    has_aux = None
    for q in range(Q):
        pp, pr = _read_opening(r, ext=True)
        mp, mr = _read_opening(r, ext=False)
        if bool(r.u8()) != (pre_root is not None):
            raise ValueError("precomputed openings without a precomputed commitment, or the reverse")
        if pre_root is not None:
            pp_, pr_ = _read_opening(r, ext=False)
            pre_paths.append(pp_)
            pre_rows.append(pr_)
        a = bool(r.u8())
        has_aux = a if has_aux is None else has_aux
        if a != has_aux:
            raise ValueError("aux openings present for some queries only")
        parts_paths.append(pp)
        parts_rows.append(pr)
        main_paths.append(mp)
        main_rows.append(mr)
        if a:
            ap, ar = _read_opening(r, ext=True)
            aux_paths.append(ap)
            aux_rows.append(ar)
    nonce = r.u64() if r.u8() else None
    L = tuple(int(x) for x in r.fe3()) if r.u8() else None
    iotas = r.vec_fe()
    if len(iotas) != Q:
        raise ValueError("query index count")
    return AirTableProof(
        trace_rows=n, main_root=main_root, aux_root=aux_root, table_contribution=L, composition_root=comp_root,
        ood_block0=b0, ood_block1=b1, parts_ood=parts_ood, fri_roots=fri_roots, fri_coeffs=fri_coeffs, nonce=nonce,
        iotas=iotas, fri_sym=fri_sym, fri_paths=[np.stack(x) if Q else x for x in fri_paths],
        main_rows=np.stack(main_rows) if Q else np.zeros((0, 2, 0), np.uint64),
        main_paths=np.stack(main_paths) if Q else main_paths,
        aux_rows=np.stack(aux_rows) if has_aux else None, aux_paths=np.stack(aux_paths) if has_aux else None,
        parts_rows=np.stack(parts_rows) if Q else np.zeros((0, 2, 0), np.uint64),
        parts_paths=np.stack(parts_paths) if Q else parts_paths, precomputed_root=pre_root,
        precomputed_rows=(np.stack(pre_rows) if Q else np.zeros((0, 2, 0), np.uint64)) if pre_root else None,
        precomputed_paths=(np.stack(pre_paths) if Q else pre_paths) if pre_root else None)


def serialize_multi_proof(proofs) -> bytes:
    """MultiProof: the number of tables, then each proof (length-prefixed)."""
    w = _W()
    w.u64(len(proofs))
    for p in proofs:
        b = serialize_table_proof(p)
        w.u64(len(b))
        w.raw(b)
    return bytes(w.b)


def deserialize_multi_proof(data: bytes):
    r = _R(data)
    out = []
    for _ in range(r.u64()):
        k = r.u64()
        out.append(deserialize_table_proof(r.raw(k)))
    if r.i != len(data):
        raise ValueError("trailing bytes after the proofs")
    return out


def proof_digest(data: bytes) -> str:
    return hashlib.sha256(data).hexdigest()
