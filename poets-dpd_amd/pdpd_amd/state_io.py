"""WorldState files for the Python World of pdpd_amd.world, in the reference's
three formats (dpd_state_io.hpp:109-588): text "v0", and "v1binary" /
"v2binary" (text header, then one packed record per polymer). The C++ tools use
include/pdpd/world_state_io.hpp, whose header comment gives the record layout;
both write the same bytes, and those are the bytes the reference writes for the
same state (text numbers in C++'s default 6-significant-digit format, constant
parameters as std::to_string's %f, expressions with every binary operation in
parentheses). Files ending in .gz / .bz2 are compressed.
Binary records are parsed with one Python step per polymer."""
from __future__ import annotations

import bz2
import gzip
import struct

import numpy as np

from . import expr
from .world import PolymerType, World


def _g(v):
    """A double as C++'s default ostream writes it."""
    return format(float(v), "g")


def _f(v):
    """A constant parameter as std::to_string writes it."""
    return f"{float(v):f}"


def _open(name, mode):
    name = str(name)
    if name.endswith(".gz"):
        return gzip.open(name, mode + "b")
    if name.endswith(".bz2"):
        return bz2.open(name, mode + "b")
    return open(name, mode + "b")


def _cround(x):
    """C's round(): to nearest, ties away from zero (t - floor(t) is exact)."""
    t = np.abs(x)
    r = np.floor(t)
    r = r + (t - r >= 0.5)
    return np.copysign(r, x)


def pack_half3(vecs) -> np.ndarray:
    """(n, 3) doubles -> (n,) uint32 in the v2binary vector form: bits 0-4 hold
    e+15 for a common scale 2^e, e in [-14, 15] (0 = zero vector), bits 5-13 /
    14-22 / 23-31 the components as 9-bit two's complement m, component =
    m * 2^(e-8). e is the largest frexp exponent of the components, m =
    round(c * 2^(8-e)); a component rounding to +-256 raises e by one.
    Subnormal components count as zero; vectors below 2^-15 become zero."""
    c = np.array(vecs, dtype=np.float64).reshape(-1, 3)
    if not np.isfinite(c).all():
        raise ValueError("cannot pack NaN or infinity into a v2binary vector")
    c[np.abs(c) < np.finfo(np.float64).tiny] = 0.0
    _, ex = np.frexp(c)
    e = np.where(c != 0.0, ex, -(1 << 20)).max(axis=1)
    nz = (c != 0.0).any(axis=1)
    if (nz & (e > 15)).any():
        raise ValueError("vector too large for a v2binary record (magnitude 65536 or more)")
    keep = nz & (e >= -14)
    e = np.where(keep, e, 0)
    m = _cround(np.ldexp(c, (8 - e)[:, None]))
    carry = keep & (np.abs(m) == 256.0).any(axis=1)
    if (carry & (e + 1 > 15)).any():
        raise ValueError("vector too large for a v2binary record (magnitude 65536 or more)")
    e = e + carry
    m = np.where(carry[:, None], _cround(np.ldexp(c, (8 - e)[:, None])), m).astype(np.int64)
    bits = (e + 15).astype(np.uint64)
    for d in range(3):
        bits |= (m[:, d].astype(np.uint64) & np.uint64(0x1FF)) << np.uint64(5 + 9 * d)
    return np.where(keep, bits, 0).astype(np.uint32)


def unpack_half3(bits) -> np.ndarray:
    b = np.asarray(bits, dtype=np.uint32).reshape(-1).astype(np.int64)
    e = b & 31
    out = np.empty((b.size, 3))
    for d in range(3):
        m = (b >> (5 + 9 * d)) & 0x1FF
        m = np.where(m & 0x100, m - 0x200, m)
        out[:, d] = np.ldexp(m.astype(np.float64), e - 23)
    return out


def unpack_c3(words) -> np.ndarray:
    """(n, 4) int16 v1binary vectors (e, m0, m1, m2) -> (n, 3): m * 2^(e-15)."""
    w = np.asarray(words, dtype=np.int16).reshape(-1, 4).astype(np.int64)
    return np.ldexp(w[:, 1:].astype(np.float64), w[:, :1] - 15)


def pack_pos(x, box) -> np.ndarray:
    """(n, 3) positions relative to the origin -> uint32 in units of 2^-16; a
    coordinate rounding up to the box edge wraps to 0."""
    r = _cround(np.asarray(x, np.float64) * 65536.0)
    lim = np.asarray(box, np.float64) * 65536.0
    r = np.where(r == lim, 0.0, r)
    if not ((r >= 0) & (r < lim)).all():
        raise ValueError("Bead is out of bounds.")
    return r.astype(np.uint32)


def _binary_beads(w: World) -> bytes:
    out = bytearray()
    pos = pack_pos(w.x, w.box)
    vb, fb = pack_half3(w.v), pack_half3(w.f)
    prev_p, prev_b = 0xFFFFFFFF, 0xFFFFFFFF
    starts, ids = w.polymer_bead_start, w.polymer_bead_ids
    for pid in range(len(w.polymer_type_of)):
        bead_ids = [int(i) for i in ids[starts[pid]:starts[pid + 1]]]
        pt = int(w.polymer_type_of[pid])
        if pt >= 1 << 13:
            raise ValueError("polymer type does not fit the 13-bit binary field")
        jp = int(pid != (prev_p + 1) & 0xFFFFFFFF)
        jf = int(bead_ids[0] != (prev_b + 1) & 0xFFFFFFFF)
        ji = int(any(bead_ids[k] != bead_ids[k - 1] + 1 for k in range(1, len(bead_ids))))
        out += struct.pack("<H", pt | jp << 13 | jf << 14 | ji << 15)
        if jp:
            out += struct.pack("<I", (pid - prev_p) & 0xFFFFFFFF)
        prev_p = pid
        for k, b in enumerate(bead_ids):
            if jf if k == 0 else ji:
                out += struct.pack("<I", (b - prev_b) & 0xFFFFFFFF)
            prev_b = b
            out += struct.pack("<5I", pos[b, 0], pos[b, 1], pos[b, 2], vb[b], fb[b])
    return bytes(out)


def dumps(w: World, precise=False, bead_radius=0.5, binary=False):
    """The file contents: str for text v0, bytes for binary=True (v2binary, as
    the reference's write_world_state(..., true)). precise=True writes the text
    numbers with 17 significant digits, as the C++ writer's precise mode does."""
    g = (lambda v: format(float(v), ".17g")) if precise else _g
    origin = (0.0, 0.0, 0.0) if w.origin is None else tuple(float(o) for o in w.origin)
    nt = w.n_types
    fmt = "v2binary" if binary else "v0"
    out = [f"WorldState {fmt} {nt} {len(w.polymer_types)} {w.n_beads} {len(w.polymer_type_of)}",
           f"T {int(w.t)} {g(w.dt)}", f"Lambda {g(w.lambda_)}",
           "Origin " + " ".join(g(o) for o in origin), "Box " + " ".join(g(b) for b in w.box),
           f"Seed {int(w.seed)}"]
    ex = w.exprs or {}
    for i in range(nt):
        out.append(f"ConservativeStrength {i} " + " ".join(ex.get(("conservative", i, j)) or _f(w.conservative[i, j])
                                                           for j in range(nt)))
    for i in range(nt):
        out.append(f"DissipativeStrength {i} " + " ".join(ex.get(("dissipative", i, j)) or _f(w.dissipative[i, j])
                                                          for j in range(nt)))
    out += ["", "# BeadTypes"]
    stat = w.stationary if w.stationary is not None else np.zeros(nt, np.uint8)
    for i, name in enumerate(w.bead_type_names):
        out.append(f"BeadType {i} {name} {g(bead_radius)}" + (" stationary" if stat[i] else ""))
    out += ["", "# polymerTypes"]
    for i, p in enumerate(w.polymer_types):
        out.append(f"PolymerType {i} {p.name} {len(p.bead_types)} {len(p.bonds)} {len(p.bond_pairs)}")
        out.append("BeadTypeIndices " + " ".join(str(int(t)) for t in p.bead_types))
        for h, t, k, r0 in p.bonds:
            out.append(f"Bond {int(h)} {int(t)} {_f(k)} {_f(r0)}")
        for h, t, k, th in p.bond_pairs:
            out.append(f"BondPair {int(h)} {int(t)} {_f(k)} {_f(th)}")
        out.append("")
    if binary:
        out += ["", "BinaryData"]
        return ("\n".join(out) + "\n").encode() + b"\0" + _binary_beads(w)
    out += ["", "# Beads"]
    ptype = w.polymer_type_of
    for i in range(w.n_beads):
        pid = int(w.polymer_id[i])
        x = " ".join(g(w.x[i, d] + origin[d]) for d in range(3))
        v = " ".join(g(c) for c in w.v[i])
        f = " ".join(g(c) for c in w.f[i])
        out.append(f"B {i} {pid} {int(ptype[pid])} {int(w.polymer_offset[i])}  {x}  {v}  {f}")
    out.append("")
    return "\n".join(out) + "\n"


def write(name, w: World, precise=False, binary=False):
    data = dumps(w, precise=precise, binary=binary)
    with _open(name, "w") as fh:
        fh.write(data if binary else data.encode())


class _Lines:
    def __init__(self, text):
        self.lines = text.split("\n")
        self.i = 0

    def expect(self, prefix, n=None):
        while self.i < len(self.lines):
            parts = self.lines[self.i].replace("\r", " ").replace("\t", " ").split()
            self.i += 1
            if not parts or parts[0].startswith("#"):
                continue
            if parts[0] != prefix:
                raise ValueError(f"expected '{prefix}' but got '{parts[0]}' at line {self.i}")
            if n is not None and len(parts) != n:
                raise ValueError(f"expected {n} fields for '{prefix}' at line {self.i}")
            return parts
        raise ValueError(f"expected '{prefix}' but reached the end of the file")


def _num(s):
    try:
        return float(s)
    except ValueError:
        raise ValueError(f"not a number (parameter expressions are not supported): '{s}'") from None


def _strength(tok):
    """(value, None) for a number or a constant expression, (0.0, text) for
    an expression of variables, text in the reference's printed form."""
    try:
        return float(tok), None
    except ValueError:
        pass
    tree = expr.parse(tok)
    if tree[0] == "c":
        return tree[1], None
    return 0.0, expr.text(tree)


def _strengths(L, name, nt, kind, exprs):
    rows = []
    for _ in range(nt):
        p = L.expect(name, nt + 2)
        i = int(p[1])
        row = []
        for j, tok in enumerate(p[2:]):
            val, text = _strength(tok)
            if text is not None:
                exprs[(kind, i, j)] = text
            row.append(val)
        rows.append(row)
    return np.array(rows, dtype=np.float64)


def _read_binary_beads(blob, v2, box, ptypes, ptype_of, members, x, v, f):
    """The per-polymer binary records (layout in include/pdpd/world_state_io.hpp)."""
    nb, npoly = len(x), len(members)
    rec = 20 if v2 else 28
    where = np.full(nb, -1, np.int64)
    off, prev_p, prev_b = 0, 0xFFFFFFFF, 0xFFFFFFFF
    try:
        for _ in range(npoly):
            (h,) = struct.unpack_from("<H", blob, off)
            off += 2
            pt = h & 0x1FFF
            if (h >> 13) & 1:
                (d,) = struct.unpack_from("<I", blob, off)
                off += 4
            else:
                d = 1
            pid = (prev_p + d) & 0xFFFFFFFF
            prev_p = pid
            if pid >= npoly:
                raise ValueError("Polymer id out of range.")
            if ptype_of[pid] >= 0:
                raise ValueError("Polymer appears twice.")
            if pt >= len(ptypes):
                raise ValueError("polymer type out of range in binary data")
            ptype_of[pid] = pt
            for k in range(len(ptypes[pt].bead_types)):
                if (h >> (14 if k == 0 else 15)) & 1:
                    (d,) = struct.unpack_from("<I", blob, off)
                    off += 4
                else:
                    d = 1
                bid = (prev_b + d) & 0xFFFFFFFF
                prev_b = bid
                if bid >= nb:
                    raise ValueError("bead id out of range in binary data")
                if where[bid] >= 0:
                    raise ValueError("Bead appears twice.")
                where[bid] = off
                members[pid][k] = bid
                off += rec
            if off > len(blob):
                raise struct.error("short")
    except struct.error:
        raise ValueError("binary bead data ends early") from None
    if (where < 0).any():
        raise ValueError("a bead is missing from the binary data")
    raw = np.frombuffer(blob, np.uint8)[where[:, None] + np.arange(rec)]
    pos = raw[:, :12].copy().view("<u4").astype(np.float64)
    x[:] = np.ldexp(pos, -16)
    if not (x < np.asarray(box, np.float64)).all():
        raise ValueError("Bead was out of box.")
    if v2:
        vf = raw[:, 12:20].copy().view("<u4")
        v[:] = unpack_half3(vf[:, 0])
        f[:] = unpack_half3(vf[:, 1])
    else:
        vf = raw[:, 12:28].copy().view("<i2")
        v[:] = unpack_c3(vf[:, :4])
        f[:] = unpack_c3(vf[:, 4:])


def _split_binary(data: bytes):
    """Header text and the bytes after the NUL that follows the BinaryData line."""
    k = data.find(b"\nBinaryData")
    if k < 0:
        raise ValueError("expected 'BinaryData' but reached the end of the file")
    e = data.find(b"\n", k + 1)
    if e < 0:
        raise ValueError("Missing null before binary data.")
    i = e + 1
    while i < len(data) and data[i] != 0:
        if not chr(data[i]).isspace():
            raise ValueError("Unexpected char before null char.")
        i += 1
    if i >= len(data):
        raise ValueError("Missing null before binary data.")
    return data[:e + 1].decode(), data[i + 1:]


def loads(data) -> World:
    """A World from file contents (str or bytes) in any of the three formats."""
    if isinstance(data, str):
        data = data.encode()
    first = data[:data.find(b"\n") if b"\n" in data else len(data)].decode(errors="replace").split()
    fmt = first[1] if len(first) > 1 and first[0] == "WorldState" else ""
    binary = fmt in ("v1binary", "v2binary")
    text, blob = _split_binary(data) if binary else (data.decode(), None)
    L = _Lines(text)
    h = L.expect("WorldState", 6)
    if not binary and h[1] != "v0":
        raise ValueError(f"WorldState format '{h[1]}' is not supported (v0, v1binary, v2binary)")
    nt, npt, nb, npoly = (int(h[k]) for k in range(2, 6))
    p = L.expect("T", 3)
    t, dt = int(float(p[1])), _num(p[2])
    lam = _num(L.expect("Lambda", 2)[1])
    origin = np.array([_num(c) for c in L.expect("Origin", 4)[1:]])
    box = np.array([_num(c) for c in L.expect("Box", 4)[1:]])
    seed = int(L.expect("Seed", 2)[1])
    exprs = {}
    cons = _strengths(L, "ConservativeStrength", nt, "conservative", exprs)
    diss = _strengths(L, "DissipativeStrength", nt, "dissipative", exprs)
    names, stat = [], np.zeros(nt, np.uint8)
    for i in range(nt):
        p = L.expect("BeadType")
        if len(p) not in (4, 5):
            raise ValueError("expected 4 or 5 fields for BeadType")
        names.append(p[2])
        if len(p) == 5:
            if p[4] != "stationary":
                raise ValueError(f"unknown BeadType flag '{p[4]}'")
            stat[i] = 1
    ptypes = [None] * npt
    for _ in range(npt):
        p = L.expect("PolymerType", 6)
        pid, nbt, nbond, npair = int(p[1]), int(p[3]), int(p[4]), int(p[5])
        bt = [int(c) for c in L.expect("BeadTypeIndices", nbt + 1)[1:]]
        bonds = []
        for _ in range(nbond):
            q = L.expect("Bond", 5)
            bonds.append((int(q[1]), int(q[2]), _num(q[3]), _num(q[4])))
        pairs = []
        for _ in range(npair):
            q = L.expect("BondPair", 5)
            pairs.append((int(q[1]), int(q[2]), _num(q[3]), _num(q[4])))
        ptypes[pid] = PolymerType(p[2], bt, bonds, pairs)
    x = np.zeros((nb, 3)); v = np.zeros((nb, 3)); f = np.zeros((nb, 3))
    bead_type = np.zeros(nb, np.uint8); mono = np.zeros(nb, np.uint8)
    polymer_id = np.zeros(nb, np.uint32); offset = np.zeros(nb, np.uint8)
    ptype_of = np.full(npoly, -1, np.int64)
    members = [dict() for _ in range(npoly)]
    if binary:
        L.expect("BinaryData", 1)
        _read_binary_beads(blob, fmt == "v2binary", box, ptypes, ptype_of, members, x, v, f)
        for pid in range(npoly):
            for off, bid in members[pid].items():
                pt = int(ptype_of[pid])
                bead_type[bid] = ptypes[pt].bead_types[off]
                mono[bid] = 1 if len(ptypes[pt].bead_types) == 1 else 0
                polymer_id[bid] = pid
                offset[bid] = off
    for _ in range(0 if binary else nb):
        q = L.expect("B", 14)
        bid, pid, pt, off = int(q[1]), int(q[2]), int(q[3]), int(q[4])
        if ptype_of[pid] >= 0 and ptype_of[pid] != pt:
            raise ValueError("polymer type changes within a polymer")
        ptype_of[pid] = pt
        if off in members[pid]:
            raise ValueError("bead repeated within a polymer")
        members[pid][off] = bid
        x[bid] = [_num(q[5 + d]) - origin[d] for d in range(3)]
        x[bid][x[bid] == box] = 0.0
        v[bid] = [_num(q[8 + d]) for d in range(3)]
        f[bid] = [_num(q[11 + d]) for d in range(3)]
        bead_type[bid] = ptypes[pt].bead_types[off]
        mono[bid] = 1 if len(ptypes[pt].bead_types) == 1 else 0
        polymer_id[bid] = pid
        offset[bid] = off
    starts = np.zeros(npoly + 1, np.uint32)
    ids = []
    for pid in range(npoly):
        m = members[pid]
        if ptype_of[pid] < 0:
            raise ValueError("a polymer is missing beads")
        n = len(ptypes[int(ptype_of[pid])].bead_types)
        if sorted(m) != list(range(n)):
            raise ValueError("a polymer is missing beads")
        ids += [m[k] for k in range(n)]
        starts[pid + 1] = len(ids)
    if len(set(ids)) != nb or len(ids) != nb:
        raise ValueError("beads are missing or repeated")
    return World(box=box, dt=dt, lambda_=lam, seed=seed, t=t, bead_type_names=names, conservative=cons,
                 dissipative=diss, polymer_types=ptypes, polymer_type_of=ptype_of.astype(np.uint32),
                 polymer_bead_start=starts, polymer_bead_ids=np.asarray(ids, np.uint32), x=x, v=v, f=f,
                 bead_type=bead_type, is_monomer=mono, polymer_id=polymer_id, polymer_offset=offset,
                 stationary=stat if stat.any() else None, origin=origin, exprs=exprs or None)


def read(name) -> World:
    with _open(name, "r") as fh:
        return loads(fh.read())
