"""WorldState in flat numpy arrays, and the seeded generators of the workloads.

`World` carries exactly the fields of the reference's WorldState
(include/dpd/core/dpd_state.hpp:141-264 in poets-dpd): box, dt, lambda, seed,
t, the bead-type interaction matrix, the polymer types with their Hookean bonds
and bond pairs, the polymers, and per-bead x/v/f with the identity fields from
which BeadHash is built (dpd_state.hpp:18-60).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def bead_hash(bead_type, is_monomer, polymer_id, polymer_offset):
    """BeadHash::construct (dpd_state.hpp:52-60): bbbb 1ppp.. / bbbb 0ooo oooo pppp.."""
    bead_type = np.asarray(bead_type, dtype=np.uint32)
    is_monomer = np.asarray(is_monomer, dtype=np.uint32)
    polymer_id = np.asarray(polymer_id, dtype=np.uint32)
    polymer_offset = np.asarray(polymer_offset, dtype=np.uint32)
    return ((polymer_offset << np.uint32(20)) | polymer_id | (is_monomer << np.uint32(27))
            | (bead_type << np.uint32(28))).astype(np.uint32)


@dataclass
class PolymerType:
    name: str
    bead_types: list
    bonds: list = field(default_factory=list)        # (head_offset, tail_offset, kappa, r0)
    bond_pairs: list = field(default_factory=list)   # (head_bond, tail_bond, kappa, theta0)


@dataclass
class World:
    box: np.ndarray
    dt: float
    lambda_: float
    seed: int
    t: int
    bead_type_names: list
    conservative: np.ndarray       # (nt, nt)
    dissipative: np.ndarray        # (nt, nt)
    polymer_types: list            # [PolymerType]
    polymer_type_of: np.ndarray    # (P,) polymer type of each polymer
    polymer_bead_start: np.ndarray  # (P+1,)
    polymer_bead_ids: np.ndarray   # bead ids of polymers, concatenated
    x: np.ndarray                  # (n, 3) float64
    v: np.ndarray
    f: np.ndarray
    bead_type: np.ndarray          # (n,) uint8
    is_monomer: np.ndarray         # (n,) uint8
    polymer_id: np.ndarray         # (n,) uint32
    polymer_offset: np.ndarray     # (n,) uint8
    stationary: np.ndarray = None  # (nt,) uint8 or None
    origin: np.ndarray = None      # (3,) offset of the box corner in files, or None for 0
    # strengths given as expressions of the cell position (x, y, z, ux, uy, uz):
    # {("conservative" | "dissipative", i, j): expression text}; those entries of
    # the constant matrices are unused (0). None or {}: constant strengths.
    exprs: dict = None

    @property
    def n_beads(self):
        return int(self.x.shape[0])

    @property
    def n_types(self):
        return int(self.conservative.shape[0])

    @property
    def spatial(self):
        return bool(self.exprs)

    def set_expression(self, kind, i, j, src, symmetric=True):
        """Make interaction (i, j) (and (j, i)) of kind "conservative" or
        "dissipative" the expression src; a bare number sets the constant
        matrix instead (an expression of no variable stays an expression, as
        in the reference's Parameter). The text is kept in the reference's printed
        form (every binary operation in parentheses)."""
        from . import expr
        if kind not in ("conservative", "dissipative"):
            raise ValueError(f"unknown strength kind '{kind}'")
        tree = expr.parse(src)
        unknown = expr.variables(tree) - set(expr.SPACE)
        if unknown:
            raise ValueError(f"variables other than the cell position: {sorted(unknown)}")
        for a, b in {(i, j), (j, i)} if symmetric else {(i, j)}:
            mat = self.conservative if kind == "conservative" else self.dissipative
            if self.exprs is None:
                self.exprs = {}
            if tree[0] == "c":
                mat[a, b] = tree[1]
                self.exprs.pop((kind, a, b), None)
            else:
                mat[a, b] = 0.0
                self.exprs[(kind, a, b)] = expr.text(tree)

    def cell_strengths(self):
        """(conservative, dissipative), each (n_cells, nt, nt) float64: the
        strengths at every cell, cell index (x * Ly + y) * Lz + z, expressions
        evaluated with the cell's bindings in double precision
        (naive_dpd_engine.hpp:158-183)."""
        from . import expr
        L = [int(b) for b in self.box]
        nc, nt = L[0] * L[1] * L[2], self.n_types
        cons = np.broadcast_to(self.conservative, (nc, nt, nt)).copy()
        diss = np.broadcast_to(self.dissipative, (nc, nt, nt)).copy()
        trees = {k: expr.parse(v) for k, v in (self.exprs or {}).items()}
        if not trees:
            return cons, diss
        for cx in range(L[0]):
            for cy in range(L[1]):
                for cz in range(L[2]):
                    b = expr.cell_bindings((cx, cy, cz), L)
                    c = (cx * L[1] + cy) * L[2] + cz
                    for (kind, i, j), tree in trees.items():
                        (cons if kind == "conservative" else diss)[c, i, j] = expr.evaluate(tree, b)
        return cons, diss

    @property
    def hash(self):
        return bead_hash(self.bead_type, self.is_monomer, self.polymer_id, self.polymer_offset)

    def copy(self):
        import copy
        return copy.deepcopy(self)

    # Flattened topology in the reference's iteration order: per polymer, the
    # bonds of its type then the bond pairs (naive_dpd_engine.hpp:567-582).
    def flat_topology(self):
        """Per polymer (in polymer order) the bonds of its type, then its bond
        pairs as (head, middle, tail) bead indices, with per-polymer start
        offsets. Built per polymer type with array indexing; the order is the
        per-polymer loop's."""
        n_poly = len(self.polymer_type_of)
        ptype = np.asarray(self.polymer_type_of, dtype=np.int64)
        pstart = np.asarray(self.polymer_bead_start, dtype=np.int64)
        ids = np.asarray(self.polymer_bead_ids, dtype=np.int64)
        nb_t = np.array([len(pt.bonds) for pt in self.polymer_types], dtype=np.int64)
        na_t = np.array([len(pt.bond_pairs) for pt in self.polymer_types], dtype=np.int64)
        nb_p = nb_t[ptype] if n_poly else np.zeros(0, np.int64)
        na_p = na_t[ptype] if n_poly else np.zeros(0, np.int64)
        bond_start = np.concatenate([[0], np.cumsum(nb_p)]).astype(np.int64)
        angle_start = np.concatenate([[0], np.cumsum(na_p)]).astype(np.int64)
        nb, na = int(bond_start[-1]), int(angle_start[-1])
        bh = np.zeros(nb, np.int64); bt = np.zeros(nb, np.int64)
        bk = np.zeros(nb, np.float64); br = np.zeros(nb, np.float64)
        ah = np.zeros(na, np.int64); am = np.zeros(na, np.int64); at = np.zeros(na, np.int64)
        ak = np.zeros(na, np.float64); a0 = np.zeros(na, np.float64)
        for t, pt in enumerate(self.polymer_types):
            polys = np.nonzero(ptype == t)[0]
            if polys.size == 0:
                continue
            base = pstart[polys]
            for j, (h, tl, k, r0) in enumerate(pt.bonds):
                dst = bond_start[polys] + j
                bh[dst] = ids[base + h]; bt[dst] = ids[base + tl]; bk[dst] = k; br[dst] = r0
            for j, (hb, tb, k, th) in enumerate(pt.bond_pairs):
                dst = angle_start[polys] + j
                ah[dst] = ids[base + pt.bonds[hb][0]]; am[dst] = ids[base + pt.bonds[hb][1]]
                at[dst] = ids[base + pt.bonds[tb][1]]; ak[dst] = k; a0[dst] = th
        i32 = lambda a: np.asarray(a, dtype=np.int32)
        return dict(poly_bond_start=i32(bond_start), poly_angle_start=i32(angle_start),
                    bond_head=i32(bh), bond_tail=i32(bt), bond_kappa=bk, bond_r0=br,
                    angle_head=i32(ah), angle_mid=i32(am), angle_tail=i32(at),
                    angle_kappa=ak, angle_theta0=a0)


class WorldBuilder:
    """Incremental construction (the role of WorldStateBuilder,
    dpd_state_builder.hpp:18-232): bead types, interactions, polymer types,
    polymers. Bead ids and polymer ids are contiguous from zero."""

    def __init__(self, box, dt=0.01, lambda_=0.5, seed=1, t=0):
        self.box = np.asarray(box, dtype=np.float64)
        if np.any(np.round(self.box) != self.box):
            raise ValueError("Dimensions must be integer.")
        self.dt, self.lambda_, self.seed, self.t = dt, lambda_, seed, t
        self.type_names = []
        self.cons = {}
        self.diss = {}
        self.ptypes = []
        self.poly_type_of = []
        self.poly_beads = []
        self.xs, self.vs, self.fs = [], [], []
        self.btype, self.mono, self.pid, self.poff = [], [], [], []
        self.default = (1.0, 0.1)   # builder defaults, dpd_state_builder.hpp:21-22

    def add_bead_type(self, name):
        if name in self.type_names:
            raise ValueError("Bead type name already exists.")
        self.type_names.append(name)
        return len(self.type_names) - 1

    def set_interaction(self, a, b, conservative, dissipative):
        ia = a if isinstance(a, int) else self.type_names.index(a)
        ib = b if isinstance(b, int) else self.type_names.index(b)
        self.cons[(ia, ib)] = self.cons[(ib, ia)] = float(conservative)
        self.diss[(ia, ib)] = self.diss[(ib, ia)] = float(dissipative)

    def add_polymer_type(self, name, bead_types, bonds=(), bond_pairs=()):
        bt = [b if isinstance(b, int) else self.type_names.index(b) for b in bead_types]
        self.ptypes.append(PolymerType(name, bt, list(bonds), list(bond_pairs)))
        return len(self.ptypes) - 1

    def add_polymer(self, ptype, xs, vs=None, fs=None):
        pt = self.ptypes[ptype]
        n = len(pt.bead_types)
        xs = np.asarray(xs, dtype=np.float64).reshape(n, 3)
        vs = np.zeros((n, 3)) if vs is None else np.asarray(vs, dtype=np.float64).reshape(n, 3)
        fs = np.zeros((n, 3)) if fs is None else np.asarray(fs, dtype=np.float64).reshape(n, 3)
        pid = len(self.poly_type_of)
        base = len(self.btype)
        ids = []
        for i in range(n):
            if np.any(xs[i] < 0) or np.any(xs[i] >= self.box):
                raise ValueError("Bead out of box.")
            ids.append(base + i)
            self.xs.append(xs[i]); self.vs.append(vs[i]); self.fs.append(fs[i])
            self.btype.append(pt.bead_types[i]); self.mono.append(1 if n == 1 else 0)
            self.pid.append(pid); self.poff.append(i)
        self.poly_type_of.append(ptype)
        self.poly_beads.append(ids)
        return pid

    def build(self):
        nt = len(self.type_names)
        cons = np.zeros((nt, nt))
        diss = np.zeros((nt, nt))
        for i in range(nt):
            for j in range(nt):
                cons[i, j] = self.cons.get((i, j), self.default[0])
                diss[i, j] = self.diss.get((i, j), self.default[1])
        starts = np.zeros(len(self.poly_beads) + 1, dtype=np.uint32)
        starts[1:] = np.cumsum([len(p) for p in self.poly_beads])
        ids = np.asarray([i for p in self.poly_beads for i in p], dtype=np.uint32)
        n = len(self.btype)
        return World(
            box=self.box.copy(), dt=self.dt, lambda_=self.lambda_, seed=self.seed, t=self.t,
            bead_type_names=list(self.type_names), conservative=cons, dissipative=diss,
            polymer_types=self.ptypes, polymer_type_of=np.asarray(self.poly_type_of, dtype=np.uint32),
            polymer_bead_start=starts, polymer_bead_ids=ids,
            x=np.asarray(self.xs, dtype=np.float64).reshape(n, 3),
            v=np.asarray(self.vs, dtype=np.float64).reshape(n, 3),
            f=np.asarray(self.fs, dtype=np.float64).reshape(n, 3),
            bead_type=np.asarray(self.btype, dtype=np.uint8),
            is_monomer=np.asarray(self.mono, dtype=np.uint8),
            polymer_id=np.asarray(self.pid, dtype=np.uint32),
            polymer_offset=np.asarray(self.poff, dtype=np.uint8))


# ----------------------------------------------------------------- workloads --

CHAIN_TYPES = ["H", "H", "T", "T", "T", "T"]
CHAIN_BONDS = [(i, i + 1, 128.0, 0.5) for i in range(5)]
CHAIN_ANGLES = [(i, i + 1, 20.0, 0.0) for i in range(4)]


def _gaussian_velocities(rng, n):
    v = rng.standard_normal((n, 3))
    if n > 0:
        v -= v.mean(axis=0)        # zero net momentum
    return v


def make_world(L, density=3.0, bonded_fraction=0.0, seed=1, a_matrix=None, gamma=4.5, dt=0.01,
               n_beads=None, theta0=0.0):
    """Water (type W) plus H2T4 chains with Hookean k=128, r0=0.5 and angle
    bonds k=20 with preferred angle theta0 (0 = straight). Positions uniform from
    the seed, velocities Gaussian at kT=1 with zero net momentum, forces zero.
    Chains are numbered first (polymer ids below 2^20, as BeadHash requires),
    then the monomers."""
    rng = np.random.default_rng(seed)
    box = np.asarray(L, dtype=np.float64)
    n = int(round(density * float(np.prod(box)))) if n_beads is None else int(n_beads)
    clen = len(CHAIN_TYPES)
    n_chains = int(round(bonded_fraction * n / clen))
    if n_chains >= (1 << 20):
        raise ValueError("too many chains for the 20-bit polymer id of BeadHash")
    three = n_chains > 0 or (a_matrix is not None and np.shape(a_matrix)[0] == 3)
    names = ["W", "H", "T"] if three else ["W"]
    nt = len(names)
    if a_matrix is None:
        a_matrix = np.full((nt, nt), 25.0)
    cons = np.array(np.asarray(a_matrix, dtype=np.float64)[:nt, :nt])
    cons = np.triu(cons) + np.triu(cons, 1).T       # symmetric from the upper triangle
    diss = np.full((nt, nt), float(gamma))
    ptypes = [PolymerType("W", [0])]
    if n_chains:
        angles = [(h, t, k, float(theta0)) for h, t, k, _ in CHAIN_ANGLES]
        ptypes.append(PolymerType("HHTTTT", [1, 1, 2, 2, 2, 2], list(CHAIN_BONDS), angles))
    n_mono = n - n_chains * clen
    v = _gaussian_velocities(rng, n)
    # chains: random walks of step 0.5 from a uniform start, wrapped
    start = rng.random((n_chains, 3)) * box
    steps = rng.standard_normal((n_chains, clen - 1, 3))
    steps *= 0.5 / np.linalg.norm(steps, axis=2, keepdims=True)
    chain_x = np.concatenate([start[:, None, :], start[:, None, :] + np.cumsum(steps, axis=1)], axis=1)
    chain_x = np.mod(chain_x, box) % box
    mono_x = rng.random((n_mono, 3)) * box
    x = np.concatenate([chain_x.reshape(-1, 3), mono_x]) if n_chains else mono_x
    x = np.minimum(x, np.nextafter(box, 0))
    nb = n_chains * clen
    bead_type = np.zeros(n, np.uint8)
    is_mono = np.ones(n, np.uint8)
    pid = np.zeros(n, np.uint32)
    poff = np.zeros(n, np.uint8)
    if n_chains:
        bead_type[:nb] = np.tile(np.array([1, 1, 2, 2, 2, 2], np.uint8), n_chains)
        is_mono[:nb] = 0
        pid[:nb] = np.repeat(np.arange(n_chains, dtype=np.uint32), clen)
        poff[:nb] = np.tile(np.arange(clen, dtype=np.uint8), n_chains)
    pid[nb:] = np.arange(n_chains, n_chains + n_mono, dtype=np.uint32)
    n_poly = n_chains + n_mono
    ptype_of = np.zeros(n_poly, np.uint32)
    ptype_of[:n_chains] = 1
    starts = np.concatenate([np.arange(0, nb + 1, clen, dtype=np.uint32)[:n_chains + 1] if n_chains
                             else np.zeros(1, np.uint32),
                             np.arange(nb + 1, n + 1, dtype=np.uint32)])
    return World(box=box, dt=dt, lambda_=0.5, seed=int(seed) & 0xFFFFFFFFFFFFFFFF, t=0, bead_type_names=names,
                 conservative=cons, dissipative=diss, polymer_types=ptypes, polymer_type_of=ptype_of,
                 polymer_bead_start=starts, polymer_bead_ids=np.arange(n, dtype=np.uint32),
                 x=x, v=v, f=np.zeros((n, 3)), bead_type=bead_type, is_monomer=is_mono,
                 polymer_id=pid, polymer_offset=poff)


def make_lattice_world(L=4, spacing=0.5, nudge=0.3, seed=3):
    """Monomers on a cubic lattice of the given spacing (density 8 at 0.5), a
    fraction `nudge` of them moved up by one FP32 ulp per axis: many pairs at
    exactly r = 1 and beads on cell faces, edges and corners, where the pair
    test and neighbour-cell culling decide at the boundary."""
    b = WorldBuilder((L, L, L), dt=0.01, seed=seed)
    b.add_bead_type("A")
    b.set_interaction("A", "A", 25.0, 4.5)
    m = b.add_polymer_type("A", ["A"])
    rng = np.random.default_rng(seed)
    g = np.arange(int(round(L / spacing))) * spacing
    for x in g:
        for y in g:
            for z in g:
                p = np.array([x, y, z])
                if rng.random() < nudge:
                    p = np.nextafter(p.astype(np.float32), np.float32(L)).astype(np.float64)
                b.add_polymer(m, p, rng.standard_normal(3) * 0.1)
    return b.build()


def add_wall(w, thickness=1.0, a_wall=25.0, gamma_wall=4.5):
    """Turn the monomers with x < thickness into a new stationary bead type "S"
    (BeadType::stationary; a wall): velocities set to zero, interactions of S
    with every type a_wall / gamma_wall. Returns a new World."""
    w = w.copy()
    nt = w.n_types
    sel = (w.is_monomer == 1) & (w.x[:, 0] < thickness)
    cons = np.full((nt + 1, nt + 1), float(a_wall))
    diss = np.full((nt + 1, nt + 1), float(gamma_wall))
    cons[:nt, :nt] = w.conservative
    diss[:nt, :nt] = w.dissipative
    w.conservative, w.dissipative = cons, diss
    w.bead_type_names = list(w.bead_type_names) + ["S"]
    w.stationary = np.zeros(nt + 1, np.uint8)
    w.stationary[nt] = 1
    w.polymer_types = list(w.polymer_types) + [PolymerType("S", [nt])]
    w.bead_type = w.bead_type.copy()
    w.bead_type[sel] = nt
    ptype_of = w.polymer_type_of.copy()
    ptype_of[w.polymer_id[sel]] = len(w.polymer_types) - 1
    w.polymer_type_of = ptype_of
    w.v = w.v.copy()
    w.v[sel] = 0.0
    return w


def random_a_matrix(rng, choices=(25.0, 35.0, 50.0, 75.0), nt=3):
    a = np.zeros((nt, nt))
    for i in range(nt):
        for j in range(i, nt):
            a[i, j] = a[j, i] = rng.choice(choices)
    return a


def config_water(seed=1):
    """16x16x16 cells, density 3, one bead type, a=25, gamma=4.5, dt=0.01."""
    return make_world((16, 16, 16), seed=seed)


def config_polymer(seed=1, L=32):
    """32^3 cells, density 3, W/H/T with 10% of beads in H2T4 chains, a_ij drawn
    from {25,35,50,75} per type pair."""
    rng = np.random.default_rng(seed + 1000)
    return make_world((L, L, L), bonded_fraction=0.10, seed=seed, a_matrix=random_a_matrix(rng))


def config_large(seed=1, L=128):
    """128^3 cells, density 3, 5% bonded."""
    rng = np.random.default_rng(seed + 1000)
    return make_world((L, L, L), bonded_fraction=0.05, seed=seed, a_matrix=random_a_matrix(rng))


def batch_world_spec(i, seed=1, lo=8, hi=32):
    """Shape and parameters of world i of the batched ensemble: sides uniform in
    [lo, hi] cells (independently), a_ij uniform in [25, 80], bonded fraction
    uniform in [0, 0.2], its own seed."""
    rng = np.random.default_rng([seed, i])
    L = tuple(int(v) for v in rng.integers(lo, hi + 1, size=3))
    a = np.zeros((3, 3))
    for p in range(3):
        for q in range(p, 3):
            a[p, q] = a[q, p] = rng.uniform(25.0, 80.0)
    frac = float(rng.uniform(0.0, 0.2))
    wseed = int(rng.integers(1, 2**62))
    return dict(L=L, a_matrix=a, bonded_fraction=frac, seed=wseed)


def config_batch(n_worlds=256, seed=1, lo=8, hi=32):
    out = []
    for i in range(n_worlds):
        s = batch_world_spec(i, seed, lo, hi)
        out.append(make_world(s["L"], bonded_fraction=s["bonded_fraction"], seed=s["seed"],
                              a_matrix=s["a_matrix"]))
    return out


def served_requests(n_requests, seed=1, rate=50.0, lo=8, hi=48, steps_lo=1000, steps_hi=20000):
    """Poisson arrivals (exponential gaps at `rate` per second), cubic sides
    log-uniform in [lo, hi], step counts uniform in [steps_lo, steps_hi]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate)
        side = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        a = np.zeros((3, 3))
        for p in range(3):
            for q in range(p, 3):
                a[p, q] = a[q, p] = rng.uniform(25.0, 80.0)
        reqs.append(dict(id=i, arrival=t, L=(side, side, side), steps=steps, a_matrix=a,
                         bonded_fraction=float(rng.uniform(0.0, 0.2)), seed=int(rng.integers(1, 2**62))))
    return reqs
