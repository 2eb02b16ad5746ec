"""Synthetic decks: the reference's structured test meshes and the five
workloads the path is measured on, generated from seeded SplitMix64 streams.

Structured builders follow crates/solvix-core/src/assembly.rs:554-624
(`quad_mesh`: row-major node ids from 1; `hex_mesh`: k-major node ids, C3D8
order bottom face CCW then top face CCW). Uniform doubles are
`(x >> 11) * 2^-53`, the mapping of assembly.rs:807. Jitter moves interior
nodes only, by `(2u - 1) * a * h_axis`; boundary, clamped and loaded nodes
stay exact.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import _abi
from .model import Model

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

SEED_BASE = 0x53170000  # workload k uses seed SEED_BASE + k


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014), vectorised over a counter."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def next_u64(self, k: int) -> np.ndarray:
        idx = np.arange(self.count + 1, self.count + k + 1, dtype=np.uint64)
        self.count += k
        with np.errstate(over="ignore"):
            z = self.state + idx * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))

    def uniform(self, k: int = 1) -> np.ndarray:
        return (self.next_u64(k) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def u(self) -> float:
        return float(self.uniform(1)[0])

    def log_uniform(self, lo: float, hi: float) -> float:
        return math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * self.u())

    def between(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.u()


# ------------------------------------------------------------ structured
def quad_grid(nx: int, ny: int, width: float, height: float):
    """Node ids/coords and CPS4 connectivity of assembly.rs:554-580."""
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    i, j = i.ravel(), j.ravel()
    coords = np.zeros((i.size, 3))
    coords[:, 0] = width * i.astype(np.float64) / nx
    coords[:, 1] = height * j.astype(np.float64) / ny
    ids = (j * (nx + 1) + i + 1).astype(np.uint32)
    ej, ei = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ei, ej = ei.ravel(), ej.ravel()
    nid = lambda a, b: (b * (nx + 1) + a + 1)
    conn = np.stack([nid(ei, ej), nid(ei + 1, ej), nid(ei + 1, ej + 1), nid(ei, ej + 1)], 1).astype(np.uint32)
    return ids, coords, conn


def hex_grid(nx: int, ny: int, nz: int, lx: float, ly: float, lz: float):
    """Node ids/coords and C3D8 connectivity of assembly.rs:584-624."""
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    coords = np.stack([lx * i.astype(np.float64) / nx, ly * j.astype(np.float64) / ny,
                       lz * k.astype(np.float64) / nz], 1)
    nid = lambda a, b, c: (c * (ny + 1) * (nx + 1) + b * (nx + 1) + a + 1)
    ids = nid(i, j, k).astype(np.uint32)
    ek, ej, ei = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()
    conn = np.stack([nid(ei, ej, ek), nid(ei + 1, ej, ek), nid(ei + 1, ej + 1, ek), nid(ei, ej + 1, ek),
                     nid(ei, ej, ek + 1), nid(ei + 1, ej, ek + 1), nid(ei + 1, ej + 1, ek + 1),
                     nid(ei, ej + 1, ek + 1)], 1).astype(np.uint32)
    return ids, coords, conn


def build_model(ids, coords, conn, kind, E, nu, thickness=1.0, bcs=(), cl_node=(), cl_dof=(), cl_val=(),
                name="model", kinds=None) -> Model:
    """Single-material model over elset ALL; conn is [n_elem, nen] node ids
    (or a list of arrays when kinds differ per element)."""
    if kinds is None:
        conn = np.asarray(conn, np.uint32)
        n_elem = conn.shape[0]
        kinds = np.full(n_elem, kind, np.uint8)
        conn_ptr = np.arange(n_elem + 1, dtype=np.int64) * conn.shape[1]
        conn_ids = conn.ravel()
    else:
        n_elem = len(conn)
        conn_ptr = np.zeros(n_elem + 1, np.int64)
        conn_ptr[1:] = np.cumsum([len(c) for c in conn])
        conn_ids = np.concatenate(conn).astype(np.uint32)
    elem_ids = np.arange(1, n_elem + 1, dtype=np.uint32)
    return Model(
        node_ids=np.asarray(ids, np.uint32), coords=np.asarray(coords, np.float64),
        elem_ids=elem_ids, kinds=np.asarray(kinds, np.uint8), conn_ptr=conn_ptr, conn_ids=conn_ids,
        materials=[("MAT", float(E), float(nu))], elsets=[("ALL", elem_ids)], assignments=[("ALL", "MAT")],
        thickness=float(thickness),
        bcs=[(np.asarray(n, np.uint32), np.asarray(d, np.int32)) for n, d in bcs],
        cload_node=np.asarray(cl_node, np.uint32), cload_dof=np.asarray(cl_dof, np.int32),
        cload_value=np.asarray(cl_val, np.float64), name=name)


def _jitter(coords, interior, rng: SplitMix64, amp: float, h: Tuple[float, ...]):
    idx = np.nonzero(interior)[0]
    nd = len(h)
    u = rng.uniform(idx.size * nd).reshape(idx.size, nd)
    for a in range(nd):
        coords[idx, a] += (2.0 * u[:, a] - 1.0) * amp * h[a]


def quad_model(nx, ny, width, height, E=210e9, nu=0.3, thickness=1.0, jitter=0.0, rng=None, name="quad"):
    ids, coords, conn = quad_grid(nx, ny, width, height)
    if jitter > 0:
        i = (ids - 1) % (nx + 1)
        j = (ids - 1) // (nx + 1)
        _jitter(coords, (i > 0) & (i < nx) & (j > 0) & (j < ny), rng, jitter, (width / nx, height / ny))
    return build_model(ids, coords, conn, _abi.SVX_CPS4, E, nu, thickness, name=name)


def hex_model(nx, ny, nz, lx, ly, lz, E=210e9, nu=0.3, jitter=0.0, rng=None, name="hex"):
    ids, coords, conn = hex_grid(nx, ny, nz, lx, ly, lz)
    if jitter > 0:
        t = ids.astype(np.int64) - 1
        i = t % (nx + 1)
        j = (t // (nx + 1)) % (ny + 1)
        k = t // ((nx + 1) * (ny + 1))
        inter = (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)
        _jitter(coords, inter, rng, jitter, (lx / nx, ly / ny, lz / nz))
    return build_model(ids, coords, conn, _abi.SVX_C3D8, E, nu, name=name)


def left_edge_ids(nx, ny):
    return (np.arange(ny + 1) * (nx + 1) + 1).astype(np.uint32)


def x0_face_ids(nx, ny, nz):
    k, j = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), indexing="ij")
    return (k.ravel() * (ny + 1) * (nx + 1) + j.ravel() * (nx + 1) + 1).astype(np.uint32)


# ------------------------------------------------------------ the README decks
def readme_cantilever(nx: int, ny: int) -> Model:
    """The CPS4 cantilever family of docs/benchmarks/cantilever.md:8-16:
    2 m x 0.1 m, t = 0.01, steel, left edge fixed in dofs 1-2, nominal
    1000 N downward split over the ny + 1 tip nodes (-200 N on four-row
    meshes, -111.111 N on eight-row meshes). Coordinates pass through the
    deck's 6-decimal text, as in examples/cantilever.inp."""
    ids, coords, conn = quad_grid(nx, ny, 2.0, 0.1)
    coords = np.array([[float(f"{v:.6f}") for v in c] for c in coords])
    tip = (np.arange(ny + 1) * (nx + 1) + nx + 1).astype(np.uint32)
    val = {4: -200.0, 8: -111.111}[ny]
    return build_model(ids, coords, conn, _abi.SVX_CPS4, 210e9, 0.3, 0.01,
                       bcs=[(left_edge_ids(nx, ny), [1, 2])], cl_node=tip,
                       cl_dof=np.full(tip.size, 2), cl_val=np.full(tip.size, val), name=f"cantilever_{nx}x{ny}")


# ------------------------------------------------------------ workloads
def cfg1(seed: int = SEED_BASE + 1, nx: int = 400, ny: int = 40) -> Model:
    """2D CPS4 cantilever: 10 m x 1 m on 400x40 (16,000 elements, 16,441
    nodes, n = 32,882), interior jitter +/-0.15 h, E = 210e9, nu = 0.3,
    t = 0.01, left edge fixed in dofs 1-2, -1e4 N in dof 2 at the tip
    mid-height node. Other nx x ny: the same cantilever on a finer mesh."""
    rng = SplitMix64(seed)
    m = quad_model(nx, ny, 10.0, 1.0, 210e9, 0.3, 0.01, jitter=0.15, rng=rng, name=f"cfg1_cps4_{nx}x{ny}")
    m.bcs = [(left_edge_ids(nx, ny), np.array([1, 2], np.int32))]
    tip = (ny // 2) * (nx + 1) + nx + 1
    m.cload_node, m.cload_dof, m.cload_value = (np.array([tip], np.uint32), np.array([2], np.int32),
                                                np.array([-1e4]))
    return m


def face_traction_loads(nx, ny, nz, lx, ly, lz, pressure, dof):
    """Consistent nodal loads of a uniform traction on the x = lx face."""
    k, j = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), indexing="ij")
    k, j = k.ravel(), j.ravel()
    ids = (k * (ny + 1) * (nx + 1) + j * (nx + 1) + nx + 1).astype(np.uint32)
    wj = np.where((j == 0) | (j == ny), 0.5, 1.0)
    wk = np.where((k == 0) | (k == nz), 0.5, 1.0)
    area = (ly / ny) * (lz / nz)
    return ids, np.full(ids.size, dof, np.int32), pressure * area * wj * wk


def cfg2(seed: int = SEED_BASE + 2, nx: int = 200, ny: int = 20, nz: int = 20) -> Model:
    """3D C3D8 block 20 x 2 x 2 m on 200x20x20 (80,000 elements, 88,641
    nodes, n = 265,923), interior jitter +/-0.1 h, E = 70e9, nu = 0.33, x = 0
    face clamped, uniform -1e6 Pa z-traction on the x = 20 face as
    consistent nodal loads (corners -2,500 N, edges -5,000 N, interior
    -10,000 N). Smaller nx/ny/nz keep the same shape for tests."""
    lx, ly, lz = 20.0 * nx / 200, 2.0 * ny / 20, 2.0 * nz / 20
    rng = SplitMix64(seed)
    m = hex_model(nx, ny, nz, lx, ly, lz, 70e9, 0.33, jitter=0.1, rng=rng, name=f"cfg2_c3d8_{nx}x{ny}x{nz}")
    m.bcs = [(x0_face_ids(nx, ny, nz), np.array([1, 2, 3], np.int32))]
    m.cload_node, m.cload_dof, m.cload_value = face_traction_loads(nx, ny, nz, lx, ly, lz, -1e6, 3)
    return m


def plane_instance(rng: SplitMix64, n_lo=2e3, n_hi=2e5, name="inst") -> Model:
    """One instance of the batched 2D distribution: n ~ logU[n_lo, n_hi],
    aspect ~ U(2, 20), ny = round(sqrt(n / (2 aspect))), nx = round(aspect ny),
    height 1 m with square cells; CPS4 with probability 0.6, else CPS3 with a
    random diagonal per quad; E ~ U[70e9, 210e9], nu ~ U[0.25, 0.35],
    t ~ U[0.005, 0.05]; left edge clamped; tip mid-node load |P| ~
    logU[1e2, 1e5] at angle ~ U[0, 2 pi); interior jitter +/-0.15 h."""
    n_target = rng.log_uniform(n_lo, n_hi)
    aspect = rng.between(2.0, 20.0)
    ny = max(1, int(round(math.sqrt(n_target / (2.0 * aspect)))))
    nx = max(1, int(round(aspect * ny)))
    quad = rng.u() < 0.6
    E = rng.between(70e9, 210e9)
    nu = rng.between(0.25, 0.35)
    t = rng.between(0.005, 0.05)
    P = rng.log_uniform(1e2, 1e5)
    th = rng.between(0.0, 2.0 * math.pi)
    width, height = float(nx) / ny, 1.0
    ids, coords, conn = quad_grid(nx, ny, width, height)
    kind = _abi.SVX_CPS4
    if not quad:
        diag = rng.uniform(conn.shape[0]) < 0.5
        # split along 1-3 (diag) or 2-4, both counter-clockwise triangles
        tri_a = np.where(diag[:, None], conn[:, [0, 1, 2]], conn[:, [0, 1, 3]])
        tri_b = np.where(diag[:, None], conn[:, [0, 2, 3]], conn[:, [1, 2, 3]])
        conn = np.stack([tri_a, tri_b], 1).reshape(-1, 3)
        kind = _abi.SVX_CPS3
    i = (ids.astype(np.int64) - 1) % (nx + 1)
    j = (ids.astype(np.int64) - 1) // (nx + 1)
    _jitter(coords, (i > 0) & (i < nx) & (j > 0) & (j < ny), rng, 0.15, (width / nx, height / ny))
    tip = (ny // 2) * (nx + 1) + nx + 1
    return build_model(ids, coords, conn, kind, E, nu, t, bcs=[(left_edge_ids(nx, ny), [1, 2])],
                       cl_node=[tip, tip], cl_dof=[1, 2], cl_val=[P * math.cos(th), P * math.sin(th)],
                       name=f"{name}_{'cps4' if quad else 'cps3'}_{nx}x{ny}")


def cfg3(count: int = 512, seed: int = SEED_BASE + 3, n_lo=2e3, n_hi=2e5) -> List[Model]:
    """512 uneven plane-stress instances; instance i draws from
    SplitMix64(seed ^ i)."""
    return [plane_instance(SplitMix64(seed ^ i), n_lo, n_hi, name=f"cfg3_{i}") for i in range(count)]


def random_unit(rng: SplitMix64) -> np.ndarray:
    z = 2.0 * rng.u() - 1.0
    phi = 2.0 * math.pi * rng.u()
    s = math.sqrt(max(0.0, 1.0 - z * z))
    return np.array([s * math.cos(phi), s * math.sin(phi), z])


def cfg4(seed: int = SEED_BASE + 4, nrhs: int = 64, nx=120, ny=12, nz=12):
    """Shared C3D8 cantilever 120x12x12 on 2 x 0.2 x 0.2 m (n = 61,347),
    E = 210e9, nu = 0.3, x = 0 face clamped; `nrhs` load cases, each with
    1-8 point loads at random nodes of the free boundary faces, direction
    uniform on the sphere, |F| ~ U[1e3, 1e5]. Returns (model without loads,
    list of (node_ids, dofs, values) per load case)."""
    rng = SplitMix64(seed)
    lx, ly, lz = 2.0 * nx / 120, 0.2 * ny / 12, 0.2 * nz / 12
    m = hex_model(nx, ny, nz, lx, ly, lz, 210e9, 0.3, name=f"cfg4_c3d8_{nx}x{ny}x{nz}")
    m.bcs = [(x0_face_ids(nx, ny, nz), np.array([1, 2, 3], np.int32))]
    t = m.node_ids.astype(np.int64) - 1
    i = t % (nx + 1)
    j = (t // (nx + 1)) % (ny + 1)
    k = t // ((nx + 1) * (ny + 1))
    free = (i > 0) & ((i == nx) | (j == 0) | (j == ny) | (k == 0) | (k == nz))
    cand = m.node_ids[free]
    cases = []
    for _ in range(nrhs):
        nl = 1 + min(7, int(8 * rng.u()))
        nodes, dofs, vals = [], [], []
        for _ in range(nl):
            nid = cand[min(cand.size - 1, int(cand.size * rng.u()))]
            d = random_unit(rng)
            mag = rng.between(1e3, 1e5)
            for a in range(3):
                nodes.append(nid)
                dofs.append(a + 1)
                vals.append(mag * d[a])
        cases.append((np.array(nodes, np.uint32), np.array(dofs, np.int32), np.array(vals)))
    return m, cases


def solid_instance(rng: SplitMix64, n_lo=1e4, n_hi=3e5, name="blk") -> Model:
    """A served C3D8 block: n ~ logU[n_lo, n_hi], aspect ~ U(2, 10), square
    1 m section with ny = nz = round((n / (3 aspect))^(1/3)), nx = round(aspect
    ny); E ~ U[70e9, 210e9], nu ~ U[0.25, 0.35]; x = 0 face clamped; tip
    centre load |P| ~ logU[1e2, 1e5] in a direction uniform on the sphere;
    interior jitter +/-0.1 h."""
    n_target = rng.log_uniform(n_lo, n_hi)
    aspect = rng.between(2.0, 10.0)
    ny = max(1, int(round((n_target / (3.0 * aspect)) ** (1.0 / 3.0))))
    nx = max(1, int(round(aspect * ny)))
    nz = ny
    E = rng.between(70e9, 210e9)
    nu = rng.between(0.25, 0.35)
    P = rng.log_uniform(1e2, 1e5)
    d = random_unit(rng)
    lx, ly, lz = float(nx) / ny, 1.0, 1.0
    m = hex_model(nx, ny, nz, lx, ly, lz, E, nu, jitter=0.1, rng=rng, name=f"{name}_c3d8_{nx}x{ny}x{nz}")
    m.bcs = [(x0_face_ids(nx, ny, nz), np.array([1, 2, 3], np.int32))]
    tip = (nz // 2) * (ny + 1) * (nx + 1) + (ny // 2) * (nx + 1) + nx + 1
    m.cload_node = np.full(3, tip, np.uint32)
    m.cload_dof = np.array([1, 2, 3], np.int32)
    m.cload_value = P * d
    return m


@dataclass
class Request:
    arrival_s: float
    model: Model


def cfg5_stream(duration_s: float = 70.0, rate: float = 200.0, seed: int = SEED_BASE + 5,
                n_lo_solid=1e4, n_hi_solid=3e5, n_lo_plane=2e3, n_hi_plane=2e5,
                max_requests: Optional[int] = None):
    """Poisson arrivals (rate per second); 20% plane instances of the batched
    distribution, 80% C3D8 blocks. Yields Request objects lazily."""
    rng = SplitMix64(seed)
    t = 0.0
    i = 0
    while True:
        t += -math.log(1.0 - rng.u()) / rate
        if t > duration_s or (max_requests is not None and i >= max_requests):
            return
        sub = SplitMix64(seed ^ (0x9E37 + i))
        if rng.u() < 0.2:
            m = plane_instance(sub, n_lo_plane, n_hi_plane, name=f"req{i}")
        else:
            m = solid_instance(sub, n_lo_solid, n_hi_solid, name=f"req{i}")
        yield Request(t, m)
        i += 1
