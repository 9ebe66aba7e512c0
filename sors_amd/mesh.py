"""Seeded structured meshes for the beam / cube / prism workloads.

The reference builds structured meshes with ``IO::make_voxel_hex_grid``
(cpp/src/io.cpp:931-987) and splits hexes with ``IO::hex_to_5_tets``
(io.cpp:899-928).  That 5-tet split uses one pattern in every cell, so
neighbouring cells pick different face diagonals and the tet mesh is not
conforming.  This module keeps the reference's local hex corner order
(io.cpp:963-976: 0=(0,0,0) 1=(1,0,0) 2=(1,1,0) 3=(0,1,0), top face +z) but
splits each cell into the 6 Kuhn tets that share the cell's main diagonal;
every face is then cut along the diagonal from its lowest to its highest
corner, which is the same for both cells sharing it.

Node ids run with x slowest: ``id = k + NZ * (j + NY * i)``, which keeps the
matrix profile narrow (the CPU skyline LDL^T depends on it).

Random material fields use SplitMix64 (Steele, Lea, Flood 2014) so that the
C++ and Python sides can reproduce them bit for bit.
"""
from __future__ import annotations

import dataclasses
import functools
import math

import numpy as np

_M64 = (1 << 64) - 1

# Kuhn split of the unit cell in the reference's local corner order: one tet
# per permutation of the axes, walking 000 -> ... -> 111.  Odd permutations
# swap their middle corners so every tet has positive orientation.
KUHN_TETS = np.array(
    [
        [0, 1, 2, 6],  # x y z
        [0, 5, 1, 6],  # x z y (odd)
        [0, 2, 3, 6],  # y x z (odd)
        [0, 3, 7, 6],  # y z x
        [0, 4, 5, 6],  # z x y
        [0, 7, 4, 6],  # z y x (odd)
    ],
    dtype=np.int32,
)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` successive SplitMix64 outputs as uint64."""
    out = np.empty(count, dtype=np.uint64)
    x = seed & _M64
    for i in range(count):
        x = (x + 0x9E3779B97F4A7C15) & _M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        out[i] = z ^ (z >> 31)
    return out


def splitmix64_np(seed: int, count: int) -> np.ndarray:
    """Vectorised SplitMix64 (identical to :func:`splitmix64`)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed & _M64) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, count: int) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64_np(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def log_uniform(seed: int, count: int, lo: float, hi: float) -> np.ndarray:
    u = uniform01(seed, count)
    return np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))


def box_grid(nx: int, ny: int, nz: int, lx: float, ly: float, lz: float):
    """Vertices (N x 3) and hex connectivity (E x 8) of an nx*ny*nz box grid."""
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    i, j, k = np.meshgrid(np.arange(NX), np.arange(NY), np.arange(NZ), indexing="ij")
    X = np.stack([i * (lx / nx), j * (ly / ny), k * (lz / nz)], axis=-1).reshape(-1, 3).astype(np.float64)

    def nid(a, b, c):
        return c + NZ * (b + NY * a)

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    hexes = np.stack(
        [
            nid(ci, cj, ck),
            nid(ci + 1, cj, ck),
            nid(ci + 1, cj + 1, ck),
            nid(ci, cj + 1, ck),
            nid(ci, cj, ck + 1),
            nid(ci + 1, cj, ck + 1),
            nid(ci + 1, cj + 1, ck + 1),
            nid(ci, cj + 1, ck + 1),
        ],
        axis=1,
    ).astype(np.int32)
    return X, hexes


def kuhn_split(hexes: np.ndarray) -> np.ndarray:
    """6 conforming tets per hex (E*6 x 4), tets of one cell are contiguous."""
    return hexes[:, KUHN_TETS].reshape(-1, 4).astype(np.int32)


def hex_to_5_tets(hexes: np.ndarray) -> np.ndarray:
    """The reference's 5-tet split (io.cpp:899-928), for callers that want it."""
    pat = np.array([[0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]], dtype=np.int32)
    return hexes[:, pat].reshape(-1, 4).astype(np.int32)


@dataclasses.dataclass
class Scene:
    """Everything one rollout needs, as plain arrays (node-major FP64)."""

    name: str
    X: np.ndarray          # N x 3 rest positions
    conn: np.ndarray       # E x ed
    youngs: np.ndarray     # E
    poisson: np.ndarray    # E
    density: np.ndarray    # E
    bcmask: np.ndarray     # 3N int32 (1 = Dirichlet DOF)
    bcvalue: np.ndarray    # 3N prescribed velocity
    v0: np.ndarray         # 3N initial velocity
    grav: np.ndarray       # 3
    dt: float
    steps: int
    alpha: float = 0.0
    tip: np.ndarray | None = None  # node ids of the tip face

    @property
    def ed(self) -> int:
        return int(self.conn.shape[1])

    @property
    def n_nodes(self) -> int:
        return int(self.X.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.conn.shape[0])

    @property
    def bbox_diag(self) -> float:
        return float(np.linalg.norm(self.X.max(0) - self.X.min(0)))


def beam(nx=20, ny=4, nz=4, lx=0.10, ly=0.02, lz=0.02, element="tet", E=1e6, nu=0.45, rho=1100.0,
         e_range=None, seed=0, tip_speed=0.5, dt=1e-3, steps=200, name="beam") -> Scene:
    """Cantilever beam clamped at x = 0, gravity -z, initial velocity
    v_z = tip_speed * x / L (zero at the clamp, tip_speed at the tip).

    ``e_range=(lo, hi)`` draws Young's modulus per hex cell, log-uniform from
    SplitMix64(seed); the 6 tets of a cell inherit the cell's value so the
    hex and tet meshes carry the same material field.
    """
    X, hexes = box_grid(nx, ny, nz, lx, ly, lz)
    ncell = hexes.shape[0]
    if e_range is None:
        Ecell = np.full(ncell, float(E))
    else:
        Ecell = log_uniform(seed, ncell, float(e_range[0]), float(e_range[1]))
    if element == "tet":
        conn = kuhn_split(hexes)
        Eel = np.repeat(Ecell, 6)
    elif element == "hex":
        conn = hexes
        Eel = Ecell
    else:
        raise ValueError(f"unknown element kind {element!r}")
    nel = conn.shape[0]
    N = X.shape[0]
    clamp = X[:, 0] < X[:, 0].min() + 1e-4 * lx
    bcmask = np.repeat(clamp.astype(np.int32), 3)
    v0 = np.zeros((N, 3))
    v0[:, 2] = tip_speed * X[:, 0] / lx
    tip = np.nonzero(X[:, 0] > X[:, 0].max() - 1e-4 * lx)[0].astype(np.int32)
    return Scene(
        name=name,
        X=X,
        conn=conn.astype(np.int32),
        youngs=Eel.astype(np.float64),
        poisson=np.full(nel, float(nu)),
        density=np.full(nel, float(rho)),
        bcmask=bcmask,
        bcvalue=np.zeros(3 * N),
        v0=v0.reshape(-1),
        grav=np.array([0.0, 0.0, -9.81]),
        dt=dt,
        steps=steps,
        tip=tip,
    )


def cantilever_small(**kw) -> Scene:
    """Plain FP64 cantilever: 20x4x4 cells, 1,920 tets, 525 nodes, E=1e6,
    nu=0.45, rho=1100, 200 steps of dt=1e-3."""
    return beam(**kw)


def cantilever_large(element="tet", seed=0, steps=20, **kw) -> Scene:
    """Scaled beam: 200x40x40 cells (320,000 hex8 or 1,920,000 tet4, 337,881
    nodes), E log-uniform in [5e5, 5e6] per cell, 20 steps of dt=1e-3."""
    return beam(nx=200, ny=40, nz=40, element=element, e_range=(5e5, 5e6), seed=seed, steps=steps,
                name=f"beam200x40x40_{element}", **kw)


def serving_request(index: int, seed: int = 0, n_lo: int = 2_000, n_hi: int = 200_000, e_lo: float = 1e5,
                    e_hi: float = 1e7, nu_lo: float = 0.30, nu_hi: float = 0.48, steps_lo: int = 10,
                    steps_hi: int = 50, dt: float = 1e-3) -> Scene:
    """Request `index` of the seeded serving stream: a beam (0.10 x 0.02 x
    0.02 m, clamped at x = 0) or a cube (0.05 m, clamped at z = 0) tet mesh
    whose node count is log-uniform in [n_lo, n_hi], one Young's modulus
    log-uniform in [e_lo, e_hi], Poisson ratio uniform in [nu_lo, nu_hi],
    steps uniform in [steps_lo, steps_hi], gravity -z.  The beam starts with
    v_z = 0.5 x / L, the cube with v_x = 0.5 z / L."""
    u = uniform01((seed + 1) * 0x9E3779B97F4A7C15 + index, 5)
    kind = "beam" if u[0] < 0.5 else "cube"
    nodes = math.exp(math.log(n_lo) + u[1] * (math.log(n_hi) - math.log(n_lo)))
    E = math.exp(math.log(e_lo) + u[2] * (math.log(e_hi) - math.log(e_lo)))
    nu = nu_lo + u[3] * (nu_hi - nu_lo)
    steps = steps_lo + min(int(u[4] * (steps_hi - steps_lo + 1)), steps_hi - steps_lo)
    if kind == "beam":
        # (5k + 1)(k + 1)^2 nodes
        k = max(1, int(round((nodes / 5.0) ** (1.0 / 3.0))))
        g = _request_geometry("beam", k)
    else:
        k = max(1, int(round(nodes ** (1.0 / 3.0))) - 1)
        g = _request_geometry("cube", k)
    X, conn, bcmask, v0 = g
    nel = conn.shape[0]
    return Scene(name=f"req{index}_{kind}", X=X, conn=conn, youngs=np.full(nel, E), poisson=np.full(nel, nu),
                 density=np.full(nel, 1100.0), bcmask=bcmask, bcvalue=np.zeros(X.size), v0=v0,
                 grav=np.array([0.0, 0.0, -9.81]), dt=dt, steps=steps)


@functools.lru_cache(maxsize=None)
def _request_geometry(kind: str, k: int):
    """Mesh arrays of a serving request, shared (read-only) by every request
    of the same kind and size: the beam of beam() with 5k x k x k cells, or
    the 0.05 m cube with k^3 cells clamped at z = 0, v_x = 0.5 z / L."""
    if kind == "beam":
        sc = beam(nx=5 * k, ny=k, nz=k)
        X, conn, bcmask, v0 = sc.X, sc.conn, sc.bcmask, sc.v0
    else:
        L = 0.05
        X, hexes = box_grid(k, k, k, L, L, L)
        conn = kuhn_split(hexes).astype(np.int32)
        clamp = X[:, 2] < 1e-4 * L
        v = np.zeros_like(X)
        v[:, 0] = 0.5 * X[:, 2] / L
        bcmask, v0 = np.repeat(clamp.astype(np.int32), 3), v.reshape(-1)
    for a in (X, conn, bcmask, v0):
        a.flags.writeable = False
    return X, conn, bcmask, v0


def serving_arrivals(count: int, rate: float = 200.0, seed: int = 0) -> np.ndarray:
    """Poisson arrival times (s) of `count` requests at `rate` per second."""
    u = uniform01(seed * 0x2545F4914F6CDD1D + 7, count)
    return np.cumsum(-np.log1p(-u) / rate)


def box_surface_nodes(X: np.ndarray, tol: float = 1e-9) -> np.ndarray:
    """Nodes on the boundary of an axis-aligned box grid (the surface vertex
    set the reference extracts for contact, io.cpp:570-681)."""
    lo, hi = X.min(0), X.max(0)
    on = np.any((np.abs(X - lo) < tol) | (np.abs(X - hi) < tol), axis=1)
    return np.nonzero(on)[0].astype(np.int32)


@dataclasses.dataclass
class PokeScene:
    """Contact poke: a cube resting on the ground plane z = 0, indented from
    above by a rigid disk (normal -z) moving down with a seeded lateral
    jitter.  disk_center(t) gives the disk centre at the end of step t."""

    scene: Scene
    surface: np.ndarray
    stiffness: float
    friction: float
    eps_v: float
    radius: float
    centers: np.ndarray   # [steps + 1][3]: initial, then after each step


def contact_poke(n: int = 40, side: float = 0.05, E: float = 2e5, nu: float = 0.4, rho: float = 1100.0,
                 stiffness: float = 1e5, friction: float = 0.5, radius: float = 0.01, speed: float = 0.02,
                 jitter: float = 1e-3, jitter_every: int = 50, gap: float = 5e-4, steps: int = 500,
                 dt: float = 5e-4, seed: int = 0, eps_v: float = 1e-3) -> PokeScene:
    """n^3-cell Kuhn tet cube (6 n^3 tets).  The disk starts `gap` above the
    top face's centre and moves -z at `speed`; its lateral offset is a seeded
    Gaussian (std `jitter`) drawn every `jitter_every` steps and linearly
    interpolated in between."""
    X, hexes = box_grid(n, n, n, side, side, side)
    conn = kuhn_split(hexes).astype(np.int32)
    N, nel = X.shape[0], conn.shape[0]
    sc = Scene(name=f"poke{n}", X=X, conn=conn, youngs=np.full(nel, E), poisson=np.full(nel, nu),
               density=np.full(nel, rho), bcmask=np.zeros(3 * N, np.int32), bcvalue=np.zeros(3 * N),
               v0=np.zeros(3 * N), grav=np.array([0.0, 0.0, -9.81]), dt=dt, steps=steps)
    knots = steps // jitter_every + 2
    u = uniform01(seed * 0x632BE59BD9B4E019 + 11, 4 * knots).reshape(knots, 4)
    # Box-Muller from pairs of uniforms
    r = np.sqrt(-2.0 * np.log1p(-u[:, 0]))
    g = np.stack([r * np.cos(2 * np.pi * u[:, 1]), r * np.sin(2 * np.pi * u[:, 1])], axis=1) * jitter
    g[0] = 0.0
    t = np.arange(steps + 1) / jitter_every
    k0 = np.minimum(np.floor(t).astype(int), knots - 2)
    f = (t - k0)[:, None]
    lateral = (1 - f) * g[k0] + f * g[k0 + 1]
    c = np.zeros((steps + 1, 3))
    c[:, 0] = 0.5 * side + lateral[:, 0]
    c[:, 1] = 0.5 * side + lateral[:, 1]
    c[:, 2] = side + gap - speed * dt * np.arange(steps + 1)
    return PokeScene(scene=sc, surface=box_surface_nodes(X), stiffness=stiffness, friction=friction, eps_v=eps_v,
                     radius=radius, centers=c)


@dataclasses.dataclass
class LegScene:
    """Fibre-actuated leg: a prism standing on z, its top face clamped, its
    foot just above the ground plane z = 0 (frictionless penalty contact),
    with per-element muscle groups, one seeded fibre direction, a seeded
    activation schedule act[t, g] and a seeded target for the foot centroid."""

    scene: Scene
    group: np.ndarray      # [E] muscle group
    fibre: np.ndarray      # [E, 3]
    sigma_max: float
    act: np.ndarray        # [steps, n_groups]
    foot: np.ndarray       # node ids of the foot face
    target: np.ndarray     # [3]
    surface: np.ndarray    # contact candidates
    stiffness: float


def leg_prism(nx: int = 14, ny: int = 14, nz: int = 100, lx: float = 0.03, ly: float = 0.03, lz: float = 0.2,
              E_lo: float = 3e5, E_hi: float = 1e6, nu: float = 0.4, rho: float = 1100.0, sigma_max: float = 2e5,
              seg: int = 8, steps: int = 100, dt: float = 2e-3, gap: float = 1e-3, stiffness: float = 1e5,
              act_max: float = 0.5, seed: int = 0) -> LegScene:
    """nx*ny*nz Kuhn-split cells (6 tets each; 117,600 tets and 22,725 nodes at
    14 x 14 x 100).  Muscle groups: `seg` segments along z times 4 quadrants in
    xy.  E per element log-uniform in [E_lo, E_hi]."""
    X, hexes = box_grid(nx, ny, nz, lx, ly, lz)
    X = X.copy()
    X[:, 2] += gap
    conn = kuhn_split(hexes).astype(np.int32)
    N, nel = X.shape[0], conn.shape[0]
    Eel = log_uniform(seed * 0x5851F42D4C957F2D + 3, nel, E_lo, E_hi)
    top = X[:, 2] > X[:, 2].max() - 1e-9
    c = X[conn].mean(axis=1)
    zs = np.minimum(((c[:, 2] - gap) / lz * seg).astype(int), seg - 1)
    quad = (c[:, 0] > 0.5 * lx).astype(int) + 2 * (c[:, 1] > 0.5 * ly).astype(int)
    group = (4 * zs + quad).astype(np.int32)
    ng = 4 * seg
    u = uniform01(seed * 0x2545F4914F6CDD1D + 5, 2 + 2 * ng + 3)
    f = np.array([0.2 * (u[0] - 0.5), 0.2 * (u[1] - 0.5), 1.0])
    f /= np.linalg.norm(f)
    freq = 0.5 + u[2:2 + ng]
    phase = 2 * np.pi * u[2 + ng:2 + 2 * ng]
    t = np.arange(steps)[:, None] / max(steps, 1)
    act = act_max * 0.5 * (1.0 + np.sin(2 * np.pi * freq[None, :] * t + phase[None, :]))
    foot = np.nonzero(X[:, 2] < gap + 1e-9)[0].astype(np.int32)
    foot_c = X[foot].mean(axis=0)
    off = u[-3:] - 0.5
    target = foot_c + np.array([0.02 * off[0], 0.02 * off[1], 0.005 + 0.01 * abs(off[2])])
    sc = Scene(name=f"leg{nx}x{ny}x{nz}", X=X, conn=conn, youngs=Eel, poisson=np.full(nel, nu),
               density=np.full(nel, rho), bcmask=np.repeat(top.astype(np.int32), 3), bcvalue=np.zeros(3 * N),
               v0=np.zeros(3 * N), grav=np.array([0.0, 0.0, -9.81]), dt=dt, steps=steps)
    return LegScene(scene=sc, group=group, fibre=np.tile(f, (nel, 1)), sigma_max=sigma_max, act=act, foot=foot,
                    target=target, surface=box_surface_nodes(X), stiffness=stiffness)
