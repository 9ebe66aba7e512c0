"""Pins the CPU oracle's system level: the reference's own C++ tests
(cpp/src/_test.cpp), brute-force triplet assembly, finite differences and
closed-form dynamics."""
import numpy as np
import pytest

from sors_amd import mesh

# TestEnergy mesh: cpp/src/_test.cpp:83-92
TE_X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, -1, 0]], float)
TE_CONN = np.array([[0, 1, 2, 3], [0, 1, 4, 3]], np.int32)
# TestCube mesh: cpp/src/_test.cpp:290-308
CUBE_X = np.array([[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1], [1, 0, 1]], float)
CUBE_CONN = np.array([[0, 1, 3, 4], [1, 4, 5, 6], [1, 2, 3, 6], [3, 4, 6, 7], [1, 3, 4, 6]], np.int32)


def dummy_system(orc, scheme):
    # Testahedron elements (polynomial energy), density 1, default settings
    return orc.System(TE_X, TE_CONN, np.zeros(2), np.zeros(2), np.ones(2), scheme=scheme, method=orc.NEWTON,
                      ed=orc.TEST_POLY)


def test_dummy_energy_gradient_hessian(orc):
    # cpp/src/_test.cpp:136-176; default SimulationSettings use Crank-Nicolson
    s = dummy_system(orc, orc.CN)
    q = TE_X.reshape(-1)
    assert s.energy(q, 0.0) == 3.0
    gradGT = np.array([[0, 0, 0], [6, 0, 0], [0, 3, 0], [0, 0, 6], [0, -3, 0]], float).reshape(-1)
    assert np.allclose(s.gradient(q, 0.0), gradGT)
    hessDiagGT = np.array([[-2, -2, -2], [22, -2, -2], [-1, 11, -1], [-2, -2, 22], [-1, 11, -1]], float).reshape(-1)
    assert np.allclose(s.hessian_dense(q, 0.0), np.diag(hessDiagGT))


def test_dummy_newton_optimization(orc):
    # cpp/src/_test.cpp:180-199: minimize_newton with dt = 1 reaches ||g||^2 < 1e-9
    s = dummy_system(orc, orc.CN)
    q, st = s.minimize(TE_X.reshape(-1), 1.0)
    assert (s.gradient(q, 1.0) ** 2).sum() < 1e-9


def cube_system(orc, grav, density=1.0, E=5e5, nu=0.4):
    return orc.System(CUBE_X, CUBE_CONN, np.full(5, E), np.full(5, nu), np.full(5, density), grav=grav,
                      scheme=orc.BE, method=orc.NEWTON)


def com_accel(s, q, dt, axis, steps=10, **kw):
    com = [q.reshape(-1, 3).mean(0)[axis]]
    for _ in range(1, steps):
        q, _ = s.step(q, dt, **kw)
        com.append(q.reshape(-1, 3).mean(0)[axis])
    com = np.array(com)
    acc = (-2 * com[1:-1] + com[2:] + com[:-2]) / dt**2
    return acc[1:steps - 3].mean()


def test_cube_gravity_free_fall(orc):
    # cpp/src/_test.cpp:395-443: COM acceleration -10 m/s^2 within 1e-3
    s = cube_system(orc, (0, 0, -10))
    assert com_accel(s, CUBE_X.reshape(-1).copy(), 0.01, 2) == pytest.approx(-10.0, abs=1e-3)


def test_cube_external_vertex_force(orc):
    # cpp/src/_test.cpp:507-560: -10/8 N on every DOF of the 8 nodes
    s = cube_system(orc, (0, 0, 0))
    s.set_external_force(np.full(24, -10.0 / 8))
    assert com_accel(s, CUBE_X.reshape(-1).copy(), 0.01, 2) == pytest.approx(-10.0, abs=1e-3)


def brute_force_pattern(conn, mask, n, dt, sqp):
    """Literal triplet list of energy.cpp:315-350 + integrator.cpp:94-106 +
    energy.cpp:381-388, then sorted unique (setFromTriplets)."""
    ed = conn.shape[1]
    dof = (3 * conn[:, :, None] + np.arange(3)).reshape(len(conn), 3 * ed)
    R = np.repeat(dof, 3 * ed, axis=1).reshape(-1)
    C = np.tile(dof, (1, 3 * ed)).reshape(-1)
    keep = np.ones_like(R, bool) if sqp else (mask[R] == 0) & (mask[C] == 0)
    rows, cols = list(R[keep]), list(C[keep])
    if not keep.all():  # default Triplet() = (0, 0, 0.0)
        rows.append(0)
        cols.append(0)
    if dt != 0:
        rows += list(range(n))
        cols += list(range(n))
    if not sqp:
        m = np.nonzero(mask)[0]
        rows += list(m)
        cols += list(m)
    key = np.unique(np.array(rows, np.int64) * n + np.array(cols, np.int64))
    r, c = key // n, key % n
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    return np.cumsum(rp).astype(np.int32), c.astype(np.int32)


@pytest.mark.parametrize("case", ["clamp_dt", "clamp_dt0", "partial_dt0", "sqp", "hex_clamp"])
def test_pattern_matches_brute_force_triplets(orc, case):
    element = "hex" if case == "hex_clamp" else "tet"
    sc = mesh.beam(nx=4, ny=2, nz=3, element=element)
    mask = sc.bcmask.copy()
    dt = 0.0 if case.endswith("dt0") else 1e-3
    method = orc.SQP if case == "sqp" else orc.NEWTON
    if case == "partial_dt0":
        mask[:] = 0
        mask[[0, 7, 8, 20]] = 1  # masks DOF 0: (0,0) survives only through the default triplets
    s = orc.System(sc.X, sc.conn, sc.youngs, sc.poisson, sc.density, mask, sc.bcvalue, sc.grav, method=method)
    rp, cols = s.pattern(dt)
    brp, bcols = brute_force_pattern(sc.conn, mask, 3 * sc.n_nodes, dt, method == orc.SQP)
    assert np.array_equal(rp, brp)
    assert np.array_equal(cols, bcols)
    smap = s.scatter_map(dt)
    ed = sc.ed
    for e in range(0, sc.n_elems, 7):
        dof = (3 * sc.conn[e][:, None] + np.arange(3)).reshape(-1)
        for r in range(3 * ed):
            for c in range(3 * ed):
                idx = smap[e, r, c]
                R, Cc = dof[r], dof[c]
                masked = method != orc.SQP and (mask[R] or mask[Cc])
                if masked:
                    assert idx == -1
                else:
                    assert rp[R] <= idx < rp[R + 1] and cols[idx] == Cc


def test_config_pattern_sizes(orc):
    # 20x4x4 Kuhn beam, x = 0 clamp: 52,437 entries after masking,
    # 55,053 without the mask (SQP semantics)
    sc = mesh.cantilever_small()
    s = orc.System.from_scene(sc)
    assert len(s.pattern(sc.dt)[1]) == 52437
    s2 = orc.System.from_scene(sc, method=orc.SQP)
    assert len(s2.pattern(sc.dt)[1]) == 55053


def test_energy_gradient_hessian_consistency(orc):
    # finite differences of the FP64 energy / gradient on free DOFs (BE)
    sc = mesh.beam(nx=3, ny=2, nz=2, E=2e5)
    s = orc.System.from_scene(sc, method=orc.NEWTON)
    rng = np.random.default_rng(7)
    q = sc.X.reshape(-1) + rng.normal(scale=2e-4, size=3 * sc.n_nodes) * (1 - sc.bcmask)
    dt = 1e-3
    g = s.gradient(q, dt)
    H = s.hessian_dense(q, dt)
    free = np.nonzero(sc.bcmask == 0)[0]
    h = 1e-7
    for j in free[::5]:
        e = np.zeros_like(q)
        e[j] = h
        fd = (s.energy(q + e, dt) - s.energy(q - e, dt)) / (2 * h)
        assert fd == pytest.approx(g[j], rel=1e-5, abs=1e-7 * np.abs(g).max())
        gfd = (s.gradient(q + e, dt) - s.gradient(q - e, dt)) / (2 * h)
        assert np.allclose(gfd[free], H[free, j], rtol=1e-5, atol=1e-6 * np.abs(H).max())
    # Dirichlet rows: override gradient and unit diagonal (energy.cpp:299-306, 381-388)
    m = np.nonzero(sc.bcmask)[0]
    assert np.allclose(g[m], q[m] - sc.X.reshape(-1)[m])
    assert np.allclose(np.diag(H)[m], 1.0)
    assert np.allclose(H[np.ix_(m, free)], 0)


def test_total_mass_closed_form(orc):
    sc = mesh.beam(nx=5, ny=2, nz=3, rho=1100.0)
    s = orc.System.from_scene(sc)
    assert s.mass().sum() / 3 == pytest.approx(1100.0 * 0.1 * 0.02 * 0.02, rel=1e-12)
    sh = orc.System.from_scene(mesh.beam(nx=5, ny=2, nz=3, element="hex"))
    assert sh.mass().sum() / 3 == pytest.approx(1100.0 * 0.1 * 0.02 * 0.02, rel=1e-12)


def test_ldlt_and_pcg_reach_same_state(orc):
    sc = mesh.cantilever_small()
    a = orc.System.from_scene(sc, method=orc.NEWTON_LS)
    b = orc.System.from_scene(sc, method=orc.NEWTON_LS)
    qa = sc.X.reshape(-1).copy()
    qb = qa.copy()
    for _ in range(3):
        qa, _ = a.step(qa, sc.dt)
        qb, _ = b.step(qb, sc.dt, linear_solver=orc.PCG, project=1, pcg_rtol=1e-10)
    assert np.abs(qa - qb).max() < 1e-9 * sc.bbox_diag


def test_kuhn_split_is_conforming_and_positive(orc):
    X, hexes = mesh.box_grid(3, 2, 2, 3.0, 2.0, 2.0)
    tets = mesh.kuhn_split(hexes)
    P = X[tets]
    vol = np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0])) / 6
    assert (vol > 0).all() and vol.sum() == pytest.approx(12.0)
    # conforming: every interior triangle is shared by exactly two tets
    faces = np.sort(np.concatenate([tets[:, [0, 1, 2]], tets[:, [0, 1, 3]], tets[:, [0, 2, 3]], tets[:, [1, 2, 3]]]),
                    axis=1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    assert counts.max() == 2
    n_boundary = (counts == 1).sum()
    assert n_boundary == 2 * 2 * (3 * 2 + 3 * 2 + 2 * 2)  # two triangles per boundary quad
