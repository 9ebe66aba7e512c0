"""The C++ host interfaces over the device (tests/cpp/test_facade.cpp): the
reference's Energy<3,E>::compute_energy / compute_gradient / compute_hessian
/ update_state, Solver<3,E>::minimize_* / step / calculate_CFL_timestep, the
ElementEnergy plugins and the Element classes (cpp/headers/energy.h:73-205,
solver.h:27-95, energies/elementEnergy.h:16-114, geometries/element.h),
checked against the FP64 oracle on the inputs the program prints.

Scenario "tet": tet4 Neo-Hookean, backward Euler, Newton + line search.
Scenario "hex": hex8 Stable Neo-Hookean with SoftCon muscles on half the
elements (two groups, activations 0.3 / 0.7), Crank-Nicolson, plain Newton.
"""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(__file__), "cpp", "test_facade")


@pytest.fixture(scope="module")
def out():
    assert os.path.exists(BIN), "run python -m sors_amd.build"
    print("running", BIN, flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def relmax(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def oracle_system(orc, s):
    ed = 4 if s["meshType"] == "Tetrahedron" else 8
    X = np.array(s["X"]).reshape(-1, 3)
    conn = np.array(s["conn"], np.int32).reshape(-1, ed)
    scheme = orc.BE if s["scheme"] == "backward_euler" else orc.CN
    method = orc.NEWTON_LS if s["method"] == "minimize_newton_linesearch" else orc.NEWTON
    o = orc.System(X, conn, s["youngs"], s["poisson"], s["density"], bcmask=np.array(s["mask"], np.int32),
                   bcvalue=np.zeros(X.size), grav=(0.0, 0.0, -9.81), alpha=0.0, scheme=scheme, method=method)
    if ed == 8:
        o.set_material(orc.MAT_SNH)
        o.set_muscle(np.array(s["muscleGroup"], np.int32), np.array(s["muscleK"]),
                     np.array(s["muscleDir"]).reshape(-1, 3))
        o.set_muscle_activation(np.array(s["activation"]))
    o.set_initial_velocity(np.array(s["v0"]))
    return o, X, conn


@pytest.mark.parametrize("name", ["tet", "hex"])
def test_energy_gradient_hessian_match_oracle(orc, out, name):
    s = out[name]
    o, X, conn = oracle_system(orc, s)
    q, dt = np.array(s["q"]), s["dt"]
    e = o.energy(q, dt)
    g = o.gradient(q, dt)
    rp, cols, vals = o.hessian(q, dt, project=False)
    H = s["hessian"]
    print(name, "energy", s["energy"], e, "grad rel", relmax(s["gradient"], g), "nnz", len(vals), flush=True)
    assert abs(s["energy"] - e) <= 1e-12 * abs(e)
    assert relmax(s["gradient"], g) <= 1e-11
    assert np.array_equal(H["outer"], rp) and np.array_equal(H["inner"], cols)
    assert relmax(H["values"], vals) <= 1e-11


@pytest.mark.parametrize("name", ["tet", "hex"])
def test_elements_and_plugins_match_oracle(orc, out, name):
    s = out[name]
    ed = 4 if name == "tet" else 8
    mat = orc.MAT_NH if name == "tet" else orc.MAT_SNH
    X = np.array(s["X"]).reshape(-1, 3)
    conn = np.array(s["conn"]).reshape(-1, ed)
    q = np.array(s["q"]).reshape(-1, 3)
    for el in s["elements"]:
        e = el["e"]
        Xe, xe = X[conn[e]], q[conn[e]]
        g = s["muscleGroup"][e]
        muscle = None
        if g >= 0:
            muscle = (np.array(s["muscleDir"][3 * e:3 * e + 3]), s["muscleK"][e], s["activation"][g])
        E, nu = s["youngs"][e], s["poisson"][e]
        oe, og, oh = orc.elem_eval_ex(Xe, xe, E, nu, material=mat, muscle=muscle)
        print(name, "element", e, "energy", el["energy"], oe, "hess rel", relmax(el["hessian"], oh), flush=True)
        assert abs(el["energy"] - oe) <= 1e-12 * abs(oe)
        assert relmax(el["gradient"], og) <= 1e-11
        assert relmax(el["hessian"], oh) <= 1e-11
        # geometry: volume (|det|/6 per tet of the element), mass, shortest edge
        if ed == 4:
            V0 = abs(np.linalg.det((Xe[:3] - Xe[3]).T)) / 6.0
            assert abs(el["volume"] - V0) <= 1e-14 * V0
        assert abs(el["mass"] - s["density"][e] * el["volume"]) <= 1e-14 * el["mass"]
        edges = [np.linalg.norm(xe[i] - xe[j]) for i in range(ed) for j in range(i + 1, ed)]
        assert abs(el["min_edge"] - min(edges)) <= 1e-15
        # each plugin alone (per unit undeformed volume, as the reference's plugins return it)
        V0 = el["volume"]
        for pname, pv in el["plugins"].items():
            if pname == "softconMuscle":
                pe, pg, _ = orc.elem_eval_ex(Xe, xe, 0.0, 0.3, material=orc.MAT_NH, muscle=muscle)
            else:
                pe, pg, _ = orc.elem_eval_ex(Xe, xe, E, nu, material=mat)
            assert abs(pv["energy"] * V0 - pe) <= 1e-12 * abs(pe) + 1e-18
            assert abs(pv["specific"] - pe) <= 1e-12 * abs(pe) + 1e-18
            assert relmax(np.array(pv["gradient"]) * V0, pg) <= 1e-11
        assert set(el["plugins"]) == ({"neohookean"} if name == "tet" else
                                      {"stableneohookean", "softconMuscle"} if g >= 0 else {"stableneohookean"})


@pytest.mark.parametrize("name", ["tet", "hex"])
def test_minimize_step_and_update_state(orc, out, name):
    s = out[name]
    o, X, conn = oracle_system(orc, s)
    dt = s["dt"]
    bbox = np.linalg.norm(X.max(axis=0) - X.min(axis=0))
    guess = np.array(s["guess"])
    qo, _ = o.minimize(guess, dt)
    for key in ("minimize_newton", "minimize_newton_linesearch"):
        err = np.abs(np.array(s[key]) - qo).max() / bbox
        print(name, key, "vs oracle", err, flush=True)
        assert err <= 1e-8
    # Solver::step == minimize + Energy::update_state, bit for bit, two steps
    assert s["step_equals_minimize_update"]
    o2, _, _ = oracle_system(orc, s)
    qs = X.reshape(-1).copy()
    for _ in range(2):
        qs, _ = o2.step(qs, dt)
    _, vprev = o2.state()
    print(name, "two steps vs oracle", np.abs(np.array(s["step2"]) - qs).max() / bbox, flush=True)
    assert np.abs(np.array(s["step2"]) - qs).max() <= 1e-8 * bbox
    assert np.abs(np.array(s["vprev2"]) - vprev).max() <= 1e-8 * bbox / dt
    # calculate_CFL_timestep at the rest positions with v_prev = v0
    ref = orc.cfl_timestep(X.reshape(-1), conn, np.array(s["v0"]), dt, 1.0)
    assert abs(s["cfl"] - ref) <= 1e-14 * ref


@pytest.mark.parametrize("name", ["tet", "hex"])
def test_plugin_registry(out, name):
    s = out[name]
    assert s["registry"].split() == ["neohookean", "softconMuscle", "stableneohookean"]
    assert "Unknown energy type: pseudostrainX" in s["unknown_energy_error"]
