"""The reference's unit-test scenarios run through the C++ host layer
(include/platoanalyze_amd.hpp, program tests/cpp/reference_cases), configured
from the reference's own input decks, against the reference's golden values at
its own tolerances and against the oracle:

* DerivativeTests ElastostaticResidual3D / InternalElasticEnergy3D /
  ThermostaticResidual3D / InternalThermalEnergy3D goldens, plus the full
  Jacobians, the transposed Jacobian and the CRS graph against the oracle;
* an elliptic compliance Problem (clamped x-, traction on x+, SIMP p = 3)
  against a direct solve and the reference's composite gradient formula;
* ElastoPlasticity ElasticSolution3D and CriterionTest_3D goldens through
  PlasticityProblem with 'Time Dependent' Dirichlet functions;
* the Plato Engine operations (AnalyzeApp) on the compliance and plastic-work
  decks, against the Problem calls bit for bit;
* the J2 local and global functions at a converged plastic state against the
  oracle (1e-12).
"""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle

pytestmark = pytest.mark.gpu

HERE = Path(__file__).resolve().parent
BIN = HERE / "cpp" / "reference_cases"
DGOLD = json.loads((HERE / "golden" / "derivative_tests.json").read_text())["tests"]
JGOLD = json.loads((HERE / "golden" / "j2_tests.json").read_text())["tests"]


def rel(a, b):
    a, b = np.asarray(a, float).ravel(), np.asarray(b, float).ravel()
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def gold(test, key):
    return np.asarray(DGOLD[test]["gold"][key]["values"], float)


def reference_equal(got, ref, tol, zero_tol=1e-11):
    """TEST_FLOATING_EQUALITY per entry, |x| < zero_tol where the gold is 0."""
    got, ref = np.asarray(got, float)[:len(ref)], np.asarray(ref, float)
    nz = ref != 0
    return (np.all(np.abs(got[nz] - ref[nz]) <= tol * np.abs(ref[nz])) and np.all(np.abs(got[~nz]) < zero_tol))


@pytest.fixture(scope="module")
def results():
    assert BIN.exists(), "tests/cpp/reference_cases is not built (run build())"
    print("case: C++ host layer, all reference cases", flush=True)
    r = subprocess.run([str(BIN), "all"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    out = {}
    for line in r.stdout.splitlines():
        d = json.loads(line)
        assert "error" not in d, (d["case"], d.get("error"))
        out[d["case"]] = d["outputs"]
    return out


def model(physics):
    coords, conn = oracle.box_mesh(2, 2, 2)
    if physics == "elastic":
        return oracle.Model(coords, conn, youngs=1e6, poisson=0.3, penalty_exponent=1.0, penalty_min=0.0)
    return oracle.Model(coords, conn, conductivity=100.0, penalty_exponent=1.0, penalty_min=0.0)


def test_elastostatic_residual(results):
    o = results["ElastostaticResidual3D"]
    assert reference_equal(o["residual"], gold("ElastostaticResidual3D", "residual"), 1e-13)
    assert reference_equal(o["jacobian"], gold("ElastostaticResidual3D", "jacobian_head"), 1e-15, 1e-9)
    assert reference_equal(o["gradient_z"], gold("ElastostaticResidual3D", "gradient_z_head"), 2e-14)
    om = model("elastic")
    assert np.array_equal(o["rowmap"], om.rowmap) and np.array_equal(o["colidx"], om.colidx)
    assert o["num_rows"] == 81 and o["block"] == 3 and o["gz_block_cols"] == 3
    assert reference_equal(o["gradient_x"], gold("ElastostaticResidual3D", "gradient_x_head"), 1e-13)
    assert o["gx_block"] == 33
    u = 1e-4 * np.arange(1, 82)
    z = np.ones(27)
    assert rel(o["jacobian"], om.gradient_u("elastic", u, z)) < 1e-12
    assert rel(o["gradient_z"], om.gradient_z("elastic", u, z)) < 1e-12
    A = oracle.to_scipy(3, 3, om.rowmap, om.colidx, np.asarray(o["jacobian"]))
    AT = oracle.to_scipy(3, 3, om.rowmap, om.colidx, np.asarray(o["jacobian_T"]))
    assert abs(AT - A.T).max() <= 1e-15 * abs(A).max()


def test_internal_elastic_energy(results):
    o = results["InternalElasticEnergy3D"]
    assert abs(o["value"] - 48.15) <= 1e-13 * 48.15
    assert reference_equal(o["gradient_u"], gold("InternalElasticEnergy3D", "gradient_u"), 1e-13)
    assert reference_equal(o["gradient_z"], gold("InternalElasticEnergy3D", "gradient_z"), 1e-13)
    gx, ref = np.asarray(o["gradient_x"]), gold("InternalElasticEnergy3D", "gradient_x")
    nz = ref != 0.0  # zero entries are checked as |value| < 1e-13, as the reference does
    assert reference_equal(gx[:ref.size][nz], ref[nz], 1e-13)
    assert np.max(np.abs(gx[:ref.size][~nz])) < 1e-13


def test_thermostatic_residual_and_energy(results):
    o = results["ThermostaticResidual3D"]
    assert reference_equal(o["residual"], gold("ThermostaticResidual3D", "residual"), 1e-13)
    assert reference_equal(o["jacobian"], gold("ThermostaticResidual3D", "jacobian_head"), 1e-15, 1e-9)
    assert reference_equal(o["gradient_z"], gold("ThermostaticResidual3D", "gradient_z_head"), 1e-14)
    assert reference_equal(o["gradient_x"], gold("ThermostaticResidual3D", "gradient_x_head"), 1e-13)
    e = results["InternalThermalEnergy3D"]
    assert abs(e["value"] - float(gold("InternalThermalEnergy3D", "value"))) < 1e-11
    assert reference_equal(e["gradient_u"], gold("InternalThermalEnergy3D", "gradient_u"), 1e-13)
    assert reference_equal(e["gradient_z"], gold("InternalThermalEnergy3D", "gradient_z"), 1e-13)
    assert reference_equal(e["gradient_x"], gold("InternalThermalEnergy3D", "gradient_x"), 1e-13)


def test_compliance_problem(results):
    o = results["Compliance"]
    coords, conn = oracle.box_mesh(6, 3, 3, 2.0, 1.0, 1.0)
    fc, fl = oracle.side_set(coords, conn, 0, 2.0)
    om = oracle.Model(coords, conn, youngs=1.0, poisson=0.3, penalty_exponent=3.0, penalty_min=1e-9,
                      face_cells=fc, face_lnodes=fl, flux=(0.0, 0.0, -1.0))
    z = np.asarray(o["z"])
    dofs = np.asarray(o["bc_dofs"], np.int32)
    X = coords.reshape(-1, 3)
    left = np.nonzero(np.abs(X[:, 0]) < 1e-12)[0]
    assert np.array_equal(dofs, np.sort((3 * left[:, None] + np.arange(3)).ravel()))
    n = 3 * om.N
    vals = om.gradient_u("elastic", np.zeros(n), z)
    rhs = -om.value("elastic", np.zeros(n), z)
    oracle.apply_block_constraints(3, om.rowmap, om.colidx, vals, rhs, dofs, np.zeros(dofs.size))
    ud = spla.spsolve(oracle.to_scipy(3, 3, om.rowmap, om.colidx, vals).tocsc(), rhs)
    assert o["rel_residual"] <= 1e-12
    assert rel(o["u"], ud) < 1e-7
    assert abs(o["value"] - om.energy_value("elastic", ud, z)) <= 1e-9 * abs(o["value"])
    u = np.asarray(o["u"])
    comp = om.energy_gradient_z("elastic", u, z) + oracle.block_spmv(1, 3, om.rowmap, om.colidx,
                                                                     om.gradient_z("elastic", u, z), -u)
    assert rel(o["gradient"], comp) < 1e-12
    # self-adjoint criterionGradientX = -df/dx
    assert rel(o["gradient_x"], -om.energy_gradient_x("elastic", u, z)) < 1e-12


def test_analyze_app_operations(results):
    """The compliance deck through the Plato Engine operations (AnalyzeApp,
    the reference's MPMD_App, src/Analyze_App.cpp): the same state, value and
    gradients as the Problem calls of the Compliance case, the value minus the
    operation's Target, the coordinate gradient by component ("name@i"), and
    an operation Parameter that doubles the traction (compliance x4, linear
    elasticity)."""
    a, o = results["AnalyzeApp"], results["Compliance"]
    z = np.asarray(o["z"])
    assert a["num_ids"] == z.size
    assert np.array_equal(a["topology"], z)
    u = np.asarray(o["u"]).reshape(-1, 3)
    for d, name in enumerate(("displacement X", "displacement Y", "displacement Z")):
        assert np.array_equal(a[name], u[:, d])
    assert a["objective_value"] == o["value"]
    assert np.array_equal(a["objective_gradient"], o["gradient"])
    gx = np.asarray(o["gradient_x"]).reshape(-1, 3)
    for d in range(3):
        assert np.array_equal(a[f"gradient_x{d}"], gx[:, d])
    assert a["constraint_value"] == o["value"] - 2.5
    assert a["objective_value_after_reinit"] == o["value"]
    assert abs(a["scaled_value"] - 4.0 * o["value"]) <= 1e-12 * abs(4.0 * o["value"])
    for k in ("unknown_operation_throws", "undefined_parameter_throws", "write_output_throws",
              "reassigned_name_throws"):
        assert a[k] == 1.0, k


def test_analyze_app_plasticity(results):
    """The reference's plastic-work criterion deck through the Plato Engine
    operations (AnalyzeApp, Physics 'Plasticity'): ComputeCriterion gives the
    value and adjoint gradient of the PlasticityProblem calls bit for bit (and
    so the reference's golden gradient), the solution exports by the mixed
    form's dof names, and the coordinate gradient raises an error."""
    a, o = results["AnalyzeAppPlasticity"], results["ElastoPlasticity_CriterionTest_3D"]
    assert a["objective_value"] == o["value"]
    assert np.array_equal(a["objective_gradient"], o["gradient"])
    state = np.asarray(o["final_state"]).reshape(-1, 4)
    for d, name in enumerate(("displacement X", "displacement Y", "displacement Z", "pressure")):
        assert np.array_equal(a[name], state[:, d])
    assert a["gradient_x_throws"] == 1.0


def test_j2_elastic_solution_golden(results):
    g = JGOLD["ElastoPlasticity_ElasticSolution3D"]
    o = results["ElastoPlasticity_ElasticSolution3D"]
    assert o["converged"] == 1.0
    u, ref = np.asarray(o["state"]), np.asarray(g["gold"])
    nz = ref != 0
    assert np.all(np.abs(u[nz] - ref[nz]) <= g["inputs"]["tolerance"] * np.abs(ref[nz]))


def test_j2_criterion_golden(results):
    g = JGOLD["ElastoPlasticity_CriterionTest_3D"]
    o = results["ElastoPlasticity_CriterionTest_3D"]
    tol = g["inputs"]["tolerance"]
    assert abs(o["value"] - g["gold_value"]) <= tol * abs(g["gold_value"])
    ref = np.asarray(g["gold"])
    assert np.all(np.abs(np.asarray(o["gradient"]) - ref) <= tol * np.abs(ref))


def test_j2_functions_against_oracle(results):
    o = results["J2Worksets"]
    inp = JGOLD["ElastoPlasticity_CriterionTest_3D"]["inputs"]
    coords, conn = oracle.box_mesh(2, 2, 2)
    keys = ["youngs", "poisson", "hardening_isotropic", "hardening_kinematic", "initial_yield_stress",
            "elastic_penalty_exponent", "elastic_min_ersatz", "plastic_penalty_exponent", "plastic_min_ersatz",
            "global_penalty_exponent", "global_min_ersatz"]
    om = oracle.J2Model(coords, conn, **{k: inp[k] for k in keys})
    u, c, cp, pi = (np.asarray(o[k]) for k in ("u", "c", "cp", "pi"))
    z = np.ones(om.N)
    assert (c.reshape(-1, 14)[:, 1] > 0).sum() > 0  # a plastic state
    assert rel(o["c_updated"], om.update_local_state(u, cp, z)) < 1e-12
    H = np.asarray(o["local_value"])
    assert np.max(np.abs(H - om.local_value(u, c, cp, z))) <= 1e-12 * max(1.0, np.max(np.abs(c)))
    assert rel(o["local_gradient_c"], om.local_gradient("c", u, c, cp, z)) < 1e-12
    assert rel(o["local_gradient_u"], om.local_gradient("u", u, c, cp, z)) < 1e-12
    assert rel(o["global_value"], om.global_value(u, c, pi, z)) < 1e-12
    pi_ws = oracle.gather_nodal(om.conn, pi, 3)
    assert rel(o["global_gradient_u"], om.global_gradient_ws("u", u, c, pi_ws, z)) < 1e-12
    assert rel(o["global_gradient_z"], om.global_gradient_ws("z", u, c, pi_ws, z)) < 1e-12
    vals, R = om.condensed_system(u, c, cp, pi, z)
    assert rel(o["tangent"], vals) < 1e-12
    assert rel(o["condensed_residual"], R) < 1e-12
