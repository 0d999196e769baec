"""Extract the J2 plasticity golden vectors of the reference's unit tests into JSON.

Run once in the build container (where the reference checkout is mounted):

    python tests/golden/make_j2_goldens.py /root/reference

It reads the reference test files as text and takes the brace-initialised
gold arrays (and the scalar gold of the criterion test) of six tests; the
inputs of each test are recorded beside them.  Nothing from the reference is
executed.  Entries like ``2.0*0.18231319`` are evaluated as the product.
"""
import json
import re
import sys
from pathlib import Path

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"
TERM = re.compile(rf"{NUM}(?:\s*\*\s*{NUM})*")


def test_body(text, name):
    start = text.index(f"TEUCHOS_UNIT_TEST(PlatoAnalyzeUnitTests, {name})")
    nxt = text.find("TEUCHOS_UNIT_TEST(", start + 10)
    return text[start: nxt if nxt > 0 else len(text)], text[:start].count("\n") + 1


def brace_block(body, var):
    i = body.index(var)
    j = body.index("{", i)
    depth, k = 0, j
    while True:
        if body[k] == "{":
            depth += 1
        elif body[k] == "}":
            depth -= 1
            if depth == 0:
                break
        k += 1
    return body[j:k + 1], j


def values(chunk):
    out = []
    for t in TERM.findall(chunk):
        v = 1.0
        for f in t.split("*"):
            v *= float(f)
        out.append(v)
    return out


def main(ref_root):
    root = Path(ref_root) / "unit_tests"
    plast = (root / "PlasticityTests.cpp").read_text()
    elasto = (root / "ElastoPlasticityTests.cpp").read_text()
    out = {"tests": {}}

    def gold(text, src, test, var, inputs, rows=None):
        body, l0 = test_body(text, test)
        chunk, j = brace_block(body, var)
        vals = values(chunk)
        rec = {"source": f"unit_tests/{src}:{l0 + body[:j].count(chr(10))}", "inputs": inputs, "gold": vals}
        if rows:
            rec["shape"] = rows
        out["tests"][test] = rec

    j2_two = {"youngs": 520.0, "poisson": 0.3, "hardening_isotropic": 20.0, "hardening_kinematic": 15.0,
              "initial_yield_stress": 3.0, "elastic_penalty_exponent": 3.0, "elastic_min_ersatz": 1e-9,
              "plastic_penalty_exponent": 2.5, "plastic_min_ersatz": 1e-9, "control": 0.9, "mesh_intervals": 1,
              "note": "global state: += value on node sets (x+: uy,uz; y+: ux,uz; z+: ux,uy)"}
    gold(plast, "PlasticityTests.cpp", "J2Plasticity_Evaluate3D", "tGold",
         dict(j2_two, boundary_value=0.1, tolerance=1e-5,
              local_state=[0.3, 0.5, -0.1, 0.0, 0.1, 0.2, 0.2, 0.2, 12.0, -4.0, -8.0, 1.0, 1.0, 1.0],
              prev_local_state=[0.2, 0.75, 0.0, -0.2, 0.2, 0.2, 0.2, 0.2, 2.0, 40.0, -42.0, 2.0, 2.0, 2.0]))
    gold(plast, "PlasticityTests.cpp", "J2Plasticity_UpdateLocalState3D", "tGold",
         dict(j2_two, boundary_value=0.2, tolerance=1e-5,
              prev_local_state=[0.2, 0.75, 0.0, -0.2, 0.2, 0.2, 0.2, 0.2, 2.0, 40.0, -42.0, 2.0, 2.0, 2.0]))
    gold(elasto, "ElastoPlasticityTests.cpp", "ElastoPlasticity_Residual3D_Elastic", "tGold",
         {"youngs": 1e11, "poisson": 0.35, "global_penalty_exponent": 3.0, "global_min_ersatz": 1e-9,
          "control": 1.0, "mesh_intervals": 1, "tolerance": 1e-4,
          "global_state": "node n: (1e-7 n, 2e-7 n, 3e-7 n, 4e-7 n)",
          "projected_pressure_grad_workset": "cell e, node a, dim d: 4e-7 (a+1)(d+1)(e+1)",
          "local_state": "zero"}, rows=[6, 16])
    gold(elasto, "ElastoPlasticityTests.cpp", "ElastoPlasticity_ElasticSolution3D", "tGold",
         {"youngs": 1e11, "poisson": 0.35, "hardening_isotropic": 1e9, "hardening_kinematic": 1e9,
          "initial_yield_stress": 1e9, "elastic_penalty_exponent": 3.0, "elastic_min_ersatz": 1e-6,
          "plastic_penalty_exponent": 2.5, "plastic_min_ersatz": 1e-9, "global_penalty_exponent": 3.0,
          "global_min_ersatz": 1e-9, "control": 1.0, "mesh_intervals": 1, "num_steps": 1, "max_newton": 20,
          "dirichlet": "x- nodes dofs 0,1,2 = 0; x+ nodes dof 1 = -0.001 t", "tolerance": 1e-4})
    crit = {"youngs": 1e6, "poisson": 0.3, "hardening_isotropic": 1e3, "hardening_kinematic": 1e3,
            "initial_yield_stress": 1e3, "elastic_penalty_exponent": 3.0, "elastic_min_ersatz": 1e-6,
            "plastic_penalty_exponent": 2.5, "plastic_min_ersatz": 1e-9, "global_penalty_exponent": 3.0,
            "global_min_ersatz": 1e-6, "work_penalty_exponent": 3.0, "work_min_ersatz": 1e-9, "multiplier": -1.0,
            "control": 1.0, "mesh_intervals": 2, "num_steps": 4, "max_newton": 20,
            "dirichlet": "x- dof 0, y- dof 1, y+ dof 1, z- dof 2 = 0; x+ dof 0 = 0.002 t", "tolerance": 1e-4}
    gold(elasto, "ElastoPlasticityTests.cpp", "ElastoPlasticity_CriterionTest_3D", "tGold", crit)
    body, l0 = test_body(elasto, "ElastoPlasticity_CriterionTest_3D")
    m = re.search(rf"TEST_FLOATING_EQUALITY\(tObjValue,\s*({NUM})", body)
    out["tests"]["ElastoPlasticity_CriterionTest_3D"]["gold_value"] = float(m.group(1))
    out["tests"]["ElastoPlasticity_CriterionTest_3D"]["value_source"] = (
        f"unit_tests/ElastoPlasticityTests.cpp:{l0 + body[:m.start()].count(chr(10))}")
    body, l0 = test_body(elasto, "ElastoPlasticity_InverseMatrixWorkset")
    chunk, j = brace_block(body, "tGoldMatrixInverse")
    out["tests"]["ElastoPlasticity_InverseMatrixWorkset"] = {
        "source": f"unit_tests/ElastoPlasticityTests.cpp:{l0 + body[:j].count(chr(10))}",
        "inputs": {"matrix": "cell n: s [[-2, 1.5], [1, -0.5]] with s = 1/(1+n), n = 0..2",
                   "tolerance": 1e-6},
        "gold": values(chunk), "note": "inverse of cell n is (1+n) * gold"}
    dst = Path(__file__).resolve().parent / "j2_tests.json"
    dst.write_text(json.dumps(out, indent=1))
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
