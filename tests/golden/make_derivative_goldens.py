"""Extract the golden vectors of unit_tests/DerivativeTests.cpp into JSON.

Run once in the build container (where the reference checkout is mounted):

    python tests/golden/make_derivative_goldens.py /root/reference

It reads the reference test file as text, takes the brace-initialised gold
arrays of five tests and writes them with their inputs to
tests/golden/derivative_tests.json.  Nothing from the reference is executed.
The values are data (expected outputs); the mesh, material and state inputs
are recorded next to them.
"""
import json
import re
import sys
from pathlib import Path

FLOAT = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def test_body(text, name):
    start = text.index(f"TEUCHOS_UNIT_TEST( DerivativeTests, {name} )")
    nxt = text.find("TEUCHOS_UNIT_TEST(", start + 10)
    return text[start: nxt if nxt > 0 else len(text)], text[:start].count("\n") + 1


def gold_array(body, first_line, var):
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
    chunk = body[j:k + 1]
    line = first_line + body[:j].count("\n")
    return [float(x) for x in FLOAT.findall(chunk)], line


def scalar(body, first_line, pattern):
    m = re.search(pattern, body)
    return float(m.group(1)), first_line + body[:m.start()].count("\n")


def main(ref_root):
    src = Path(ref_root) / "unit_tests" / "DerivativeTests.cpp"
    text = src.read_text()
    out = {"source": "unit_tests/DerivativeTests.cpp",
           "mesh": {"box": [1.0, 1.0, 1.0], "intervals": 2, "element": "TET4"},
           "tests": {}}

    def add(test, entries, inputs):
        body, l0 = test_body(text, test)
        rec = {"inputs": inputs, "gold": {}}
        for key, var in entries.items():
            if var.startswith("re:"):
                v, line = scalar(body, l0, var[3:])
            else:
                v, line = gold_array(body, l0, var)
            rec["gold"][key] = {"values": v, "line": line}
        out["tests"][test] = rec

    add("3D", {"gradient": "gradient_gold", "strain": "strain_gold", "stress": "stress_gold",
               "result": "result_gold"},
        {"youngs": 1.0e6, "poisson": 0.3, "state": "u_k = 1e-4*(k+1)"})
    add("ElastostaticResidual3D", {"residual": "residual_gold", "jacobian_head": "gold_jac_entries",
                                   "gradient_z_head": "gold_grad_entries",
                                   "gradient_x_head": "gold_grad_x_entries"},
        {"youngs": 1.0e6, "poisson": 0.3, "penalty_exponent": 1.0, "penalty_min": 0.0,
         "state": "u_k = 1e-4*(k+1)", "control": 1.0})
    add("InternalElasticEnergy3D", {"value": r"re:value_gold = ([-+0-9.eE]+);",
                                    "gradient_u": "grad_u_gold", "gradient_z": "grad_z_gold",
                                    "gradient_x": "grad_x_gold"},
        {"youngs": 1.0e6, "poisson": 0.3, "penalty_exponent": 1.0, "penalty_min": 0.0,
         "state": "u_k = 1e-4*(k+1)", "control": 1.0})
    add("ThermostaticResidual3D", {"residual": "residual_gold", "jacobian_head": "gold_jac_entries",
                                   "gradient_z_head": "gold_grad_entries",
                                   "gradient_x_head": "gold_grad_x_entries"},
        {"conductivity": 100.0, "penalty_exponent": 1.0, "penalty_min": 0.0,
         "state": "T_k = 0.1*(k+1)", "control": 1.0})
    add("InternalThermalEnergy3D", {"value": r"re:value_gold = ([-+0-9.eE]+);",
                                    "gradient_u": "grad_u_gold", "gradient_z": "grad_z_gold",
                                    "gradient_x": "grad_x_gold"},
        {"conductivity": 100.0, "penalty_exponent": 1.0, "penalty_min": 0.0,
         "state": "T_k = 0.1*(k+1)", "control": 1.0})
    dst = Path(__file__).with_name("derivative_tests.json")
    dst.write_text(json.dumps(out, indent=1))
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
