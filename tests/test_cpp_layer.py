"""CPU checks of the C++ host layer (include/platoanalyze_amd.hpp): the test
program builds against the header and the library, lists its cases and passes
its parameter-list and expression checks (none of which touch a device)."""
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
BIN = HERE / "cpp" / "reference_cases"

CASES = ["ElastostaticResidual3D", "InternalElasticEnergy3D", "ThermostaticResidual3D", "InternalThermalEnergy3D",
         "Compliance", "ElastoPlasticity_ElasticSolution3D", "ElastoPlasticity_CriterionTest_3D", "J2Worksets",
         "ExodusMesh", "AnalyzeApp", "AnalyzeAppPlasticity"]


@pytest.fixture(scope="module")
def program():
    if not BIN.exists():
        import importlib.util
        spec = importlib.util.spec_from_file_location("pa_cpp_build", HERE / "cpp" / "build.py")
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build(verbose=False)
    return BIN


def test_case_list(program):
    out = subprocess.run([str(program), "--list"], capture_output=True, text=True, timeout=60, check=True).stdout
    assert out.split() == CASES


def test_parameter_list_and_expression_checks(program):
    r = subprocess.run([str(program), "--parser"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "parser checks passed 0" in r.stdout


def test_unknown_case_is_an_error(program):
    r = subprocess.run([str(program), "NoSuchCase"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
