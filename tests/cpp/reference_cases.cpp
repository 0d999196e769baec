// The reference's unit-test scenarios for this path, run through the C++
// host layer (include/platoanalyze_amd.hpp) with the reference's own input
// decks: DerivativeTests ElastostaticResidual3D, InternalElasticEnergy3D,
// ThermostaticResidual3D, InternalThermalEnergy3D
// (unit_tests/DerivativeTests.cpp:250-625, 1560-1882), ElastoPlasticity
// ElasticSolution3D and CriterionTest_3D
// (unit_tests/ElastoPlasticityTests.cpp:2104-2217, 3547-3700), plus an
// elliptic compliance Problem and the J2 local/global worksets.
//
// Usage: reference_cases <case>|all   prints one JSON object per case
//        (ExodusMesh reads the mesh file named by PA_EXODUS_MESH and is not
//        part of all)
//        reference_cases --list       case names (no device use)
//        reference_cases --parser     checks of the parameter-list reader and
//                                     the expression evaluator (no device use)
//        reference_cases --exodus <f> the host-side Exodus II reader on one file
//                                     (no device use): counts, or status 3 and
//                                     the reader's error
// The outputs are compared with the reference's golden values and the oracle
// by the Python tests that run this program.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/platoanalyze_amd.hpp"

namespace pa = Plato::amd;

namespace {

struct Json {
    std::ostringstream s;
    bool first = true;
    Json() {
        s.precision(17);
        s << "{";
    }
    void key(const std::string& k) {
        s << (first ? "" : ", ") << "\"" << k << "\": ";
        first = false;
    }
    void put(const std::string& k, double v) {
        key(k);
        s << v;
    }
    void put(const std::string& k, const std::string& v) {
        key(k);
        s << "\"" << v << "\"";
    }
    template <class T>
    void put(const std::string& k, const std::vector<T>& v, size_t n = (size_t)-1) {
        key(k);
        s << "[";
        const size_t m = std::min(n, v.size());
        for (size_t i = 0; i < m; ++i) s << (i ? ", " : "") << v[i];
        s << "]";
    }
    std::string str() { return s.str() + "}"; }
};

const char* kElasticDeck =
    "<ParameterList name='Plato Problem'>\n"
    "  <ParameterList name='Spatial Model'>\n"
    "    <ParameterList name='Domains'>\n"
    "      <ParameterList name='Design Volume'>\n"
    "        <Parameter name='Element Block' type='string' value='body'/>\n"
    "        <Parameter name='Material Model' type='string' value='Unobtainium'/>\n"
    "      </ParameterList>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <Parameter name='PDE Constraint' type='string' value='Elliptic'/>\n"
    "  <Parameter name='Self-Adjoint' type='bool' value='true'/>\n"
    "  <ParameterList name='Elliptic'>\n"
    "    <ParameterList name='Penalty Function'>\n"
    "      <Parameter name='Exponent' type='double' value='1.0'/>\n"
    "      <Parameter name='Minimum Value' type='double' value='0.0'/>\n"
    "      <Parameter name='Type' type='string' value='SIMP'/>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <ParameterList name='Material Models'>\n"
    "    <ParameterList name='Unobtainium'>\n"
    "      <ParameterList name='Isotropic Linear Elastic'>\n"
    "        <Parameter name='Poissons Ratio' type='double' value='0.3'/>\n"
    "        <Parameter name='Youngs Modulus' type='double' value='1.0e6'/>\n"
    "      </ParameterList>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <ParameterList name='Criteria'>\n"
    "    <ParameterList name='Internal Elastic Energy'>\n"
    "      <Parameter name='Type' type='string' value='Scalar Function'/>\n"
    "      <Parameter name='Scalar Function Type' type='string' value='Internal Elastic Energy'/>\n"
    "      <ParameterList name='Penalty Function'>\n"
    "        <Parameter name='Exponent' type='double' value='1.0'/>\n"
    "        <Parameter name='Minimum Value' type='double' value='0.0'/>\n"
    "        <Parameter name='Type' type='string' value='SIMP'/>\n"
    "      </ParameterList>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "</ParameterList>\n";

const char* kThermalDeck =
    "<ParameterList name='Plato Problem'>\n"
    "  <ParameterList name='Spatial Model'>\n"
    "    <ParameterList name='Domains'>\n"
    "      <ParameterList name='Design Volume'>\n"
    "        <Parameter name='Element Block' type='string' value='body'/>\n"
    "        <Parameter name='Material Model' type='string' value='Unobtainium'/>\n"
    "      </ParameterList>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <Parameter name='PDE Constraint' type='string' value='Elliptic'/>\n"
    "  <Parameter name='Self-Adjoint' type='bool' value='true'/>\n"
    "  <ParameterList name='Elliptic'>\n"
    "    <ParameterList name='Penalty Function'>\n"
    "      <Parameter name='Exponent' type='double' value='1.0'/>\n"
    "      <Parameter name='Minimum Value' type='double' value='0.0'/>\n"
    "      <Parameter name='Type' type='string' value='SIMP'/>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <ParameterList name='Material Models'>\n"
    "    <ParameterList name='Unobtainium'>\n"
    "      <ParameterList name='Thermal Conduction'>\n"
    "        <Parameter name='Thermal Conductivity' type='double' value='100.0'/>\n"
    "      </ParameterList>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <ParameterList name='Criteria'>\n"
    "    <ParameterList name='Internal Thermal Energy'>\n"
    "      <Parameter name='Type' type='string' value='Scalar Function'/>\n"
    "      <Parameter name='Scalar Function Type' type='string' value='Internal Thermal Energy'/>\n"
    "      <ParameterList name='Penalty Function'>\n"
    "        <Parameter name='Exponent' type='double' value='1.0'/>\n"
    "        <Parameter name='Minimum Value' type='double' value='0.0'/>\n"
    "        <Parameter name='Type' type='string' value='SIMP'/>\n"
    "      </ParameterList>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "</ParameterList>\n";

// configuration 1 of the benchmark at a small size: clamped x-, traction
// (0,0,-1) on x+, SIMP p = 3
const char* kComplianceDeck =
    "<ParameterList name='Plato Problem'>\n"
    "  <Parameter name='Physics' type='string' value='Mechanical'/>\n"
    "  <Parameter name='PDE Constraint' type='string' value='Elliptic'/>\n"
    "  <Parameter name='Self-Adjoint' type='bool' value='true'/>\n"
    "  <ParameterList name='Elliptic'>\n"
    "    <ParameterList name='Penalty Function'>\n"
    "      <Parameter name='Type' type='string' value='SIMP'/>\n"
    "      <Parameter name='Exponent' type='double' value='3.0'/>\n"
    "      <Parameter name='Minimum Value' type='double' value='1.0e-9'/>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <ParameterList name='Material Models'>\n"
    "    <ParameterList name='Unobtainium'>\n"
    "      <ParameterList name='Isotropic Linear Elastic'>\n"
    "        <Parameter name='Poissons Ratio' type='double' value='0.3'/>\n"
    "        <Parameter name='Youngs Modulus' type='double' value='1.0'/>\n"
    "      </ParameterList>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <ParameterList name='Criteria'>\n"
    "    <ParameterList name='Compliance'>\n"
    "      <Parameter name='Type' type='string' value='Scalar Function'/>\n"
    "      <Parameter name='Scalar Function Type' type='string' value='Internal Elastic Energy'/>\n"
    "      <ParameterList name='Penalty Function'>\n"
    "        <Parameter name='Type' type='string' value='SIMP'/>\n"
    "        <Parameter name='Exponent' type='double' value='3.0'/>\n"
    "        <Parameter name='Minimum Value' type='double' value='1.0e-9'/>\n"
    "      </ParameterList>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <ParameterList name='Natural Boundary Conditions'>\n"
    "    <ParameterList name='Traction Vector Boundary Condition'>\n"
    "      <Parameter name='Type' type='string' value='Uniform'/>\n"
    "      <Parameter name='Values' type='Array(double)' value='{0.0, 0.0, -1.0}'/>\n"
    "      <Parameter name='Sides' type='string' value='x+'/>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <ParameterList name='Essential Boundary Conditions'>\n"
    "    <ParameterList name='X Fixed Displacement Boundary Condition'>\n"
    "      <Parameter name='Type' type='string' value='Zero Value'/>\n"
    "      <Parameter name='Index' type='int' value='0'/>\n"
    "      <Parameter name='Sides' type='string' value='x-'/>\n"
    "    </ParameterList>\n"
    "    <ParameterList name='Y Fixed Displacement Boundary Condition'>\n"
    "      <Parameter name='Type' type='string' value='Zero Value'/>\n"
    "      <Parameter name='Index' type='int' value='1'/>\n"
    "      <Parameter name='Sides' type='string' value='x-'/>\n"
    "    </ParameterList>\n"
    "    <ParameterList name='Z Fixed Displacement Boundary Condition'>\n"
    "      <Parameter name='Type' type='string' value='Zero Value'/>\n"
    "      <Parameter name='Index' type='int' value='2'/>\n"
    "      <Parameter name='Sides' type='string' value='x-'/>\n"
    "    </ParameterList>\n"
    "  </ParameterList>\n"
    "  <ParameterList name='Linear Solver'>\n"
    "    <Parameter name='Tolerance' type='double' value='1e-12'/>\n"
    "    <Parameter name='Iterations' type='int' value='20000'/>\n"
    "  </ParameterList>\n"
    "</ParameterList>\n";

std::string j2_deck(const std::string& aMaterial, const std::string& aPenaltyMin, int aSteps,
                    const std::string& aCriteria, const std::string& aBCs) {
    return "<ParameterList name='Plato Problem'>\n"
           "  <ParameterList name='Spatial Model'>\n"
           "    <ParameterList name='Domains'>\n"
           "      <ParameterList name='Design Volume'>\n"
           "        <Parameter name='Element Block' type='string' value='body'/>\n"
           "        <Parameter name='Material Model' type='string' value='Unobtainium'/>\n"
           "      </ParameterList>\n"
           "    </ParameterList>\n"
           "  </ParameterList>\n"
           "  <Parameter name='Physics'          type='string'  value='Plasticity'/>\n"
           "  <Parameter name='PDE Constraint'   type='string'  value='Elliptic'/>\n"
           "  <ParameterList name='Material Models'>\n"
           "    <ParameterList name='Unobtainium'>\n" +
           aMaterial +
           "    </ParameterList>\n"
           "  </ParameterList>\n"
           "  <ParameterList name='Elliptic'>\n"
           "    <ParameterList name='Penalty Function'>\n"
           "      <Parameter name='Type' type='string' value='SIMP'/>\n"
           "      <Parameter name='Exponent' type='double' value='3.0'/>\n"
           "      <Parameter name='Minimum Value' type='double' value='" +
           aPenaltyMin +
           "'/>\n"
           "    </ParameterList>\n"
           "  </ParameterList>\n" +
           aCriteria +
           "  <ParameterList name='Time Stepping'>\n"
           "    <Parameter name='Initial Num. Pseudo Time Steps' type='int' value='" +
           std::to_string(aSteps) +
           "'/>\n"
           "    <Parameter name='Maximum Num. Pseudo Time Steps' type='int' value='" +
           std::to_string(aSteps) +
           "'/>\n"
           "  </ParameterList>\n"
           "  <ParameterList name='Newton-Raphson'>\n"
           "    <Parameter name='Stop Measure' type='string' value='residual'/>\n"
           "    <Parameter name='Maximum Number Iterations' type='int' value='20'/>\n"
           "  </ParameterList>\n"
           "  <ParameterList  name='Essential Boundary Conditions'>\n" +
           aBCs +
           "  </ParameterList>\n"
           "</ParameterList>\n";
}

std::string j2_material(const std::string& aE, const std::string& aNu, const std::string& aH) {
    return "      <ParameterList name='Isotropic Linear Elastic'>\n"
           "        <Parameter  name='Density' type='double' value='1e3'/>\n"
           "        <Parameter  name='Poissons Ratio' type='double' value='" +
           aNu +
           "'/>\n"
           "        <Parameter  name='Youngs Modulus' type='double' value='" +
           aE +
           "'/>\n"
           "      </ParameterList>\n"
           "      <ParameterList name='Plasticity Model'>\n"
           "        <ParameterList name='J2 Plasticity'>\n"
           "          <Parameter  name='Hardening Modulus Isotropic' type='double' value='" +
           aH +
           "'/>\n"
           "          <Parameter  name='Hardening Modulus Kinematic' type='double' value='" +
           aH +
           "'/>\n"
           "          <Parameter  name='Initial Yield Stress' type='double' value='" +
           aH +
           "'/>\n"
           "          <Parameter  name='Elastic Properties Penalty Exponent' type='double' value='3'/>\n"
           "          <Parameter  name='Elastic Properties Minimum Ersatz' type='double' value='1e-6'/>\n"
           "          <Parameter  name='Plastic Properties Penalty Exponent' type='double' value='2.5'/>\n"
           "          <Parameter  name='Plastic Properties Minimum Ersatz' type='double' value='1e-9'/>\n"
           "        </ParameterList>\n"
           "      </ParameterList>\n";
}

std::string bc(const std::string& aName, const std::string& aType, int aIndex, const std::string& aSides,
               const std::string& aFunction = "") {
    std::string s = "    <ParameterList name='" + aName + "'>\n      <Parameter name='Type' type='string' value='" +
                    aType + "'/>\n      <Parameter name='Index' type='int' value='" + std::to_string(aIndex) +
                    "'/>\n      <Parameter name='Sides' type='string' value='" + aSides + "'/>\n";
    if (!aFunction.empty()) s += "      <Parameter name='Function' type='string' value='" + aFunction + "'/>\n";
    return s + "    </ParameterList>\n";
}

// state u_k = aStep * (k + 1), control 1 (the DerivativeTests inputs)
pa::ScalarVector ramp(size_t n, double aStep) {
    std::vector<double> h(n);
    double v = 0.0;
    for (auto& x : h) x = (v += aStep);
    return pa::toDevice(h, "state");
}

pa::ScalarVector ones(size_t n) {
    pa::ScalarVector z("control", n);
    pa::fill(1.0, z);
    return z;
}

std::string elastostatic_residual() {
    auto tParams = pa::getParametersFromXmlString(kElasticDeck);
    auto tMesh = pa::getBoxMesh("TET4", 2);
    auto z = ones(tMesh->NumNodes());
    auto u = ramp(3 * tMesh->NumNodes(), 0.0001);
    pa::Elliptic::VectorFunction<pa::Mechanics<3>> esVectorFunction(tMesh, *tParams,
                                                                    tParams->get<std::string>("PDE Constraint"));
    auto residual = esVectorFunction.value(u, z);
    auto jacobian = esVectorFunction.gradient_u(u, z);
    auto jacobian_T = esVectorFunction.gradient_u_T(u, z);
    auto gradient_z = esVectorFunction.gradient_z(u, z);
    Json j;
    j.put("residual", pa::toHost(residual));
    j.put("jacobian", pa::toHost(jacobian->entries()));
    j.put("jacobian_T", pa::toHost(jacobian_T->entries()));
    j.put("gradient_z", pa::toHost(gradient_z->entries()));
    j.put("rowmap", pa::toHost(jacobian->rowMap()));
    j.put("colidx", pa::toHost(jacobian->columnIndices()));
    j.put("num_rows", (double)jacobian->numRows());
    j.put("block", (double)jacobian->numRowsPerBlock());
    j.put("gz_block_cols", (double)gradient_z->numColsPerBlock());
    auto gradient_x = esVectorFunction.gradient_x(u, z);
    j.put("gradient_x", pa::toHost(gradient_x->entries()));
    j.put("gx_block", (double)gradient_x->numRowsPerBlock() * 10 + gradient_x->numColsPerBlock());
    return j.str();
}

std::string internal_elastic_energy() {
    auto tParams = pa::getParametersFromXmlString(kElasticDeck);
    auto tMesh = pa::getBoxMesh("TET4", 2);
    auto z = ones(tMesh->NumNodes());
    auto u = ramp(3 * tMesh->NumNodes(), 0.0001);
    pa::Solutions tSolution;
    tSolution.set("State", u);
    pa::Elliptic::PhysicsScalarFunction<pa::Mechanics<3>> eeScalarFunction(tMesh, *tParams, "Internal Elastic Energy");
    Json j;
    j.put("value", eeScalarFunction.value(tSolution, z));
    j.put("gradient_u", pa::toHost(eeScalarFunction.gradient_u(tSolution, z, 0)));
    j.put("gradient_z", pa::toHost(eeScalarFunction.gradient_z(tSolution, z)));
    j.put("gradient_x", pa::toHost(eeScalarFunction.gradient_x(tSolution, z)));
    return j.str();
}

std::string thermostatic_residual() {
    auto tParams = pa::getParametersFromXmlString(kThermalDeck);
    auto tMesh = pa::getBoxMesh("TET4", 2);
    auto z = ones(tMesh->NumNodes());
    auto u = ramp(tMesh->NumNodes(), 0.1);
    pa::Elliptic::VectorFunction<pa::Thermal<3>> tsVectorFunction(tMesh, *tParams,
                                                                  tParams->get<std::string>("PDE Constraint"));
    Json j;
    j.put("residual", pa::toHost(tsVectorFunction.value(u, z)));
    j.put("jacobian", pa::toHost(tsVectorFunction.gradient_u(u, z)->entries()));
    j.put("gradient_z", pa::toHost(tsVectorFunction.gradient_z(u, z)->entries()));
    j.put("gradient_x", pa::toHost(tsVectorFunction.gradient_x(u, z)->entries()));
    return j.str();
}

std::string internal_thermal_energy() {
    auto tParams = pa::getParametersFromXmlString(kThermalDeck);
    auto tMesh = pa::getBoxMesh("TET4", 2);
    auto z = ones(tMesh->NumNodes());
    auto u = ramp(tMesh->NumNodes(), 0.1);
    pa::Elliptic::PhysicsScalarFunction<pa::Thermal<3>> tsScalarFunction(tMesh, *tParams, "Internal Thermal Energy");
    Json j;
    j.put("value", tsScalarFunction.value(u, z));
    j.put("gradient_u", pa::toHost(tsScalarFunction.gradient_u(u, z)));
    j.put("gradient_z", pa::toHost(tsScalarFunction.gradient_z(u, z)));
    j.put("gradient_x", pa::toHost(tsScalarFunction.gradient_x(u, z)));
    return j.str();
}

std::string compliance_problem() {
    auto tParams = pa::getParametersFromXmlString(kComplianceDeck);
    auto tMesh = pa::getBoxMesh(6, 3, 3, 2.0, 1.0, 1.0);
    const int N = tMesh->NumNodes();
    std::vector<double> zh(N);
    for (int i = 0; i < N; ++i) zh[i] = 0.3 + 0.7 * (double)((i * 7919) % 101) / 100.0;
    auto z = pa::toDevice(zh, "control");
    pa::Elliptic::Problem<pa::Mechanics<3>> tProblem(tMesh, *tParams);
    auto tSolution = tProblem.solution(z);
    const double F = tProblem.criterionValue(z, tSolution, "Compliance");
    auto dF = tProblem.criterionGradient(z, tSolution, "Compliance");
    auto dFdx = tProblem.criterionGradientX(z, tSolution, "Compliance");
    Json j;
    j.put("gradient_x", pa::toHost(dFdx));
    j.put("z", zh);
    j.put("u", pa::toHost(tSolution.get("State")));
    j.put("value", F);
    j.put("gradient", pa::toHost(dF));
    j.put("rel_residual", tProblem.solverStats().rel_residual);
    j.put("bc_dofs", tProblem.bcDofs());
    return j.str();
}

std::string elastic_solution_j2() {
    auto tParams = pa::getParametersFromXmlString(
        j2_deck(j2_material("1e11", "0.35", "1e9"), "1.0e-9", 1, "",
                bc("X Fixed Displacement Boundary Condition", "Zero Value", 0, "x-") +
                    bc("Y Fixed Displacement Boundary Condition", "Zero Value", 1, "x-") +
                    bc("Z Fixed Displacement Boundary Condition", "Zero Value", 2, "x-") +
                    bc("Applied Displacement Boundary Condition", "Time Dependent", 1, "x+", "-0.001*t")));
    auto tMesh = pa::getBoxMesh("TET4", 1);
    using PhysicsT = pa::InfinitesimalStrainPlasticity<3>;
    pa::PlasticityProblem<PhysicsT> tPlasticityProblem(tMesh, *tParams);
    tPlasticityProblem.readEssentialBoundaryConditions(*tParams);
    pa::ScalarVector tControls("Controls", tMesh->NumNodes());
    pa::fill(1.0, tControls);
    auto tSolution = tPlasticityProblem.solution(tControls);
    Json j;
    j.put("state", pa::toHost(tSolution.get("State", 0)));
    j.put("converged", tPlasticityProblem.converged() ? 1.0 : 0.0);
    return j.str();
}

// The deck of the reference's plastic-work criterion test.
std::shared_ptr<pa::ParameterList> criterion_test_j2_deck() {
    const std::string tCriteria =
        "  <ParameterList name='Criteria'>\n"
        "    <ParameterList name='My Plastic Work'>\n"
        "      <Parameter name='Type'                 type='string' value='Scalar Function'/>\n"
        "      <Parameter name='Scalar Function Type' type='string' value='Plastic Work'/>\n"
        "      <Parameter name='Multiplier'           type='double' value='-1.0'/>\n"
        "      <Parameter name='Exponent'             type='double' value='3.0'/>\n"
        "      <Parameter name='Minimum Value'        type='double' value='1.0e-9'/>\n"
        "    </ParameterList>\n"
        "  </ParameterList>\n";
    auto tParams = pa::getParametersFromXmlString(
        j2_deck(j2_material("1.0e6", "0.3", "1.0e3"), "1.0e-6", 4, tCriteria,
                bc("X Fixed Displacement Boundary Condition", "Zero Value", 0, "x-") +
                    bc("Y Fixed Displacement Boundary Condition", "Zero Value", 1, "y-") +
                    bc("Y Fixed Displacement Boundary Condition 2", "Zero Value", 1, "y+") +
                    bc("Z Fixed Displacement Boundary Condition", "Zero Value", 2, "z-") +
                    bc("Applied Displacement Boundary Condition", "Time Dependent", 0, "x+", "0.002*t")));
    return tParams;
}

std::string criterion_test_j2() {
    auto tParams = criterion_test_j2_deck();
    auto tMesh = pa::getBoxMesh("TET4", 2);
    using PhysicsT = pa::InfinitesimalStrainPlasticity<3>;
    pa::PlasticityProblem<PhysicsT> tPlasticityProblem(tMesh, *tParams);
    tPlasticityProblem.readEssentialBoundaryConditions(*tParams);
    const std::string tCriterionName("My Plastic Work");
    pa::ScalarVector tControls("Controls", tMesh->NumNodes());
    pa::fill(1.0, tControls);
    auto tSolution = tPlasticityProblem.solution(tControls);
    const double tObjValue = tPlasticityProblem.criterionValue(tControls, tCriterionName);
    auto tObjGrad = tPlasticityProblem.criterionGradient(tControls, tCriterionName);
    Json j;
    j.put("value", tObjValue);
    j.put("gradient", pa::toHost(tObjGrad));
    j.put("final_state", pa::toHost(tSolution.get("State", 3)));
    return j.str();
}

// the J2 local and global functions at the last state of the criterion test
std::string j2_worksets() {
    const std::string tCriteria = "";
    auto tParams = pa::getParametersFromXmlString(
        j2_deck(j2_material("1.0e6", "0.3", "1.0e3"), "1.0e-6", 4, tCriteria,
                bc("X Fixed Displacement Boundary Condition", "Zero Value", 0, "x-") +
                    bc("Y Fixed Displacement Boundary Condition", "Zero Value", 1, "y-") +
                    bc("Y Fixed Displacement Boundary Condition 2", "Zero Value", 1, "y+") +
                    bc("Z Fixed Displacement Boundary Condition", "Zero Value", 2, "z-") +
                    bc("Applied Displacement Boundary Condition", "Time Dependent", 0, "x+", "0.002*t")));
    auto tMesh = pa::getBoxMesh("TET4", 2);
    using PhysicsT = pa::InfinitesimalStrainPlasticity<3>;
    pa::PlasticityProblem<PhysicsT> tProblem(tMesh, *tParams);
    pa::ScalarVector z("Controls", tMesh->NumNodes());
    pa::fill(1.0, z);
    auto S = tProblem.solution(z);
    const auto &u = S.get("State", 3), &up = S.get("State", 2), &c = S.get("Local State", 3),
               &cp = S.get("Local State", 2), &pi = S.get("Projected Pressure Gradient", 3);
    pa::LocalVectorFunctionInc<PhysicsT> tLocal(tMesh, *tParams);
    pa::GlobalVectorFunctionInc<PhysicsT> tGlobal(tMesh, *tParams);
    pa::ScalarVector c2("Local State", c.size());
    tLocal.updateLocalState(u, up, c2, cp, z);
    Json j;
    j.put("u", pa::toHost(u));
    j.put("c", pa::toHost(c));
    j.put("cp", pa::toHost(cp));
    j.put("pi", pa::toHost(pi));
    j.put("c_updated", pa::toHost(c2));
    j.put("local_value", pa::toHost(tLocal.value(u, up, c, cp, z)));
    j.put("local_gradient_c", pa::toHost(tLocal.gradient_c(u, up, c, cp, z).values));
    j.put("local_gradient_u", pa::toHost(tLocal.gradient_u(u, up, c, cp, z).values));
    j.put("global_value", pa::toHost(tGlobal.value(u, up, c, cp, pi, z)));
    j.put("global_gradient_u", pa::toHost(tGlobal.gradient_u(u, up, c, cp, pi, z).values));
    j.put("global_gradient_z", pa::toHost(tGlobal.gradient_z(u, up, c, cp, pi, z).values));
    pa::ScalarVector Rt;
    auto Kt = tGlobal.condensedSystem(u, c, cp, pi, z, &Rt);
    j.put("tangent", pa::toHost(Kt->entries()));
    j.put("condensed_residual", pa::toHost(Rt));
    return j.str();
}

// the compliance deck on a mesh read from an Exodus II file (path in
// PA_EXODUS_MESH): value, gradient_u and the solved state
std::string exodus_problem() {
    const char* path = std::getenv("PA_EXODUS_MESH");
    if (!path) throw std::runtime_error("PA_EXODUS_MESH is not set");
    auto tParams = pa::getParametersFromXmlString(kComplianceDeck);
    auto tMesh = pa::readExodusMesh(path);
    const int N = tMesh->NumNodes();
    std::vector<double> zh(N), uh(3 * N);
    for (int i = 0; i < N; ++i) zh[i] = 0.3 + 0.7 * (double)((i * 7919) % 101) / 100.0;
    for (int i = 0; i < 3 * N; ++i) uh[i] = 1e-3 * (double)((i * 104729) % 211 - 105);
    auto z = pa::toDevice(zh, "control");
    auto u = pa::toDevice(uh, "state");
    pa::Elliptic::VectorFunction<pa::Mechanics<3>> tVf(tMesh, *tParams, "Elliptic");
    pa::Elliptic::Problem<pa::Mechanics<3>> tProblem(tMesh, *tParams);
    auto tSolution = tProblem.solution(z);
    Json j;
    j.put("num_nodes", (double)N);
    j.put("num_cells", (double)tMesh->NumElements());
    j.put("x_minus", tMesh->getNodeSetNodes("x-"));
    j.put("z", zh);
    j.put("u", uh);
    j.put("residual", pa::toHost(tVf.value(u, z)));
    j.put("jacobian", pa::toHost(tVf.gradient_u(u, z)->entries()));
    j.put("state", pa::toHost(tSolution.get("State")));
    j.put("compliance", tProblem.criterionValue(z, tSolution, "Compliance"));
    return j.str();
}

// The compliance deck driven through the Plato Engine operations
// (AnalyzeApp): the engine's control import, a solution, criterion value and
// gradient, a value with a target, the coordinate gradient by component, and
// an operation Parameter that retargets the traction.
const char* kComplianceOperations = R"(<?xml version="1.0"?>
<!-- operations of the compliance deck -->
<Operation>
  <Function>ComputeSolution</Function>
  <Name>Compute Displacement Solution</Name>
</Operation>
<Operation>
  <Function>ComputeCriterion</Function>
  <Name>Compute Objective</Name>
  <Criterion>Compliance</Criterion>
  <Output><Argument>Value</Argument><ArgumentName>Objective Value</ArgumentName></Output>
  <Output><Argument>Gradient</Argument><ArgumentName>Objective Gradient</ArgumentName></Output>
</Operation>
<Operation>
  <Function>ComputeCriterionValue</Function>
  <Name>Compute Constraint Value</Name>
  <Criterion>Compliance</Criterion>
  <Target>2.5</Target>
  <Output><Argument>Value</Argument><ArgumentName>Constraint Value</ArgumentName></Output>
</Operation>
<Operation>
  <Function>ComputeCriterionGradientX</Function>
  <Name>Compute Objective Gradient X</Name>
  <Criterion>Compliance</Criterion>
  <Output><Argument>Gradient</Argument><ArgumentName>Objective Gradient X</ArgumentName></Output>
</Operation>
<Operation>
  <Function>ComputeCriterionValue</Function>
  <Name>Compute Value Scaled Load</Name>
  <Criterion>Compliance</Criterion>
  <Parameter>
    <ArgumentName>Traction Z</ArgumentName>
    <Target>[Plato Problem]:[Natural Boundary Conditions]:[Traction Vector Boundary Condition]:Values(2)</Target>
    <InitialValue>-1.0</InitialValue>
  </Parameter>
  <Output><Argument>Value</Argument><ArgumentName>Scaled Value</ArgumentName></Output>
</Operation>
<Operation>
  <Function>Reinitialize</Function>
  <Name>Reinitialize on Change</Name>
  <OnChange>true</OnChange>
</Operation>
<Operation>
  <Function>UpdateProblem</Function>
  <Name>Update Problem</Name>
</Operation>
)";

std::string analyze_app() {
    auto tParams = pa::getParametersFromXmlString(std::string("<ParameterList name='Problem'>\n") + kComplianceDeck +
                                                  "</ParameterList>\n");
    auto tMesh = pa::getBoxMesh(6, 3, 3, 2.0, 1.0, 1.0);
    const int N = tMesh->NumNodes();
    std::vector<double> zh(N);
    for (int i = 0; i < N; ++i) zh[i] = 0.3 + 0.7 * (double)((i * 7919) % 101) / 100.0;
    pa::AnalyzeApp app(tMesh, *tParams, kComplianceOperations);
    app.initialize();
    std::vector<int> ids;
    app.exportDataMap(ids);
    app.importData("Topology", zh);
    Json j;
    std::vector<double> out;
    app.exportData("Topology", out);
    j.put("topology", out);
    j.put("num_ids", (double)ids.size());
    app.compute("Compute Displacement Solution");
    for (const char* d : {"displacement X", "displacement Y", "displacement Z"}) {
        app.exportData(d, out);
        j.put(d, out);
    }
    app.compute("Compute Objective");
    app.exportData("Objective Value", out);
    j.put("objective_value", out[0]);
    app.exportData("Objective Gradient", out);
    j.put("objective_gradient", out);
    app.compute("Compute Objective Gradient X");
    for (int d = 0; d < 3; ++d) {
        app.exportData("Objective Gradient X@" + std::to_string(d), out);
        j.put("gradient_x" + std::to_string(d), out);
    }
    app.compute("Compute Constraint Value");
    app.exportData("Constraint Value", out);
    j.put("constraint_value", out[0]);
    app.compute("Update Problem");
    // OnChange: the first call reinitializes, a second with the same
    // "Parameters" does not, a changed value does
    app.importValue("Parameters", {1.0});
    app.compute("Reinitialize on Change");
    app.compute("Reinitialize on Change");
    app.importValue("Parameters", {2.0});
    app.compute("Reinitialize on Change");
    app.compute("Compute Objective");
    app.exportData("Objective Value", out);
    j.put("objective_value_after_reinit", out[0]);
    app.importParameter("Traction Z", "Compute Value Scaled Load", -2.0);
    app.compute("Compute Value Scaled Load");
    app.exportData("Scaled Value", out);
    j.put("scaled_value", out[0]);
    // errors the reference raises too
    auto throws = [&](auto&& f) {
        try {
            f();
        } catch (const std::runtime_error&) {
            return 1.0;
        }
        return 0.0;
    };
    j.put("unknown_operation_throws", throws([&] { app.compute("No Such Operation"); }));
    j.put("undefined_parameter_throws", throws([&] { app.importParameter("Nope", "Compute Objective", 1.0); }));
    j.put("write_output_throws", throws([&] {
              pa::AnalyzeApp a2(tMesh, *tParams,
                                "<Operation><Function>WriteOutput</Function><Name>Write</Name></Operation>");
          }));
    j.put("reassigned_name_throws", throws([&] {
              pa::AnalyzeApp a3(tMesh, *tParams,
                                "<Operation><Function>ComputeCriterionValue</Function><Name>A</Name>"
                                "<Criterion>Compliance</Criterion><Output><Argument>Value</Argument>"
                                "<ArgumentName>V</ArgumentName></Output></Operation>"
                                "<Operation><Function>ComputeCriterionValue</Function><Name>B</Name>"
                                "<Criterion>Other</Criterion><Output><Argument>Value</Argument>"
                                "<ArgumentName>V</ArgumentName></Output></Operation>");
          }));
    return j.str();
}

// The plastic-work criterion deck driven through the Plato Engine
// operations: ComputeCriterion (solve, value, adjoint gradient) and the
// solution exported by dof name; the coordinate gradient of a plasticity
// criterion is not on this path and raises an error.
const char* kPlasticWorkOperations = R"(<?xml version="1.0"?>
<Operation>
  <Function>ComputeCriterion</Function>
  <Name>Compute Objective</Name>
  <Criterion>My Plastic Work</Criterion>
  <Output><Argument>Value</Argument><ArgumentName>Objective Value</ArgumentName></Output>
  <Output><Argument>Gradient</Argument><ArgumentName>Objective Gradient</ArgumentName></Output>
</Operation>
<Operation>
  <Function>ComputeCriterionGradientX</Function>
  <Name>Compute Objective Gradient X</Name>
  <Criterion>My Plastic Work</Criterion>
  <Output><Argument>Gradient</Argument><ArgumentName>Objective Gradient X</ArgumentName></Output>
</Operation>
)";

std::string analyze_app_j2() {
    auto tParams = criterion_test_j2_deck();
    auto tMesh = pa::getBoxMesh("TET4", 2);
    pa::AnalyzeApp app(tMesh, *tParams, kPlasticWorkOperations);
    app.initialize();
    Json j;
    std::vector<double> out;
    app.compute("Compute Objective");
    app.exportData("Objective Value", out);
    j.put("objective_value", out[0]);
    app.exportData("Objective Gradient", out);
    j.put("objective_gradient", out);
    for (const char* d : {"displacement X", "displacement Y", "displacement Z", "pressure"}) {
        app.exportData(d, out);
        j.put(d, out);
    }
    double tThrows = 0.0;
    try {
        app.compute("Compute Objective Gradient X");
    } catch (const std::runtime_error&) {
        tThrows = 1.0;
    }
    j.put("gradient_x_throws", tThrows);
    return j.str();
}

// ---------------------------------------------------------------- parser
int parser_checks() {
    int failures = 0;
    auto expect = [&](bool ok, const std::string& what) {
        if (!ok) {
            std::printf("FAIL %s\n", what.c_str());
            ++failures;
        }
    };
    auto P = pa::getParametersFromXmlString(kComplianceDeck);
    expect(P->name() == "Plato Problem", "root name");
    expect(P->get<std::string>("PDE Constraint") == "Elliptic", "string parameter");
    expect(P->get<bool>("Self-Adjoint") == true, "bool parameter");
    expect(P->sublist("Elliptic").sublist("Penalty Function").get<double>("Exponent") == 3.0, "nested double");
    expect(P->sublist("Linear Solver").get<int>("Iterations") == 20000, "int parameter");
    const auto v = P->sublist("Natural Boundary Conditions")
                       .sublist("Traction Vector Boundary Condition")
                       .get<std::vector<double>>("Values");
    expect(v.size() == 3 && v[0] == 0.0 && v[2] == -1.0, "Array(double)");
    expect(P->isSublist("Criteria") && !P->isParameter("Criteria"), "isSublist");
    expect(P->isType<std::string>("Physics") && !P->isType<double>("Physics"), "isType");
    expect(P->get<double>("Missing", 7.5) == 7.5, "default value");
    const auto names = P->sublist("Essential Boundary Conditions").sublistNames();
    expect(names.size() == 3 && names[2] == "Z Fixed Displacement Boundary Condition", "sublist order");
    bool threw = false;
    try {
        P->get<double>("PDE Constraint");
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect(threw, "wrong type throws");
    threw = false;
    try {
        pa::getParametersFromXmlString("<ParameterList name='a'><Parameter name='x' type='int' value='1'/>");
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect(threw, "unterminated list throws");
    auto Q = pa::getParametersFromXmlString(
        "<?xml version='1.0'?>\n<!-- comment -->\n<ParameterList name=\"q\">"
        "<Parameter name=\"s\" type=\"string\" value=\"a &lt; b\"/><ParameterList name=\"e\"/></ParameterList>");
    expect(Q->get<std::string>("s") == "a < b", "double quotes and entities");
    expect(Q->isSublist("e"), "empty sublist");
    expect(pa::Expression("0.002*t")(0.5) == 0.001, "a*t");
    expect(pa::Expression("-0.001*t")(1.0) == -0.001, "-a*t");
    expect(std::fabs(pa::Expression("2*(t+1)^2 - 3/4")(1.0) - 7.25) < 1e-15, "precedence");
    expect(pa::Expression("1.5e-3")(0.3) == 1.5e-3, "constant");
    expect(std::fabs(pa::Expression("a = 2; a*sin(pi*t/2)")(1.0) - 2.0) < 1e-15, "assignment and functions");
    threw = false;
    try {
        pa::Expression("q*t")(1.0);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect(threw, "unknown symbol throws");
    // operation documents and parameter targets
    const auto ops = pa::detail::parseXmlElements(kComplianceOperations);
    expect(ops.all("Operation").size() == 7, "operation count");
    expect(ops.all("Operation")[0]->get("Function") == "ComputeSolution", "element text");
    expect(ops.all("Operation")[1]->all("Output").size() == 2, "repeated elements");
    expect(ops.all("Operation")[1]->all("Output")[1]->get("ArgumentName") == "Objective Gradient", "nested text");
    expect(ops.all("Operation")[4]->first("Parameter")->get("Target") ==
               "[Plato Problem]:[Natural Boundary Conditions]:[Traction Vector Boundary Condition]:Values(2)",
           "parameter target");
    expect(pa::detail::parseXmlElements("<a><b>x &amp; y</b><c/></a>").first("a")->get("b") == "x & y",
           "entities and self-closing elements");
    auto D = pa::getParametersFromXmlString(std::string("<ParameterList name='Problem'>\n") + kComplianceDeck +
                                            "</ParameterList>\n");
    auto D2 = D->clone();
    pa::detail::parse_inline(*D2, "[Plato Problem]:[Natural Boundary Conditions]:[Traction Vector Boundary Condition]:Values(2)",
                             -2.5);
    pa::detail::parse_inline(*D2, "[Plato Problem]:[Linear Solver]:Tolerance", 1e-10);
    const auto& tr = D2->sublist("Plato Problem").sublist("Natural Boundary Conditions");
    expect(tr.sublist("Traction Vector Boundary Condition").get<std::vector<double>>("Values")[2] == -2.5,
           "inline target, array entry");
    expect(D2->sublist("Plato Problem").sublist("Linear Solver").get<double>("Tolerance") == 1e-10,
           "inline target, scalar");
    expect(D->sublist("Plato Problem")
                   .sublist("Natural Boundary Conditions")
                   .sublist("Traction Vector Boundary Condition")
                   .get<std::vector<double>>("Values")[2] == -1.0,
           "clone is deep");
    for (const char* bad : {"<a><b></a>", "<a>", "</a>", "<!-- x"}) {
        threw = false;
        try {
            pa::detail::parseXmlElements(bad);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        expect(threw, std::string("malformed XML throws: ") + bad);
    }
    threw = false;
    try {
        pa::detail::parse_inline(*D2, "[Plato Problem]:[No Such List]:x", 1.0);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect(threw, "inline target into a missing list throws");
    std::printf("%s %d\n", failures ? "parser checks failed:" : "parser checks passed", failures);
    return failures ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::vector<std::pair<std::string, std::function<std::string()>>> cases = {
        {"ElastostaticResidual3D", elastostatic_residual},
        {"InternalElasticEnergy3D", internal_elastic_energy},
        {"ThermostaticResidual3D", thermostatic_residual},
        {"InternalThermalEnergy3D", internal_thermal_energy},
        {"Compliance", compliance_problem},
        {"ElastoPlasticity_ElasticSolution3D", elastic_solution_j2},
        {"ElastoPlasticity_CriterionTest_3D", criterion_test_j2},
        {"J2Worksets", j2_worksets},
        {"ExodusMesh", exodus_problem},
        {"AnalyzeApp", analyze_app},
        {"AnalyzeAppPlasticity", analyze_app_j2},
    };
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <case>|all|--list|--parser|--exodus <file>\n", argv[0]);
        return 2;
    }
    const std::string what = argv[1];
    if (what == "--list") {
        for (const auto& c : cases) std::printf("%s\n", c.first.c_str());
        return 0;
    }
    if (what == "--parser") return parser_checks();
    if (what == "--exodus") {
        // host-side Exodus reader only (no device work): counts, or the
        // reader's error with status 3
        if (argc < 3) return 2;
        try {
            const Plato::amd::ExodusData d = Plato::amd::readExodusData(argv[2]);
            std::printf("{\"nodes\": %zu, \"cells\": %zu, \"side_sets\": %zu, \"node_sets\": %zu}\n",
                        d.coords.size() / 3, d.conn.size() / 4, d.side_sets.size(), d.node_sets.size());
            return 0;
        } catch (const std::exception& e) {
            std::printf("{\"error\": \"%s\"}\n", e.what());
            return 3;
        }
    }
    int ran = 0;
    for (const auto& c : cases) {
        if (what != c.first && (what != "all" || c.first == "ExodusMesh")) continue;
        ++ran;
        std::fprintf(stderr, "case %s\n", c.first.c_str());
        try {
            std::printf("{\"case\": \"%s\", \"outputs\": %s}\n", c.first.c_str(), c.second().c_str());
        } catch (const std::exception& e) {
            std::printf("{\"case\": \"%s\", \"error\": \"%s\"}\n", c.first.c_str(), e.what());
        }
        std::fflush(stdout);
    }
    if (!ran) {
        std::fprintf(stderr, "unknown case %s\n", what.c_str());
        return 2;
    }
    return 0;
}
