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

std::string criterion_test_j2() {
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
