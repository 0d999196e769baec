// Device helpers shared by the cell-evaluation kernels of the elliptic path:
// material constants, the MSIMP penalty and the XCD-aware block order.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace pa {

struct Material {
    double lambda, mu;  // isotropic: C(0,1) and C(3,3) of the Voigt stiffness
    double k;           // conductivity
    double p, emin;     // MSIMP exponent, minimum value
};

// Blocks are dispatched round-robin over the 8 XCDs; renumber them so each
// XCD walks a contiguous range of rows and its L2 keeps the shared nodes.
__device__ inline int xcd_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = (nb + 7) / 8;
    const int x = b % 8, i = b / 8;
    const int full = nb % 8 == 0 ? 8 : nb % 8;  // XCDs that receive `per` blocks
    if (x < full) return x * per + i;
    return full * per + (x - full) * (per - 1) + i;
}

// MSIMP s(rho) = emin + (1 - emin) rho^p (src/Simp.hpp:130-134); integral
// exponents by products (matches pow to an ulp)
__device__ inline double penal(double rho, double p, double emin) {
    double r;
    if (p == 3.0) r = rho * rho * rho;
    else if (p == 1.0) r = rho;
    else if (p == 2.0) r = rho * rho;
    else r = pow(rho, p);
    return emin + (1.0 - emin) * r;
}

// The same with the exponent resolved at compile time: PK = 1, 2 or 3 for
// those integral exponents, 0 for any other (pow).
template <int PK>
__device__ inline double penal_k(double rho, double p, double emin) {
    double r;
    if constexpr (PK == 3) r = rho * rho * rho;
    else if constexpr (PK == 2) r = rho * rho;
    else if constexpr (PK == 1) r = rho;
    else r = pow(rho, p);
    return emin + (1.0 - emin) * r;
}

// Compile-time exponent class of a penalty exponent (see penal_k).
inline int penal_kind(double p) { return p == 3.0 ? 3 : p == 2.0 ? 2 : p == 1.0 ? 1 : 0; }

__device__ inline double dpenal(double rho, double p, double emin) {
    double r;
    if (p == 3.0) r = 3.0 * rho * rho;
    else if (p == 1.0) r = 1.0;
    else if (p == 2.0) r = 2.0 * rho;
    else r = rho == 0.0 ? 0.0 : p * pow(rho, p - 1.0);
    return (1.0 - emin) * r;
}

}  // namespace pa
