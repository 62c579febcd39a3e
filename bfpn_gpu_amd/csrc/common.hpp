// common.hpp — shared definitions of the MI355X BFPn library (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/bfpn.h"

namespace bfpn {

constexpr double kEps = 1e-10;  // Numerics::EPSILON, include/config.hpp:39

// Error raised inside the library; converted to a bfpn_status at the C ABI.
struct Error : std::runtime_error {
    bfpn_status status;
    Error(bfpn_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void set_last_error(const std::string& msg);

#define BFPN_HIP(expr)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::bfpn::Error(BFPN_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define BFPN_REQUIRE(cond, msg)                                                           \
    do {                                                                                  \
        if (!(cond)) throw ::bfpn::Error(BFPN_E_INVALID, msg);                            \
    } while (0)

// Runs `body` and maps exceptions onto status codes (never exit()).
template <class F>
bfpn_status guarded(F&& body) {
    try {
        body();
        return BFPN_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return BFPN_E_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return BFPN_E_INVALID;
    }
}

// Per-energy-group coefficients of the eq15 DG/CN sweep for one half step h.
// They depend only on (g, h, rho, dg, tables), so the host computes them once
// in FP64 (the reference recomputes them in every thread,
// src/energy_solver.cu:108-150).  Field names follow that code.
struct EnergyCoef {
    double inv_h;     // 1/h
    double half_sig;  // 0.5*sigma
    double c_up;      // 0.5*rho*S_hi/dg
    double c_cur;     // 0.5*rho*S_lo/dg
    double inv_F;     // 1/den_F
    double A;         // 1.5*rho*S_hi/dg * inv_F
    double B;         // F_from_f  = 0.5*rho*S_lo/dg * inv_F
    double Cc;        // F_from_hi = 0.5*rho*S_hi/dg * inv_F
    double k_up;      // 1.5*rho*S_hi/dg   (also "high2")
    double k_cur;     // 1.5*rho*S_lo/dg
    double k_F;       // 1.5*rho*(S_lo+S_hi)/dg
    double k_f;       // 0.5*rho*(S_hi-S_lo)/dg
    double H;         // high2 - A*0.5*rho*S_hi/dg
    double inv_coe;   // 1/max(coe, eps)
    double a_hi;      // 0.5*rho*T_hi/dg^2   (0 without straggling)
    double a_lo;      // 0.5*rho*T_lo/dg^2
    double a_g;       // rho*T_g/dg^2
    double pad;
};

// Per-group dose weights: IDD = scale * sum w_yz(s) (psi1*w1[g] + psi2*w2[g]).
struct DoseCoef {
    double w1;
    double w2;
};

}  // namespace bfpn
