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
// Per-group eq15 sweep coefficients.  The new (psi^2, psi^1) = (f, F) of a
// group are linear in x = (P_hi, R_hi, P, R, P_lo, Q, Delta_hi): old F and f
// of group g+1, old F and f of g, old F of g-1 (straggling), the source, and
// the new F - f of g+1; f = sum af[i] x[i], F = sum aF[i] x[i].
struct EnergyCoef {
    double af[7];
    double aF[7];
    double w1, w2;  // Eq.24 dose weights of the group (when the sweep accumulates the IDD)
};

// Per-group coefficients of the first-order upwind energy step (the legacy
// energy model and the lite-memory step): for the half step h,
//   F' = max((max(F,0)/h + c_hi F'_{g+1} + Q) / denom + strag * lap(F), 0)
//   f' = max((f/h + c_hi f'_{g+1}) / denom, 0)
// with c_hi = rho S_hi/dg, denom = max(1/h + rho S_lo/dg + sigma, eps) and
// strag = h/2 T_g/dg^2 (0 without straggling).
struct UpwindCoef {
    double h, c_hi, denom, strag;
};

// Per-group dose weights: IDD = scale * sum w_yz(s) (psi1*w1[g] + psi2*w2[g]).
struct DoseCoef {
    double w1;
    double w2;
};

}  // namespace bfpn
