// tables.hpp — host FP64 construction of every table the depth march reads.
#pragma once

#include <string>
#include <vector>

#include "common.hpp"

namespace bfpn {

bfpn_deltas compute_deltas(const bfpn_grid& g);

struct MaterialTables {
    std::vector<double> mu;     // Ng x 3 (data_cross.txt / data_b / data_a)
    std::vector<double> sigma;  // Ng     (cross_total.txt, optionally calibrated)
    int mu_rows = 0;
};

MaterialTables load_material_tables(const std::string& data_path, int Ng, bool calibrate);

std::vector<double> stopping_power_table(const std::string& material, int Ng, double Lg, double dg);
std::vector<double> straggling_table(int Ng);
std::vector<double> transport_xs_table(int Ng, double dg);
std::vector<double> energy_grid(int Ng, double dg);
std::vector<double> direction_table(const bfpn_grid& g, const bfpn_deltas& d);

struct SourceKernels {
    std::vector<double> e1, e2;  // Ng x Ng, row = output group, col = source group
    std::vector<int> beg, end;   // nonzero column span per row
    std::vector<double> v;       // Ng x nang x nang, [g'][in][out]
};
SourceKernels build_source_kernels(int Ng, double dg, int nang, const std::vector<double>& mu,
                                   const std::vector<double>& sigma, const std::vector<double>& om);

// Eq.23 beam, separable as F[g][s][m] = space_angle[s*nang+m] * energy[g].
struct BeamFactors {
    std::vector<double> space_angle;  // nyz x nang (already scaled)
    std::vector<double> energy;       // Ng
    double spatial_mass = 0, angular_mass = 0, energy_mass = 0, scale = 1;
};
BeamFactors beam_factors(const bfpn_grid& g, const bfpn_phys& p, const bfpn_deltas& d,
                         const std::vector<double>& om);

std::vector<EnergyCoef> energy_coefs(const std::vector<double>& S, const std::vector<double>& sigma,
                                     const double* T, bool straggling, double h, double dg,
                                     double rho, int Ng);

// Upwind energy step coefficients (src/energy_solver.cu:88-106 with the
// straggling enabled by `straggling`; src/bfpm_solver.cu:600-622 with rho = 1).
std::vector<UpwindCoef> upwind_coefs(const std::vector<double>& S, const std::vector<double>& sigma,
                                     const double* T, bool straggling, double h, double dg,
                                     double rho, int Ng);

std::vector<DoseCoef> dose_coefs(const std::vector<double>& S, const std::vector<double>& sigma,
                                 const std::vector<double>& en, double dg, bool catastrophic, int Ng);

// Crank-Nicolson DST-I factors coeff[g][nu*Nom+mu] (ratio * 4/((Nom+1)(Nmu+1)))
// and the sine bases S_u (Nom x Nom), S_v (Nmu x Nmu).
std::vector<double> angle_factors(const double* sig_tr, int Ng, int Nmu, int Nom, double du,
                                  double dv, double dt, double rho);
std::vector<double> sine_basis(int n);

}  // namespace bfpn
