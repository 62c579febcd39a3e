// tables.cpp — host FP64 tables for the in-core eq15 depth march.
//
// These run once per beam on the host.  FMA contraction is switched off so
// that every table is the same IEEE result the reference's host code computes
// (thresholds such as |K| < 1e-5 and the exact-zero tests on the kernels make
// the last bit matter).  Constants: include/config.hpp:8-44.

#pragma clang fp contract(off)

#include <thread>

#include "tables.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <limits>

namespace bfpn {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kElectronMass = 9.10956e-31;  // kg
constexpr double kProtonMass = 1.673e-27;      // kg
constexpr double kFineStructure = 1.0 / 137.0;
constexpr double kLightSpeed = 2.998e8;        // m/s
constexpr double kHbarC = 197.3;               // MeV fm
constexpr double kMeVinJ = 1.6e-13;
constexpr double kAngularCut = 1e-3;           // KERNEL_THRESHOLD
constexpr double kEnergyCut = 1e-5;            // SPARSE_THRESHOLD

struct Element {
    int Z;
    double mass_fraction;
};

// Bethe ingredients, src/bfpm_solver.cu:150-182
double excitation_ev(int Z) { return Z == 1 ? 19.0 : (Z <= 13 ? 11.2 + 11.7 * Z : 52.8 + 8.71 * Z); }

double mass_number(int Z) {
    switch (Z) {
        case 1: return 1.0;
        case 6: return 12.0;
        case 7: return 14.0;
        case 8: return 16.0;
        case 11: return 23.0;
        case 12: return 24.0;
        case 15: return 31.0;
        case 16: return 32.0;
        case 18:
        case 20: return 40.0;
        default: return 2.0 * Z;
    }
}

double bethe(double E, int Z) {
    double b2 = E * (E + 2.0 * 938.3) / ((E + 938.3) * (E + 938.3));
    b2 = std::max(b2, kEps);
    const double g2 = 1.0 / std::max(1.0 - b2, kEps);
    const double arg = std::max(2.0 * 0.51099895e6 * b2 * g2 / excitation_ev(Z), 1.0 + kEps);
    return 0.307 * (static_cast<double>(Z) / mass_number(Z)) / b2 * (std::log(arg) - b2);
}

// Mass fractions, src/bfpm_solver.cu:184-196 ("" and unknown -> water)
std::vector<Element> elements_of(const std::string& material) {
    if (material == "bone")
        return {{1, 0.04200}, {6, 0.1940}, {7, 0.04000}, {8, 0.4250}, {11, 0.001000},
                {12, 0.002000}, {15, 0.09200}, {16, 0.003000}, {20, 0.2010}};
    if (material == "air") return {{6, 0.0001248}, {7, 0.7553}, {8, 0.2318}, {18, 0.01283}};
    return {{1, 0.1111}, {8, 0.8889}};
}

double mixture_bethe(const std::vector<Element>& els, double E) {
    double acc = 0.0;
    for (const Element& el : els) acc += el.mass_fraction * bethe(E, el.Z);
    return acc;
}

}  // namespace

bfpn_deltas compute_deltas(const bfpn_grid& g) {
    bfpn_deltas d;
    d.dt = g.Lx / g.Nx;
    d.dy = g.Ly / g.Ny;
    d.dz = g.Lz / g.Nz;
    d.dg = g.Lg / g.Ng;
    d.du = g.Lu / (g.Nom + 1);  // u pairs with Nom
    d.dv = g.Lv / (g.Nmu + 1);  // v pairs with Nmu
    return d;
}

// Table files, src/bfpm_solver.cu:1436-1493: data_cross.txt (or data_b.txt /
// data_a.txt in a directory named bone / air), at least Ng-1 rows of 3
// numbers (missing rows stay zero); cross_total.txt with Ng numbers.
MaterialTables load_material_tables(const std::string& data_path, int Ng, bool calibrate) {
    MaterialTables t;
    t.mu.assign(static_cast<size_t>(Ng) * 3, 0.0);
    t.sigma.assign(Ng, 0.0);

    std::string path = data_path + "/data_cross.txt";
    std::ifstream in(path);
    if (!in) {
        const size_t cut = data_path.find_last_of("/\\");
        const std::string leaf = data_path.substr(cut == std::string::npos ? 0 : cut + 1);
        if (leaf == "bone" || leaf == "air") {
            path = data_path + "/data_" + leaf.substr(0, 1) + ".txt";
            in.clear();
            in.open(path);
        }
    }
    if (!in) throw Error(BFPN_E_IO, "Cannot open scattering data file in " + data_path);
    int rows = 0;
    double a, b, c;
    while (rows < Ng && (in >> a >> b >> c)) {
        t.mu[rows * 3 + 0] = a;
        t.mu[rows * 3 + 1] = b;
        t.mu[rows * 3 + 2] = c;
        ++rows;
    }
    if (rows < Ng - 1)
        throw Error(BFPN_E_IO, "Scattering data " + path + " has only " + std::to_string(rows) +
                                   " rows; need at least " + std::to_string(Ng - 1));
    t.mu_rows = rows;

    const std::string spath = data_path + "/cross_total.txt";
    std::ifstream sin(spath);
    if (!sin) throw Error(BFPN_E_IO, "Cannot open total cross-section file " + spath);
    int srows = 0;
    while (srows < Ng && (sin >> a)) t.sigma[srows++] = a;
    if (srows != Ng)
        throw Error(BFPN_E_IO, "Total cross-section data " + spath + " has " + std::to_string(srows) +
                                   " rows; need " + std::to_string(Ng));
    if (calibrate) {  // legacy empirical correction, off by default
        for (double& s : t.sigma) s *= 0.9;
        for (int i = 401; i < 444 && i < Ng; ++i) t.sigma[i] -= 0.0002 * (0.5 + 0.5 * (i - 401));
    }
    return t;
}

// S(E) on the Ng+1 group edges, src/bfpm_solver.cu:1570-1597
std::vector<double> stopping_power_table(const std::string& material, int Ng, double Lg, double dg) {
    const auto mix = elements_of(material);
    const auto water = elements_of("water");
    std::vector<double> S(Ng + 1);
    for (int i = 0; i <= Ng; ++i) {
        const double E = i == 0 ? 1.0 : (i == Ng ? Lg + 1.0 : (i - 0.5) * dg + 1.0);
        const double num = mixture_bethe(mix, E);
        const double den = mixture_bethe(water, E);
        const double b2 = E * (E + 2.0 * 938.3) / std::pow(E + 938.3, 2);
        const double legacy = 0.170 / b2 * (std::log(1.02e6 * b2 / (1.0 - b2)) - b2 - 4.31) + 0.02;
        S[i] = legacy * (den > kEps ? num / den : 1.0);
    }
    return S;
}

// Straggling coefficient, src/bfpm_solver.cu:1599-1615 (energy independent)
std::vector<double> straggling_table(int Ng) {
    const double eta =
        8.99e18 * 4.0 * kPi * std::pow(1.6e-19, 4) / std::pow(1.6e-13, 2) / 100.0;
    return std::vector<double>(Ng, eta * 3.34e29 * 1.2);
}

// Screened-Rutherford transport cross section at group centres,
// src/bfpm_solver.cu:1617-1636
std::vector<double> transport_xs_table(int Ng, double dg) {
    std::vector<double> sig(Ng);
    const double Zt = 3;
    for (int i = 0; i < Ng; ++i) {
        const double E = (i + 0.5) * dg + 1.0;
        const double beta = std::sqrt(2.0 * E * kMeVinJ / kProtonMass) / (3.0 * kLightSpeed);
        const double eta = std::pow(Zt, 2.0 / 3.0) * std::pow(kFineStructure, 2) *
                           std::pow(kElectronMass / kProtonMass, 2) / (beta * beta);
        const double screen = std::log((eta + 1.0) / eta) - 1.0 / (eta + 1.0);
        sig[i] = 2.0 * kPi * 3.34 * std::pow(kHbarC, 2) * std::pow(kFineStructure, 2) / 4.0 * Zt *
                 Zt / (E * E) * screen / 1000.0;
    }
    return sig;
}

std::vector<double> energy_grid(int Ng, double dg) {
    std::vector<double> en(Ng + 1);
    for (int i = 0; i <= Ng; ++i) en[i] = i * dg + 1.0;
    return en;
}

// Direction of every angular node m = iv*Nom + iu: (omega_y, omega_z) =
// (-v[Nmu-1-iv], -u[iu]), src/bfpm_solver.cu:1516-1533
std::vector<double> direction_table(const bfpn_grid& g, const bfpn_deltas& d) {
    std::vector<double> om(2 * static_cast<size_t>(g.Nmu) * g.Nom);
    for (int iv = 0; iv < g.Nmu; ++iv) {
        const double v = -0.5 * g.Lv + (g.Nmu - iv) * d.dv;  // v[Nmu-1-iv]
        for (int iu = 0; iu < g.Nom; ++iu) {
            const double u = -0.5 * g.Lu + (iu + 1) * d.du;
            const size_t m = static_cast<size_t>(iv) * g.Nom + iu;
            om[2 * m] = -v;
            om[2 * m + 1] = -u;
        }
    }
    return om;
}

// Catastrophic-collision kernels, src/bfpm_solver.cu:1638-1813.
//   energy: K0[i][j] = dg/l_j exp(-(i-j+1/2) dg/l_j) (i >= j, l_j = mu[j][2]),
//           shifted one row below the diagonal, flipped in both indices,
//           thresholded at 1e-5 and weighted by sigma[col];
//           K1 = (K + K_next_row)/2, K2 = (K_next_row - K)/2.
//   angle:  K[g'][in][out] ~ theta^(mu0-1) exp(-theta/mu1) (1+|w_in|^2)^-3/2,
//           normalised over `in` per `out`, thresholded at 1e-3.
SourceKernels build_source_kernels(int Ng, double dg, int nang, const std::vector<double>& mu,
                                   const std::vector<double>& sigma, const std::vector<double>& om) {
    const size_t n = Ng;
    auto base_kernel = [&](int i, int j) -> double {  // K0 before the shift
        if (i >= Ng - 1 || j >= Ng - 1) return 0.0;
        const double len = mu[j * 3 + 2];
        if (len == 0 || i < j) return 0.0;
        return dg / len * std::exp(-(i - j + 0.5) * dg / len);
    };
    auto shifted = [&](int i, int j) -> double {
        if (j >= Ng - 1 || i == j) return 0.0;
        return i < j ? base_kernel(i, j) : base_kernel(i - 1, j);
    };
    std::vector<double> K(n * n);
    for (int r = 0; r < Ng; ++r)
        for (int c = 0; c < Ng; ++c) {
            double v = shifted(Ng - 1 - r, Ng - 1 - c);
            if (std::abs(v) < kEnergyCut) v = 0.0;
            K[r * n + c] = v * sigma[c];
        }
    SourceKernels k;
    k.e1.assign(n * n, 0.0);
    k.e2.assign(n * n, 0.0);
    k.beg.assign(Ng, 0);
    k.end.assign(Ng, 0);
    for (int r = 0; r < Ng; ++r) {
        int lo = Ng, hi = 0;
        for (int c = 0; c < Ng; ++c) {
            const double cur = K[r * n + c];
            const double nxt = r + 1 < Ng ? K[(r + 1) * n + c] : 0.0;
            const double v1 = 0.5 * (cur + nxt);
            const double v2 = 0.5 * (-cur + nxt);
            k.e1[r * n + c] = v1;
            k.e2[r * n + c] = v2;
            if (v1 != 0.0 || v2 != 0.0) {
                lo = std::min(lo, c);
                hi = std::max(hi, c + 1);
            }
        }
        if (lo < Ng) {
            k.beg[r] = lo;
            k.end[r] = hi;
        }
    }

    const size_t na = nang;
    std::vector<double> theta(na * na);
    for (size_t o = 0; o < na; ++o) {
        const double oy = om[2 * o], oz = om[2 * o + 1];
        const double on = std::sqrt(1.0 + oy * oy + oz * oz);
        for (size_t i = 0; i < na; ++i) {
            const double iy = om[2 * i], iz = om[2 * i + 1];
            const double inorm = std::sqrt(1.0 + iy * iy + iz * iz);
            double c = (1.0 + iy * oy + iz * oz) / std::max(inorm * on, kEps);
            c = std::min(1.0, std::max(-1.0, c));
            theta[i * na + o] = std::abs(std::acos(c) * 180.0 / kPi);
        }
    }
    std::vector<double> weight_in(na);
    for (size_t i = 0; i < na; ++i) {
        const double iy = om[2 * i], iz = om[2 * i + 1];
        weight_in[i] = std::pow(1.0 + iy * iy + iz * iz, -1.5);
    }
    k.v.assign(n * na * na, 0.0);
    // Groups are independent: build them on all host threads (same libm
    // calls per entry, so the table does not depend on the thread count).
    auto build_group = [&](int g) {
        const double shape = mu[g * 3 + 0], width = mu[g * 3 + 1];
        if (shape == 0.0 || width == 0.0) return;
        double* Kg = k.v.data() + static_cast<size_t>(g) * na * na;
        for (size_t o = 0; o < na; ++o) {
            double total = 0.0;
            for (size_t i = 0; i < na; ++i) {
                const double th = theta[i * na + o];
                double w = std::pow(th, shape - 1.0) * std::exp(-th / width) * weight_in[i];
                if (!std::isfinite(w)) w = 0.0;
                Kg[i * na + o] = w;
                total += w;
            }
            if (total > kEps)
                for (size_t i = 0; i < na; ++i) {
                    double w = Kg[i * na + o] / total;
                    Kg[i * na + o] = w < kAngularCut ? 0.0 : w;
                }
        }
    };
    const int workers = std::max(1, std::min<int>(Ng - 1, static_cast<int>(std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w)
        pool.emplace_back([&, w] {
            for (int g = w; g + 1 < Ng; g += workers) build_group(g);
        });
    for (auto& t : pool) t.join();
    return k;
}

// Eq.23 pencil beam, src/bfpm_solver.cu:1815-1884 and 1978-2025
BeamFactors beam_factors(const bfpn_grid& g, const bfpn_phys& p, const bfpn_deltas& d,
                         const std::vector<double>& om) {
    BeamFactors b;
    const int nyz = (g.Ny + 1) * (g.Nz + 1);
    const int nang = g.Nmu * g.Nom;
    double cv = std::numeric_limits<double>::max(), cu = std::numeric_limits<double>::max();
    for (int j = 0; j < g.Nmu; ++j) cv = std::min(cv, std::abs(-0.5 * g.Lv + (j + 1) * d.dv));
    for (int i = 0; i < g.Nom; ++i) cu = std::min(cu, std::abs(-0.5 * g.Lu + (i + 1) * d.du));
    int nv = 0, nu = 0;
    double amass = 0.0;
    for (int j = 0; j < g.Nmu; ++j) {
        const double dvj = std::abs(-0.5 * g.Lv + (j + 1) * d.dv) - cv;
        if (std::abs(dvj) < 0.25 * d.dv) ++nv;
        for (int i = 0; i < g.Nom; ++i) {
            const double dui = std::abs(-0.5 * g.Lu + (i + 1) * d.du) - cu;
            amass += std::exp(-p.a_2 * (dvj * dvj + dui * dui)) * d.du * d.dv;
        }
    }
    for (int i = 0; i < g.Nom; ++i)
        if (std::abs(std::abs(-0.5 * g.Lu + (i + 1) * d.du) - cu) < 0.25 * d.du) ++nu;
    if (p.discrete_angle_delta) amass = 1.0;

    auto edge_weight = [&](int k, int last) { return p.trapezoidal_yz && (k == 0 || k == last) ? 0.5 : 1.0; };
    double smass = 0.0;
    for (int j = 0; j <= g.Nz; ++j) {
        const double z = -0.5 * g.Lz + j * d.dz;
        for (int i = 0; i <= g.Ny; ++i) {
            const double y = -0.5 * g.Ly + i * d.dy;
            smass += edge_weight(i, g.Ny) * edge_weight(j, g.Nz) * std::exp(-p.a_1 * (y * y + z * z));
        }
    }
    smass *= d.dy * d.dz;

    auto gauss = [&](int ek) {
        const double E = (ek + 0.5) * d.dg + 1.0;
        return 1.0 / std::sqrt(2.0 * kPi) / p.sig_E *
               std::exp(-std::pow((E - p.beam_energy) / p.sig_E / std::sqrt(2.0), 2));
    };
    b.energy.resize(g.Ng);
    double emass = 0.0;
    for (int ek = 0; ek < g.Ng; ++ek) {
        b.energy[ek] = gauss(ek);
        emass += b.energy[ek];
    }
    emass *= d.dg;

    const double total = smass * amass * emass;
    if (!(total > 0.0) || !std::isfinite(total))
        throw Error(BFPN_E_INVALID, "Initial beam has invalid discrete phase-space mass");
    b.spatial_mass = smass;
    b.angular_mass = amass;
    b.energy_mass = emass;
    b.scale = p.normalize_initial_mass ? 1.0 / total : 1.0;

    b.space_angle.assign(static_cast<size_t>(nyz) * nang, 0.0);
    for (int m = 0; m < nang; ++m) {
        const double d1 = std::abs(om[2 * m]) - cv;
        const double d2 = std::abs(om[2 * m + 1]) - cu;
        double ang = std::exp(-p.a_2 * (d1 * d1 + d2 * d2));
        if (p.discrete_angle_delta)
            ang = (std::abs(d1) < 0.25 * d.dv && std::abs(d2) < 0.25 * d.du)
                      ? 1.0 / (nv * nu * d.du * d.dv)
                      : 0.0;
        for (int j = 0; j <= g.Nz; ++j)
            for (int i = 0; i <= g.Ny; ++i) {
                const double y = -0.5 * g.Ly + i * d.dy;
                const double z = -0.5 * g.Lz + j * d.dz;
                const size_t s = static_cast<size_t>(j) * (g.Ny + 1) + i;
                b.space_angle[s * nang + m] = b.scale * std::exp(-p.a_1 * (y * y + z * z)) * ang;
            }
    }
    return b;
}

// Host precompute of the per-group eq15 sweep coefficients
// (src/energy_solver.cu:79-86, 114-150), folded into linear forms.  With
//   up = P_hi - R_hi, cur = P - R, strag = a_hi P_hi + a_lo P_lo - a_g P,
//   common = c_up up - c_cur cur - sig/2 P + P/h + strag + Q,
//   rhs = (1/h - sig/2) R + k_up up + k_cur cur - k_F P - k_f R - A common + H Delta,
//   f = rhs / coe,   F = B f + common / den_F + Cc Delta
// (the reference's sequence of operations), f and F are linear in
// x = (P_hi, R_hi, P, R, P_lo, Q, Delta); the sweep evaluates the two forms.
std::vector<EnergyCoef> energy_coefs(const std::vector<double>& S, const std::vector<double>& sigma,
                                     const double* T, bool straggling, double h, double dg,
                                     double rho, int Ng) {
    std::vector<EnergyCoef> out(Ng);
    const double dg2 = std::max(dg * dg, kEps);
    for (int g = 0; g < Ng; ++g) {
        const double sig = std::max(sigma[g], 0.0);
        const double Slo = std::max(S[g], 0.0);
        const double Shi = std::max(S[g + 1], 0.0);
        const bool st = straggling && T != nullptr;
        const double Tg = st ? std::max(T[g], 0.0) : 0.0;
        const double Thi = st && g + 1 < Ng ? std::max(T[g + 1], 0.0) : 0.0;
        const double Tlo = st && g > 0 ? std::max(T[g - 1], 0.0) : 0.0;
        const double inv_h = 1.0 / h, half_sig = 0.5 * sig;
        const double c_up = 0.5 * rho * Shi / dg, c_cur = 0.5 * rho * Slo / dg;
        const double inv_F = 1.0 / std::max(1.0 / h + 0.5 * rho * Slo / dg + 0.5 * sig, kEps);
        const double A = (1.5 * rho * Shi / dg) * inv_F;
        const double B = (0.5 * rho * Slo / dg) * inv_F;
        const double Cc = (0.5 * rho * Shi / dg) * inv_F;
        const double k_up = 1.5 * rho * Shi / dg, k_cur = 1.5 * rho * Slo / dg;
        const double k_F = 1.5 * rho * (Slo + Shi) / dg, k_f = 0.5 * rho * (Shi - Slo) / dg;
        const double H = k_up - A * 0.5 * rho * Shi / dg;
        const double coe = 1.0 / h + 0.5 * sig + 1.5 * rho * Slo / dg + 0.5 * rho * (Shi - Slo) / dg +
                           A * 0.5 * rho * Slo / dg;
        const double inv_coe = 1.0 / std::max(coe, kEps);
        const double a_hi = 0.5 * rho * Thi / dg2, a_lo = 0.5 * rho * Tlo / dg2, a_g = rho * Tg / dg2;
        // common = sum cm[i] x[i] (no Delta term)
        const double cm[6] = {c_up + a_hi, -c_up, inv_h - half_sig - c_cur - a_g, c_cur, a_lo, 1.0};
        // rhs = sum r[i] x[i] + H Delta
        const double r[6] = {k_up - A * cm[0],
                             -k_up - A * cm[1],
                             (k_cur - k_F) - A * cm[2],
                             (inv_h - half_sig - k_cur - k_f) - A * cm[3],
                             -A * cm[4],
                             -A * cm[5]};
        EnergyCoef& c = out[g];
        for (int i = 0; i < 6; ++i) {
            c.af[i] = inv_coe * r[i];
            c.aF[i] = B * c.af[i] + inv_F * cm[i];
        }
        c.af[6] = inv_coe * H;
        c.aF[6] = B * c.af[6] + Cc;
        c.w1 = c.w2 = 0.0;
    }
    return out;
}

// Eq.24 integrand per group (src/bfpm_solver.cu:279-287): S unclamped,
// sigma clamped at 0.
std::vector<UpwindCoef> upwind_coefs(const std::vector<double>& S, const std::vector<double>& sigma,
                                     const double* T, bool straggling, double h, double dg,
                                     double rho, int Ng) {
    constexpr double kEps = 1e-10;  // Numerics::EPSILON, include/config.hpp:39
    const double dg2 = std::max(dg * dg, kEps);
    std::vector<UpwindCoef> c(Ng);
    for (int g = 0; g < Ng; ++g) {
        const double sig = std::max(sigma[g], 0.0);
        const double S_lo = std::max(S[g], 0.0), S_hi = std::max(S[g + 1], 0.0);
        const double T_g = (straggling && T) ? std::max(T[g], 0.0) : 0.0;
        c[g].h = h;
        c[g].c_hi = rho * S_hi / dg;
        c[g].denom = std::max(1.0 / h + rho * S_lo / dg + sig, kEps);
        c[g].strag = T_g > 0.0 ? 0.5 * h * T_g / dg2 : 0.0;
    }
    return c;
}

std::vector<DoseCoef> dose_coefs(const std::vector<double>& S, const std::vector<double>& sigma,
                                 const std::vector<double>& en, double dg, bool catastrophic, int Ng) {
    std::vector<DoseCoef> w(Ng);
    for (int g = 0; g < Ng; ++g) {
        w[g].w1 = 0.5 * (S[g] + S[g + 1]) * dg;
        w[g].w2 = (S[g + 1] - S[g]) / 6.0 * dg;
        if (catastrophic) {
            const double sig = std::max(sigma[g], 0.0);
            w[g].w1 += 0.5 * (en[g] + en[g + 1]) * sig * dg;
            w[g].w2 += (en[g + 1] - en[g]) * sig * dg / 6.0;
        }
    }
    return w;
}

// CN ratio on the DST-I eigenvalues times the DST normalisation,
// src/angle_diffusion.cu:7-34 and 451-458
std::vector<double> angle_factors(const double* sig_tr, int Ng, int Nmu, int Nom, double du,
                                  double dv, double dt, double rho) {
    const int na = Nmu * Nom;
    const double norm = 4.0 / ((Nom + 1.0) * (Nmu + 1.0));
    std::vector<double> f(static_cast<size_t>(Ng) * na);
    for (int g = 0; g < Ng; ++g) {
        const double D = 0.5 * rho * sig_tr[g];
        for (int nu = 0; nu < Nmu; ++nu) {
            const double lv = (2.0 - 2.0 * std::cos(kPi * (nu + 1) / (Nmu + 1))) / (dv * dv);
            for (int mu = 0; mu < Nom; ++mu) {
                const double lu = (2.0 - 2.0 * std::cos(kPi * (mu + 1) / (Nom + 1))) / (du * du);
                const double lam = D * (lu + lv);
                f[static_cast<size_t>(g) * na + nu * Nom + mu] =
                    (2.0 / dt - lam) / (2.0 / dt + lam) * norm;
            }
        }
    }
    return f;
}

// S[k][j] = sin(pi (k+1)(j+1)/(n+1)), src/angle_diffusion.cu:395-408
std::vector<double> sine_basis(int n) {
    std::vector<double> s(static_cast<size_t>(n) * n);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j) s[static_cast<size_t>(k) * n + j] = std::sin(kPi * (k + 1) * (j + 1) / (n + 1));
    return s;
}

}  // namespace bfpn
