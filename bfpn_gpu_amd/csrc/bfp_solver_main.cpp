// bfp_solver — drop-in command line front end of the MI355X BFPn solver.
//
// Reference: src/main.cu:44-356.  Same executable name, flags, defaults,
// silent ignoring of unknown flags, positional Nx anywhere on the line,
// material presets that `--rho` later overrides, `Error: <what>` + rc 1 on
// failure, and the same output files in the working directory:
//   idd_output.txt   "# x_cm value" then (i+1)*dt and the IDD, 17 digits
//   dose_output.txt  same format, dg * sum psi^2 (primary)
// (BFPnSolver::saveResults, src/bfpm_solver.cu:3808-3815), and on request
//   energy_moments.txt       --save-energy-moments: Eq.25 moments per step and
//                            group (saveEnergyMoments, 3817-3835)
//   <prefix>_reqD_xX.txt     --spot-depths D1,D2,..: Y-Z dose plane after the
//                            first step whose depth reaches D - dt/2
//                            (saveSpotPlane, 3587-3612; solve loop 2244-2249)
//
// Opt-in additions: --precision fp32|fp64, --lg <MeV span>.
// The out-of-core streaming flags are accepted and run in core (the whole
// phase space stays resident in HBM).

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "bfpn.h"

namespace {

void print_devices() {
    int n = 0;
    if (bfpn_device_count(&n) != BFPN_OK) {
        std::cout << "No HIP device: " << bfpn_last_error() << std::endl;
        return;
    }
    std::cout << "Found " << n << " HIP device(s)" << std::endl;
    for (int i = 0; i < n; ++i) {
        char name[256] = {0};
        size_t total = 0;
        int cus = 0;
        if (bfpn_device_info(i, name, sizeof(name), &total, &cus) != BFPN_OK) continue;
        std::cout << "\nDevice " << i << ": " << name << std::endl;
        std::cout << "  Total Memory: " << total / (1024 * 1024) << " MB" << std::endl;
        std::cout << "  Compute Units: " << cus << std::endl;
    }
    std::cout << std::endl;
}

void save_series(const std::vector<double>& v, double dt, const std::string& path) {
    std::ofstream out(path);
    out << std::setprecision(17);
    out << "# x_cm value\n";
    for (size_t i = 0; i < v.size(); ++i) out << (i + 1) * dt << " " << v[i] << "\n";
}

// Comma-separated depths, empty items skipped (parseDepthList, src/main.cu:12-22).
std::vector<double> parse_depth_list(const std::string& text) {
    std::vector<double> values;
    std::stringstream ss(text);
    std::string item;
    while (std::getline(ss, item, ','))
        if (!item.empty()) values.push_back(std::stod(item));
    return values;
}

// BFPnSolver::saveSpotPlane (src/bfpm_solver.cu:3587-3612): file name and
// layout as the reference writes them (default stream precision).
void save_spot_plane(const std::vector<double>& plane, const std::string& prefix, double requested,
                     double actual, int ny, int nz, double ly, double lz, double dy, double dz) {
    if (prefix.empty()) return;
    std::ostringstream name;
    name << prefix << "_req" << std::fixed << std::setprecision(2) << requested << "_x"
         << std::fixed << std::setprecision(2) << actual << ".txt";
    std::ofstream out(name.str());
    out << "# requested_x_cm " << requested << "\n";
    out << "# actual_x_cm " << actual << "\n";
    out << "# y_cm z_cm dose\n";
    for (int j = 0; j <= nz; ++j) {
        const double z = -0.5 * lz + j * dz;
        for (int i = 0; i <= ny; ++i) {
            const double y = -0.5 * ly + i * dy;
            out << y << " " << z << " " << plane[static_cast<size_t>(j) * (ny + 1) + i] << "\n";
        }
        out << "\n";
    }
    std::cout << "Spot dose plane saved to " << name.str() << std::endl;
}

// BFPnSolver::saveEnergyMoments (src/bfpm_solver.cu:3817-3835).
void save_energy_moments(const std::vector<double>& values, int ng, double dt,
                         const std::string& path) {
    std::ofstream out(path);
    out << "# step x_cm g psi1_integral psi2_integral\n";
    const size_t per_step = static_cast<size_t>(ng) * 2;
    if (per_step == 0) return;
    const size_t steps = values.size() / per_step;
    for (size_t step = 0; step < steps; ++step) {
        const double x = (step + 1) * dt;
        const size_t base = step * per_step;
        for (int g = 0; g < ng; ++g)
            out << (step + 1) << " " << x << " " << g << " " << values[base + 2 * g] << " "
                << values[base + 2 * g + 1] << "\n";
    }
}

double preset_density(const std::string& material) {
    if (material == "water") return 1.0;
    if (material == "bone") return 1.757;
    if (material == "air") return 0.001205;
    throw std::runtime_error("Unknown material: " + material);
}

// --layer x_end:material[:rho[:data]]
struct LayerArg {
    double x_end;
    std::string material, data;
    double rho;
};

LayerArg parse_layer(const std::string& text) {
    std::vector<std::string> f;
    std::stringstream ss(text);
    for (std::string item; std::getline(ss, item, ':');) f.push_back(item);
    if (f.size() < 2 || f.size() > 4)
        throw std::runtime_error("--layer expects x_end:material[:rho[:data]], got " + text);
    LayerArg l{std::stod(f[0]), f[1], "", 0.0};
    l.rho = f.size() > 2 && !f[2].empty() ? std::stod(f[2]) : preset_density(l.material);
    l.data = f.size() > 3 ? f[3] : "BFPn_CPU_Solver/" + l.material;
    return l;
}

// --material-map FILE (no counterpart in the reference): a text file of
//   material <name> [<rho> [<data>]]   one line per material, index = order
//   slab <x_end | inf>                 then Nz+1 rows of Ny+1 material indices
//                                      (row iz, column iy)
// '#' starts a comment.  Slabs in increasing depth; the last covers any depth.
struct MaterialMap {
    std::vector<LayerArg> materials;
    std::vector<double> slab_end;
    std::vector<unsigned char> map;  // slab-major, node s = iz (Ny+1) + iy
};

MaterialMap read_material_map(const std::string& path, int ny, int nz) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Cannot open material map " + path);
    MaterialMap m;
    const size_t nyz = static_cast<size_t>(ny + 1) * (nz + 1);
    size_t row_values = 0;  // indices read for the current slab
    for (std::string line; std::getline(in, line);) {
        line = line.substr(0, line.find('#'));
        std::istringstream ls(line);
        std::string head;
        if (!(ls >> head)) continue;
        if (head == "material") {
            LayerArg l{0.0, "", "", 0.0};
            if (!(ls >> l.material)) throw std::runtime_error("material line without a name in " + path);
            std::string rho;
            l.rho = (ls >> rho) ? std::stod(rho) : preset_density(l.material);
            if (!(ls >> l.data)) l.data = "BFPn_CPU_Solver/" + l.material;
            m.materials.push_back(l);
        } else if (head == "slab") {
            if (!m.slab_end.empty() && row_values != nyz)
                throw std::runtime_error("slab " + std::to_string(m.slab_end.size()) + " of " + path +
                                         " has " + std::to_string(row_values) + " indices, not " +
                                         std::to_string(nyz));
            std::string x;
            if (!(ls >> x)) throw std::runtime_error("slab line without a depth in " + path);
            m.slab_end.push_back(x == "inf" ? 1e300 : std::stod(x));
            row_values = 0;
        } else {
            if (m.slab_end.empty()) throw std::runtime_error("map values before a slab line in " + path);
            std::istringstream vs(line);
            for (long v; vs >> v;) {
                if (v < 0 || v >= static_cast<long>(m.materials.size()) || row_values == nyz)
                    throw std::runtime_error("bad map entry " + std::to_string(v) + " in " + path);
                m.map.push_back(static_cast<unsigned char>(v));
                ++row_values;
            }
        }
    }
    if (m.slab_end.empty() || row_values != nyz || m.materials.empty())
        throw std::runtime_error("material map " + path + " needs materials and complete slabs of " +
                                 std::to_string(nyz) + " indices");
    return m;
}

void usage(const char* argv0) {
    std::cout << "Usage: " << argv0 << " [options]\n"
              << "Options:\n"
              << "  <Nx>             Optional x grid points (default: 4000)\n"
              << "  --data <path>    Data file directory (default: ./data)\n"
              << "  --time <value>   Final simulation time (default: 40.0)\n"
              << "  --energy <MeV>   Incoming beam energy E0 (default: 230.0)\n"
              << "  --sigma-e <MeV>  Initial energy Gaussian sigma (default: 1.0)\n"
              << "  --material <name> Material density preset: water, bone, air\n"
              << "  --rho <value>    Override material density in g/cm^3\n"
              << "  --ny/--nz <n>    Y/Z grid cells (nodes are n+1)\n"
              << "  --ly/--lz <cm>   Y/Z domain width (default: 1)\n"
              << "  --lu/--lv <v>    U/V angular domain width (default: 1)\n"
              << "  --ng <n>         Energy groups (default: 500)\n"
              << "  --nmu/--nom <n>  Angular grid points (default: 11)\n"
              << "  --lite-memory    Primary psi^1 only, energy step only (low memory)\n"
              << "  --primary-only   Skip secondary proton evolution/source\n"
              << "  --energy-model <eq15|legacy> Energy update model (default: eq15)\n"
              << "  --eq15-straggling     Enable Eq.15 straggling terms\n"
              << "  --energy-only    Only advance the primary energy subsystem\n"
              << "  --no-transport   Skip the spatial transport subsystem (primary)\n"
              << "  --no-angle       Skip the angular diffusion subsystem (primary)\n"
              << "  --no-spatial-clipping Disable positivity clipping of primary psi^1\n"
              << "  --profile-steps  Print energy/transport/angle step timing summary\n"
              << "  --save-energy-moments Save paper Eq.25 energy-moment diagnostics\n"
              << "  --spot-depths <x,...> Save YZ dose planes near requested depths\n"
              << "  --spot-prefix <path> Prefix for spot plane files (default: spot)\n"
              << "  --discrete-angle-delta Quadrature-normalized mono-directional beam\n"
              << "  --normalize-initial-mass Normalize the discrete beam mass\n"
              << "  --trapezoidal-yz Trapezoidal y/z weights in normalization and IDD\n"
              << "  --dose-catastrophic-loss Include the Eq.24 E*sigma_c*psi term (default)\n"
              << "  --stopping-power-dose-only Omit the Eq.24 catastrophic-loss term\n"
              << "  --no-cross-section-calibration Use fitted sigma_c (default)\n"
              << "  --legacy-cross-section-calibration Apply the legacy sigma_c correction\n"
              << "  --sigma-yz <cm>  Initial transverse Gaussian sigma (default: 0.1)\n"
              << "  --layer <x_end>:<material>[:<rho>[:<data>]] Depth layer up to x_end cm\n"
              << "                   (repeat in increasing depth; the last extends to any depth;\n"
              << "                   rho defaults to the material preset, data to\n"
              << "                   BFPn_CPU_Solver/<material>)\n"
              << "  --material-map <file> Materials per y-z node and depth slab: lines\n"
              << "                   'material <name> [<rho> [<data>]]', then per slab\n"
              << "                   'slab <x_end|inf>' and Nz+1 rows of Ny+1 indices\n"
              << "  --streaming-full Accepted; the state always stays resident in HBM\n"
              << "  --precision <fp32|fp64> State precision (default: fp32)\n"
              << "  --lg <MeV>       Energy span of the group grid [1, 1+Lg] (default: 259)\n"
              << "  --test           Print device info and exit\n"
              << "  --help           Show this help\n";
}

}  // namespace

int main(int argc, char** argv) {
    std::string data_path = "./data";
    bool data_set = false;
    std::string material;
    double t_final = 40.0;
    bool test_mode = false;
    int nx = 4000, ny = 10, nz = 10, ng = 500, nmu = 11, nom = 11;
    double ly = 1.0, lz = 1.0, lu = 1.0, lv = 1.0, lg = 259.0;
    double energy = 230.0, sigma_e = 1.0, sigma_yz = 0.1, rho = 1.0, rho_e = 1.0;
    bool lite = false, streaming = false, legacy_energy = false, profile = false;
    int precision = BFPN_FP32;
    std::vector<double> spot_depths;
    std::string spot_prefix = "spot";
    std::vector<LayerArg> layers;
    std::string map_path;
    bfpn_phys phys{};
    phys.dose_catastrophic_loss = 1;

    try {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            const bool has = i + 1 < argc;
            if (a == "--data" && has) { data_path = argv[++i]; data_set = true; }
            else if (a == "--time" && has) t_final = std::stod(argv[++i]);
            else if (a == "--energy" && has) energy = std::stod(argv[++i]);
            else if (a == "--sigma-e" && has) sigma_e = std::stod(argv[++i]);
            else if (a == "--rho" && has) { rho = std::stod(argv[++i]); rho_e = rho; }
            else if (a == "--ny" && has) ny = std::stoi(argv[++i]);
            else if (a == "--nz" && has) nz = std::stoi(argv[++i]);
            else if (a == "--ly" && has) ly = std::stod(argv[++i]);
            else if (a == "--lz" && has) lz = std::stod(argv[++i]);
            else if (a == "--lu" && has) lu = std::stod(argv[++i]);
            else if (a == "--lv" && has) lv = std::stod(argv[++i]);
            else if (a == "--ng" && has) ng = std::stoi(argv[++i]);
            else if (a == "--nmu" && has) nmu = std::stoi(argv[++i]);
            else if (a == "--nom" && has) nom = std::stoi(argv[++i]);
            else if (a == "--lg" && has) lg = std::stod(argv[++i]);
            else if (a == "--material" && has) {
                material = argv[++i];
                if (material == "water") rho = rho_e = 1.0;
                else if (material == "bone") rho = rho_e = 1.757;
                else if (material == "air") rho = rho_e = 0.001205;
                else {
                    std::cerr << "Unknown material: " << material << std::endl;
                    return 1;
                }
            } else if (a == "--test") test_mode = true;
            else if (a == "--lite-memory") lite = true;
            else if (a == "--streaming-full") streaming = true;
            else if (a == "--primary-only") phys.primary_only = 1;
            else if (a == "--energy-model" && has) {
                const std::string m = argv[++i];
                if (m == "legacy") legacy_energy = true;
                else if (m == "eq15") legacy_energy = false;
                else {
                    std::cerr << "Unknown energy model: " << m << std::endl;
                    return 1;
                }
            } else if (a == "--eq15-straggling") phys.eq15_straggling = 1;
            else if (a == "--save-energy-moments") phys.save_energy_moments = 1;
            else if (a == "--energy-only") phys.energy_only = 1;
            else if (a == "--no-transport") phys.no_transport = 1;
            else if (a == "--no-angle") phys.no_angle = 1;
            else if (a == "--no-spatial-clipping") phys.no_spatial_clipping = 1;
            else if (a == "--profile-steps") profile = true;
            else if (a == "--discrete-angle-delta") phys.discrete_angle_delta = 1;
            else if (a == "--normalize-initial-mass") phys.normalize_initial_mass = 1;
            else if (a == "--trapezoidal-yz") phys.trapezoidal_yz = 1;
            else if (a == "--dose-catastrophic-loss") phys.dose_catastrophic_loss = 1;
            else if (a == "--stopping-power-dose-only") phys.dose_catastrophic_loss = 0;
            else if (a == "--no-cross-section-calibration") phys.calibrate_cross_sections = 0;
            else if (a == "--legacy-cross-section-calibration") phys.calibrate_cross_sections = 1;
            else if (a == "--sigma-yz" && has) sigma_yz = std::stod(argv[++i]);
            else if (a == "--precision" && has) {
                const std::string p = argv[++i];
                if (p == "fp32") precision = BFPN_FP32;
                else if (p == "fp64") precision = BFPN_FP64;
                else {
                    std::cerr << "Unknown precision: " << p << std::endl;
                    return 1;
                }
            } else if (a == "--spot-depths" && has) spot_depths = parse_depth_list(argv[++i]);
            else if (a == "--layer" && has) layers.push_back(parse_layer(argv[++i]));
            else if (a == "--material-map" && has) map_path = argv[++i];
            else if (a == "--spot-prefix" && has) spot_prefix = argv[++i];
            else if ((a == "--lane-chunk" || a == "--energy-chunk" || a == "--stream-dir" ||
                      a == "--idd-stride") && has)
                ++i;  // streaming knobs: accepted, no effect in core
            else if (a == "--help") {
                usage(argv[0]);
                return 0;
            } else if (!a.empty() && a[0] != '-')
                nx = std::stoi(a);
        }
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }

    if (!data_set && !material.empty()) data_path = "BFPn_CPU_Solver/" + material;
    if (lite && streaming) {
        std::cerr << "--lite-memory and --streaming-full are mutually exclusive" << std::endl;
        return 1;
    }
    print_devices();
    if (test_mode) {
        std::cout << "Running in test mode..." << std::endl;
        return 0;
    }
    phys.legacy_energy = legacy_energy ? 1 : 0;
    phys.lite_memory = lite ? 1 : 0;

    bfpn_grid grid{nx, ny, nz, ng, nmu, nom, 40.0, ly, lz, lu, lv, lg};
    phys.sig_E = sigma_e;
    phys.a_1 = 1.0 / 2.0 / sigma_yz / sigma_yz;
    phys.a_2 = 1.0 / 2.0 / 1e-6 / 1e-6;
    phys.beam_energy = energy;
    phys.density = rho;
    phys.energy_density = rho_e;
    std::snprintf(phys.material_name, sizeof(phys.material_name), "%s", material.c_str());

    std::cout << "Domain widths: Lx=" << grid.Lx << ", Ly=" << ly << ", Lz=" << lz << ", Lu=" << lu
              << ", Lv=" << lv << std::endl;
    std::cout << "Initial beam sigmas: yz=" << sigma_yz << " cm, E=" << sigma_e
              << " MeV, angular=1e-6" << std::endl;
    std::cout << "Initial quadrature: angle="
              << (phys.discrete_angle_delta ? "discrete-delta" : "sampled-gaussian")
              << ", mass normalization=" << (phys.normalize_initial_mass ? "on" : "off")
              << ", y/z trapezoidal weights=" << (phys.trapezoidal_yz ? "on" : "off") << std::endl;
    std::cout << "Dose catastrophic-loss term: "
              << (phys.dose_catastrophic_loss ? "enabled" : "disabled") << std::endl;
    std::cout << "Cross-section calibration: "
              << (phys.calibrate_cross_sections ? "legacy empirical correction" : "disabled")
              << std::endl;
    std::cout << "Spatial positivity clipping: "
              << (phys.no_spatial_clipping ? "disabled" : "enabled") << std::endl;
    std::cout << "State precision: " << (precision == BFPN_FP64 ? "fp64" : "fp32")
              << (streaming ? " (streaming flags ignored: in-core)" : "") << std::endl;

    bfpn_ctx* ctx = nullptr;
    std::vector<bfpn_layer> spec;
    for (const auto& l : layers) {
        spec.push_back({l.x_end, l.material.c_str(), l.rho, l.data.c_str()});
        std::cout << "Layer: to x=" << l.x_end << " cm, " << l.material << ", rho=" << l.rho
                  << ", data " << l.data << std::endl;
    }
    bfpn_status created;
    if (!map_path.empty()) {
        if (!spec.empty()) {
            std::cerr << "Error: --layer and --material-map exclude each other" << std::endl;
            return 1;
        }
        MaterialMap mm;
        try {
            mm = read_material_map(map_path, ny, nz);
        } catch (const std::exception& e) {
            std::cerr << "Error: " << e.what() << std::endl;
            return 1;
        }
        std::vector<bfpn_layer> mats;
        for (const auto& l : mm.materials) {
            mats.push_back({0.0, l.material.c_str(), l.rho, l.data.c_str()});
            std::cout << "Map material " << mats.size() - 1 << ": " << l.material << ", rho=" << l.rho
                      << ", data " << l.data << std::endl;
        }
        std::cout << "Material map: " << mm.slab_end.size() << " depth slab(s) from " << map_path
                  << std::endl;
        created = bfpn_create_mapped(&grid, &phys, mats.data(), static_cast<int>(mats.size()),
                                     mm.slab_end.data(), static_cast<int>(mm.slab_end.size()),
                                     mm.map.data(), precision, &ctx);
    } else {
        created = spec.empty()
                      ? bfpn_create(&grid, &phys, data_path.c_str(), precision, &ctx)
                      : bfpn_create_layered(&grid, &phys, spec.data(), static_cast<int>(spec.size()),
                                            precision, &ctx);
    }
    if (created != BFPN_OK) {
        std::cerr << "Error: " << bfpn_last_error() << std::endl;
        return 1;
    }
    bfpn_deltas d{};
    bfpn_get_deltas(ctx, &d);
    const int steps = bfpn_num_steps(ctx, t_final);
    std::cout << "Memory: " << std::fixed << std::setprecision(2)
              << bfpn_device_bytes(ctx) / (1024.0 * 1024.0 * 1024.0) << " GiB resident"
              << std::defaultfloat << std::setprecision(6) << std::endl;
    std::cout << "Starting BFPn solver..." << std::endl;
    std::cout << "Grid: " << nx << "x" << ny << "x" << nz << ", Energy: " << ng << ", Angle: " << nmu
              << "x" << nom << std::endl;
    if (profile) bfpn_set_profiling(ctx, 1);

    std::vector<double> idd(steps), proxy(steps);
    // depth after each step, accumulated like the reference's t += dt
    std::vector<double> depth(static_cast<size_t>(steps) + 1, 0.0);
    for (int k = 1; k <= steps; ++k) depth[k] = depth[k - 1] + d.dt;
    // a spot plane is taken after the first step whose depth + dt/2 reaches
    // the requested depth (src/bfpm_solver.cu:2244-2249)
    std::vector<int> spot_step(spot_depths.size(), -1);
    for (size_t i = 0; i < spot_depths.size(); ++i)
        for (int k = 1; k <= steps; ++k)
            if (depth[k] + 0.5 * d.dt >= spot_depths[i]) {
                spot_step[i] = k;
                break;
            }
    const bool moments = phys.save_energy_moments != 0;
    std::vector<double> moment_history;
    std::vector<double> plane(static_cast<size_t>(ny + 1) * (nz + 1));
    auto fail = [&] {
        std::cerr << "Error: " << bfpn_last_error() << std::endl;
        bfpn_destroy(ctx);
        return 1;
    };
    double t = 0.0;
    int done = 0;
    while (done < steps) {
        // advance to the next progress line, diagnostic step or the end (a
        // call pairs the energy half steps inside it; its last step ends
        // with the full state, which the diagnostics read)
        int next = std::min(done + 100 - done % 100, steps);
        if (moments) next = done + 1;
        for (int k : spot_step)
            if (k > done && k < next) next = k;
        const int chunk = next - done;
        if (bfpn_advance(ctx, chunk) != BFPN_OK) return fail();
        for (int k = 0; k < chunk; ++k) t += d.dt;
        done += chunk;
        if (moments) {
            std::vector<double> m(2 * static_cast<size_t>(ng));
            if (bfpn_energy_moments(ctx, m.data(), 2 * ng) != BFPN_OK) return fail();
            moment_history.insert(moment_history.end(), m.begin(), m.end());
        }
        for (size_t i = 0; i < spot_depths.size(); ++i)
            if (spot_step[i] == done) {
                if (bfpn_spot_plane(ctx, plane.data(), plane.size()) != BFPN_OK) return fail();
                save_spot_plane(plane, spot_prefix, spot_depths[i], t, ny, nz, ly, lz, d.dy, d.dz);
            }
        if (done % 100 == 0) {
            int n = 0;
            if (bfpn_history(ctx, idd.data(), proxy.data(), steps, &n) != BFPN_OK) {
                std::cerr << "Error: " << bfpn_last_error() << std::endl;
                bfpn_destroy(ctx);
                return 1;
            }
            std::cout << "Step " << done << ", t = " << t << "/" << t_final
                      << ", IDD = " << idd[done - 1] << std::endl;
        }
    }
    int n = 0;
    if (bfpn_history(ctx, idd.data(), proxy.data(), steps, &n) != BFPN_OK) {
        std::cerr << "Error: " << bfpn_last_error() << std::endl;
        bfpn_destroy(ctx);
        return 1;
    }
    save_series(idd, d.dt, "idd_output.txt");
    save_series(proxy, d.dt, "dose_output.txt");
    if (moments) save_energy_moments(moment_history, ng, d.dt, "energy_moments.txt");
    std::cout << "Solver completed. Paper-style IDD saved to idd_output.txt" << std::endl;
    std::cout << "Diagnostic scalar proxy saved to dose_output.txt" << std::endl;
    if (profile) {
        double ph[5];
        bfpn_phase_stats(ctx, ph, nullptr, nullptr);
        const double tot = ph[0] + ph[1] + ph[2] + ph[3] + ph[4];
        const char* names[5] = {"energy", "transport", "angle", "source", "diagnostics"};
        std::cout << "Step profile (" << steps << " steps, device time):" << std::endl;
        for (int k = 0; k < 5; ++k)
            std::cout << "  " << names[k] << ": " << ph[k] << " s ("
                      << (tot > 0 ? 100.0 * ph[k] / tot : 0.0) << "%)" << std::endl;
    }
    bfpn_destroy(ctx);
    return 0;
}
