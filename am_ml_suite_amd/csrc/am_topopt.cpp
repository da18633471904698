// am_topopt — the reference's standalone optimizer program (cpp_engine/src/am_topopt.cpp) on
// the GPU engine: the BRK-A-01 bracket at a given resolution, optimised through the drop-in
// am::topopt::TopologyOptimizer, with the same command line (:453-511) and the same output
// files: density_field.npy (NumPy v1.0, C order, :344-396), metadata.json (:401-447) and
// history.csv (:622-629).
//
// Device options on top of the reference's: --precision fp64|mixed_fp32|mixed_bf16,
// --preconditioner jacobi|multigrid, --tol (PCG relative tolerance), --index
// reference|consistent (density index; the reference's is the default), and --dump-domain
// (print the set-up as JSON and exit, without a GPU).
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/am/problem.hpp"
#include "../../include/am/topology_optimizer.hpp"
#include "../../include/amt.h"

namespace fs = std::filesystem;

namespace {

struct Options {
  am::BracketConfig bracket;
  double volume_fraction = 0.25;
  int max_iterations = 100;
  double tolerance = 0.01;  // convergence on max |drho|
  double filter_radius = 3.5;
  double E_min = 1e-4;      // also the void density and rho_min, as the reference uses it
  double penalty = 3.0;
  double move_limit = 0.1;
  std::string output_dir = "./output";
  bool verbose = true;
  // device options
  std::string precision = "fp64", preconditioner = "jacobi", index = "reference";
  double pcg_tol = 1e-8;
  bool dump_domain = false;
};

void usage(const char* prog) {
  std::cout << "Usage: " << prog << " [options]\n\n"
            << "Options:\n"
            << "  --resolution <mm>     Grid resolution (default: 2.0)\n"
            << "  --volume <fraction>   Target volume fraction (default: 0.25)\n"
            << "  --iterations <n>      Max iterations (default: 100)\n"
            << "  --filter <voxels>     Filter radius in voxels (default: 3.5)\n"
            << "  --move <limit>        Max density change per iter (default: 0.1)\n"
            << "  --emin <value>        Minimum modulus ratio (default: 1e-4)\n"
            << "  --output <dir>        Output directory (default: ./output)\n"
            << "  --output_dir <dir>    Alias for --output\n"
            << "  -o <dir>              Alias for --output\n"
            << "  --quiet               Disable verbose output\n"
            << "  --help                Show this help\n"
            << "Device options:\n"
            << "  --precision <p>       fp64 | mixed_fp32 | mixed_bf16 (default: fp64)\n"
            << "  --preconditioner <p>  jacobi | multigrid (default: jacobi)\n"
            << "  --tol <x>             PCG relative residual tolerance (default: 1e-8)\n"
            << "  --index <i>           reference | consistent density index (default: reference)\n"
            << "  --dump-domain         Print the problem set-up as JSON and exit\n";
}

// 0: run, 1: exit with error, 2: exit after help
int parse(int argc, char** argv, Options& o) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](void) -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
    const char* v = nullptr;
    if (a == "--help" || a == "-h") {
      usage(argv[0]);
      return 2;
    } else if (a == "--quiet") {
      o.verbose = false;
    } else if (a == "--dump-domain") {
      o.dump_domain = true;
    } else if ((a == "--resolution" || a == "--volume" || a == "--iterations" ||
                a == "--filter" || a == "--move" || a == "--emin" || a == "--output" ||
                a == "--output_dir" || a == "-o" || a == "--precision" ||
                a == "--preconditioner" || a == "--tol" || a == "--index") &&
               (v = next())) {
      try {
        if (a == "--resolution") o.bracket.resolution_mm = std::stod(v);
        else if (a == "--volume") o.volume_fraction = std::stod(v);
        else if (a == "--iterations") o.max_iterations = std::stoi(v);
        else if (a == "--filter") o.filter_radius = std::stod(v);
        else if (a == "--move") o.move_limit = std::stod(v);
        else if (a == "--emin") o.E_min = std::stod(v);
        else if (a == "--precision") o.precision = v;
        else if (a == "--preconditioner") o.preconditioner = v;
        else if (a == "--tol") o.pcg_tol = std::stod(v);
        else if (a == "--index") o.index = v;
        else o.output_dir = v;
      } catch (const std::exception&) {
        std::cerr << "Bad value for " << a << ": " << v << std::endl;
        return 1;
      }
    } else {
      std::cerr << "Unknown option: " << a << std::endl;
      usage(argv[0]);
      return 1;
    }
  }
  return 0;
}

// NumPy v1.0 file of the x-fastest field in C order (shape nx, ny, nz); the header is padded
// with spaces so magic + version + length + header is a multiple of 64 bytes.
void save_npy(const std::string& path, const std::vector<double>& slot, int nx, int ny, int nz) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("Cannot open file: " + path);
  std::ostringstream h;
  h << "{'descr': '<f8', 'fortran_order': False, 'shape': (" << nx << ", " << ny << ", " << nz
    << "), }";
  std::string dict = h.str();
  const size_t pad = (64 - (10 + dict.size() + 1) % 64) % 64;
  dict.append(pad, ' ');
  dict.push_back('\n');
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  const uint16_t len = (uint16_t)dict.size();  // little-endian host
  f.write((const char*)magic, 8);
  f.write((const char*)&len, 2);
  f.write(dict.data(), (std::streamsize)dict.size());
  std::vector<double> c((size_t)nx * ny * nz);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k)
        c[((size_t)i * ny + j) * nz + k] = slot[((size_t)k * ny + j) * nx + i];
  f.write((const char*)c.data(), (std::streamsize)(c.size() * sizeof(double)));
  std::cout << "Saved: " << path << " (" << nx << "x" << ny << "x" << nz << ")" << std::endl;
}

void save_metadata(const std::string& path, const Options& o, const am::VoxelDomain& d,
                   const am::topopt::OptimizationResult& r, double elapsed) {
  std::ofstream f(path);
  if (!f) throw std::runtime_error("Cannot open file: " + path);
  char dev[128] = {0};
  int n_cu = 0;
  amt_device_info(dev, sizeof(dev), &n_cu);
  f << "{\n"
    << "  \"case\": \"BRK-A-01\",\n"
    << "  \"mode\": \"cpp_standalone\",\n"
    << "  \"provenance\": \"cpp_hpc_engine\",\n"
    << "  \"resolution_mm\": " << o.bracket.resolution_mm << ",\n"
    << "  \"domain_shape\": [" << d.nx << ", " << d.ny << ", " << d.nz << "],\n"
    << "  \"n_elements\": " << d.grid.size() << ",\n"
    << "  \"n_design\": " << d.count(am::ElementType::Design) << ",\n"
    << "  \"n_fixed\": " << d.count(am::ElementType::Fixed) << ",\n"
    << "  \"n_void\": " << d.count(am::ElementType::Void) << ",\n"
    << "  \"volume_fraction_target\": " << o.volume_fraction << ",\n";
  if (!r.volume_history.empty())
    f << "  \"volume_fraction_final\": " << r.volume_history.back() << ",\n";
  if (!r.compliance_history.empty())
    f << "  \"compliance_final\": " << r.compliance_history.back() << ",\n";
  f << "  \"iterations\": " << r.iterations << ",\n"
    << "  \"converged\": " << (r.converged ? "true" : "false") << ",\n"
    << "  \"elapsed_time_s\": " << elapsed << ",\n"
    << "  \"openmp_threads\": " << n_cu << ",\n"
    << "  \"device\": \"" << dev << "\",\n"
    << "  \"simp_config\": {\n"
    << "    \"filter_radius\": " << o.filter_radius << ",\n"
    << "    \"E_min\": " << o.E_min << ",\n"
    << "    \"penalty\": " << o.penalty << ",\n"
    << "    \"move_limit\": " << o.move_limit << "\n"
    << "  }\n"
    << "}\n";
  std::cout << "Saved: " << path << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  const int pr = parse(argc, argv, o);
  if (pr) return 1;
  if (!(o.bracket.resolution_mm > 0)) {
    std::cerr << "Resolution must be positive" << std::endl;
    return 1;
  }
  const am::VoxelDomain dom = am::bracketDomain(o.bracket);
  const am::LoadCase lc = am::bracketLoadCase(dom, o.bracket);
  if (o.dump_domain) {
    double fz = 0.0;
    for (const auto& pl : lc.point_loads) fz += pl.second[2];
    std::cout << "{\"domain_shape\": [" << dom.nx << ", " << dom.ny << ", " << dom.nz
              << "], \"n_design\": " << dom.count(am::ElementType::Design)
              << ", \"n_fixed\": " << dom.count(am::ElementType::Fixed)
              << ", \"n_void\": " << dom.count(am::ElementType::Void)
              << ", \"n_dofs\": " << lc.n_dofs << ", \"n_load_nodes\": " << lc.point_loads.size()
              << ", \"first_load_node\": " << (lc.point_loads.empty() ? -1 : lc.point_loads[0].first)
              << ", \"total_fz\": " << fz << ", \"n_fixed_dofs\": " << lc.constrained.size()
              << "}" << std::endl;
    return 0;
  }
  std::cout << "================================================================\n"
            << "AM Topology Optimization - BRK-A-01 Bracket (GPU engine)\n"
            << "================================================================\n\n"
            << "Grid: " << dom.nx << " x " << dom.ny << " x " << dom.nz << " = " << dom.grid.size()
            << " elements\n  Design: " << dom.count(am::ElementType::Design)
            << "\n  Fixed:  " << dom.count(am::ElementType::Fixed)
            << "\n  Void:   " << dom.count(am::ElementType::Void) << std::endl;

  am::topopt::SIMPConfig simp;  // am_topopt.cpp:569-577
  simp.volume_fraction = o.volume_fraction;
  simp.max_iterations = o.max_iterations;
  simp.convergence_tol = o.tolerance;
  simp.filter_radius = o.filter_radius;
  simp.rho_min = o.E_min;
  simp.penalty = o.penalty;
  simp.move_limit = o.move_limit;
  simp.verbose = o.verbose;
  simp.solver_tolerance = o.pcg_tol;
  simp.solver_max_iterations = 1000000;
  if (o.index == "consistent") simp.density_index = am::topopt::DensityIndex::Consistent;
  else if (o.index == "reference") simp.density_index = am::topopt::DensityIndex::Reference;
  else { std::cerr << "Unknown --index " << o.index << std::endl; return 1; }
  if (o.precision == "fp64") simp.precision = am::topopt::Precision::FP64;
  else if (o.precision == "mixed_fp32") simp.precision = am::topopt::Precision::MixedFP32;
  else if (o.precision == "mixed_bf16") simp.precision = am::topopt::Precision::MixedBF16;
  else { std::cerr << "Unknown --precision " << o.precision << std::endl; return 1; }
  if (o.preconditioner == "jacobi") simp.preconditioner = am::topopt::Preconditioner::Jacobi;
  else if (o.preconditioner == "multigrid") simp.preconditioner = am::topopt::Preconditioner::Multigrid;
  else { std::cerr << "Unknown --preconditioner " << o.preconditioner << std::endl; return 1; }

  const std::vector<double> density = dom.initialDensity(o.volume_fraction, o.E_min);
  const std::vector<double> forces = lc.getForceVector();
  std::cout << "Load: " << o.bracket.load_force_N << " N distributed over "
            << lc.point_loads.size() << " nodes\nFixed DOFs: " << lc.constrained.size()
            << " (base constraint)" << std::endl;
  try {
    const double h = o.bracket.resolution_mm / 1000.0;  // mm to m
    am::topopt::TopologyOptimizer opt(dom.nx, dom.ny, dom.nz, h, am::fem::MaterialProperties(),
                                      simp);
    const auto t0 = std::chrono::steady_clock::now();
    const am::topopt::OptimizationResult r = opt.run(density, forces, lc.constrained);
    const double elapsed =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "\n--- Results ---\nIterations: " << r.iterations
              << "\nConverged: " << (r.converged ? "yes" : "no") << "\nTime: " << elapsed
              << " s" << std::endl;
    fs::create_directories(o.output_dir);
    save_npy(o.output_dir + "/density_field.npy", r.final_density, dom.nx, dom.ny, dom.nz);
    save_metadata(o.output_dir + "/metadata.json", o, dom, r, elapsed);
    std::ofstream hist(o.output_dir + "/history.csv");
    hist << "iteration,compliance,volume\n";
    for (size_t i = 0; i < r.compliance_history.size(); ++i)
      hist << i << "," << r.compliance_history[i] << "," << r.volume_history[i] << "\n";
    std::cout << "Saved: " << o.output_dir << "/history.csv" << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
