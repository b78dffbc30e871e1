// example_solve.cpp -- C++ driver over the host mirror (afopr_hip.h), written
// the way benchmark_hmc_dwf/hmc_dwf_eo.cpp uses AFopr/ASolver.
//
//   example_solve <Lx Ly Lz Lt> <Ls> <mq> <M0> <b> <c> <U.bin> <src.bin> <out_prefix>
//
// U.bin: Field_G (18, V, 4) doubles; src.bin: Field(24, V/2, Ls) doubles.
// Writes <out_prefix>.deo (D_eo src), <out_prefix>.x (solution of DdagD x = src
// by mixed-precision Richardson) and prints nconv and diff.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "afopr_hip.h"

static std::vector<double> read_all(const char* path, size_t n) {
  std::vector<double> v(n);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double)));
  if (!f) throw std::runtime_error(std::string("cannot read ") + path);
  return v;
}
static void write_all(const std::string& path, const std::vector<double>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}

int main(int argc, char** argv) {
  if (argc != 13) {
    std::fprintf(stderr, "usage: %s Lx Ly Lz Lt Ls mq M0 b c U.bin src.bin out_prefix\n", argv[0]);
    return 2;
  }
  try {
    dwfhip::Params p;
    p.lattice_size = {atoi(argv[1]), atoi(argv[2]), atoi(argv[3]), atoi(argv[4])};
    p.extent_of_5th_dimension = atoi(argv[5]);
    p.quark_mass = atof(argv[6]);
    p.domain_wall_height = atof(argv[7]);
    p.coefficient_b = atof(argv[8]);
    p.coefficient_c = atof(argv[9]);
    const size_t V = (size_t)p.lattice_size[0] * p.lattice_size[1] * p.lattice_size[2] *
                     p.lattice_size[3];
    const size_t nf = 24 * (V / 2) * p.extent_of_5th_dimension;
    std::vector<double> U = read_all(argv[10], 18 * V * 4);
    std::vector<double> src = read_all(argv[11], nf);
    const std::string out = argv[12];

    dwfhip::AFopr_Domainwall_5din_eo fopr_d(p);
    dwfhip::Params pf = p;
    pf.precision = DWF_F32;
    dwfhip::AFopr_Domainwall_5din_eo fopr_f(pf);
    fopr_d.set_config(U.data());
    fopr_f.set_config(U.data());

    dwfhip::AField b = fopr_d.new_field(), v = fopr_d.new_field(), x = fopr_d.new_field();
    fopr_d.convert(b, src.data());

    fopr_d.mult(v, b, "Deo");
    std::vector<double> host(nf);
    fopr_d.reverse(host.data(), v);
    write_all(out + ".deo", host);

    fopr_d.set_mode("DdagD");
    fopr_f.set_mode("DdagD");
    dwfhip::ASolver_Richardson solver(fopr_d, fopr_f, 100, 1.0e-28);
    int nconv = 0;
    double diff = 0;
    solver.solve(x, b, nconv, diff);
    fopr_d.reverse(host.data(), x);
    write_all(out + ".x", host);
    std::printf("nconv %d diff %.6e\n", nconv, diff);
    return (nconv >= 0 && diff < 1e-13) ? 0 : 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
