// Drives the qball-shaped C++ adaptor on a small diamond-Si cell and writes inputs and
// results as raw little-endian arrays into an output directory, for comparison with the
// CPU oracle (tests/test_gpu_adaptor.py).  Usage: adaptor_check OUTDIR
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "qbh_qball.h"

using namespace qbh::qball;

template <typename T>
static void dump(const std::string& dir, const char* name, const std::vector<T>& v) {
  FILE* f = std::fopen((dir + "/" + name).c_str(), "wb");
  if (!f) { std::perror(name); std::exit(1); }
  std::fwrite(v.data(), sizeof(T), v.size(), f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: %s OUTDIR\n", argv[0]); return 2; }
  const std::string out = argv[1];
  const double alat = 10.26, ecut = 12.5;  // Si8, 25 Ry
  const double arows[9] = {alat, 0, 0, 0, alat, 0, 0, 0, alat};
  UnitCell cell(arows);
  Basis basis(cell, ecut, true);
  int np[3];
  density_grid(cell, ecut, np);
  Context ctx(0);
  FourierTransform ft(ctx, basis, np[0], np[1], np[2]);
  Basis vbasis(cell, 4.0 * ecut, true);
  FourierTransform vft(ctx, vbasis, np[0], np[1], np[2]);
  const int nst = 9, ngw = basis.localsize();
  std::printf("adaptor_check: ngw %d nst %d grid %d %d %d\n", ngw, nst, np[0], np[1], np[2]);
  std::fflush(stdout);

  // deterministic inputs (64-bit LCG)
  unsigned long long s = 12345;
  auto rnd = [&]() { s = s * 6364136223846793005ULL + 1442695040888963407ULL;
                     return ((s >> 11) * (1.0 / 9007199254740992.0)) - 0.5; };
  SlaterDet sd(basis, nst), dwf(basis, nst);
  for (auto& c : sd.c()) c = cd(rnd(), rnd());
  for (int n = 0; n < nst; ++n) sd.c()[(size_t)n * sd.mloc()] = cd(sd.c()[(size_t)n * sd.mloc()].real(), 0.0);
  for (int n = 0; n < nst; ++n) sd.occ()[n] = n < nst - 1 ? 2.0 : 0.0;
  EnergyFunctional ef(ft);
  for (auto& v : ef.v_r) v = rnd();
  ef.update_vr();
  // one species, 8 atoms, s + p projectors with tabulated twnl
  std::vector<double> tau, twnl(4 * (size_t)ngw), wt = {1.2, -0.5, -0.5, -0.5};
  for (int i = 0; i < 8; ++i) for (int d = 0; d < 3; ++d) tau.push_back(alat * (0.5 + rnd()));
  {
    int ig = 0;
    for (size_t r = 0; r < basis.rod_h().size(); ++r)
      for (int j = 0; j < basis.rod_size()[r]; ++j, ++ig) {
        const double h = basis.rod_h()[r], k = basis.rod_k()[r], l = basis.rod_lmin()[r] + j;
        double g[3];
        for (int d = 0; d < 3; ++d) g[d] = h * cell.b[d] + k * cell.b[3 + d] + l * cell.b[6 + d];
        const double gg = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        const double gi = gg > 0 ? 1.0 / gg : 0.0;
        twnl[ig] = 0.28 * std::exp(-0.5 * gg * gg);
        for (int d = 0; d < 3; ++d) twnl[(size_t)(1 + d) * ngw + ig] = 0.49 * g[d] * gi * gg * std::exp(-0.5 * gg * gg);
      }
  }
  ef.set_projectors({8}, {4}, {0, 1, 1, 1}, wt, twnl, tau);

  // FourierTransform::backward(c1, c2, f)
  std::vector<cd> f(ft.np012());
  ft.backward(&sd.c()[0], &sd.c()[sd.mloc()], f.data());
  // EnergyFunctional::energy(compute_hpsi=true)
  const double enl = ef.energy(sd, true, dwf);
  // NonLocalPotential::energy(..., compute_forces=true, fion)
  std::vector<std::vector<double>> fion;
  ef.nl_forces(sd, fion);
  // ChargeDensity::update_density
  ChargeDensity cd_(ft, vft, vbasis.localsize());
  cd_.update_density(sd);

  dump(out, "psi.bin", sd.c());
  dump(out, "v.bin", ef.v_r);
  dump(out, "twnl.bin", twnl);
  dump(out, "tau.bin", tau);
  dump(out, "f.bin", f);
  dump(out, "hpsi.bin", dwf.c());
  dump(out, "rho.bin", cd_.rhor);
  dump(out, "rhog.bin", cd_.rhog);
  dump(out, "fion.bin", fion[0]);
  std::vector<double> sc = {enl, cd_.nelectrons};
  dump(out, "scalars.bin", sc);
  std::printf("adaptor_check: enl %.15f nelectrons %.15f\n", enl, cd_.nelectrons);
  return 0;
}
