// example_tem.cpp -- the plane-wave TEM sequence of the reference's GUI
// ("CUDA Simulation GUI/MainWindow.xaml.cs":238-283, slice count as
// "GUI Simulation/ManagedMultislice.cpp":105-111) written against the drop-in
// facade: load, bin, parameterise, microscope, sampling, plane wave,
// conventional slices, exit wave, CTF image through the detector model.
//
//   example_tem <atoms.xyz> <out-prefix> [N=256] [ps=0.084844] [kV=200] [dz=1]
//               [dose=1000]
//
// Writes <prefix>_ew.bin (N*N complex64, row-major) and <prefix>_image.bin
// (N*N float32: the detector image |SimulatedImage|, GetDQEImValue) and prints
// one summary line.  Exit status 0 only if every call succeeded.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "UnmanagedMultisliceSimulation.h"

static bool ok(const UnmanagedMultisliceSimulation& s, const char* what) {
  if (s.GetLastStatus() == MSL_OK) return true;
  std::fprintf(stderr, "%s failed: %s\n", what, s.GetLastError());
  return false;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s atoms.xyz out-prefix [N] [ps] [kV] [dz] [dose]\n", argv[0]);
    return 2;
  }
  const std::string path = argv[1], prefix = argv[2];
  const int N = argc > 3 ? std::atoi(argv[3]) : 256;
  const float ps = argc > 4 ? (float)std::atof(argv[4]) : 0.084844f;
  const float kV = argc > 5 ? (float)std::atof(argv[5]) : 200.f;
  const float dz = argc > 6 ? (float)std::atof(argv[6]) : 1.f;
  const float dose = argc > 7 ? (float)std::atof(argv[7]) : 1000.f;

  UnmanagedMultisliceSimulation sim;
  if (!ok(sim, "create")) return 1;
  sim.LoadAtomFile(path);
  if (!ok(sim, "LoadAtomFile")) return 1;
  sim.UploadBinnedAtoms();
  if (!ok(sim, "UploadBinnedAtoms")) return 1;
  sim.GetParameterisation();
  if (!ok(sim, "GetParameterisation")) return 1;
  // df -50 nm, Cs 1.2 mm, 20 mrad objective aperture (units: kV, A, mrad)
  sim.ApplyMicroscopeParameters(kV, -500.f, 0.f, 0.f, 1.2e7f, 0.f, 0.f, 20.f);
  sim.SetCalculationVariables(ps, -500.f, N, N, sim.MaximumX - sim.MinimumX, sim.MaximumY - sim.MinimumY,
                              sim.MaximumZ - sim.MinimumZ, 2.f, 2.f);
  sim.PreCalculateFrequencies();
  sim.InitialiseWavefunctions(dz);
  if (!ok(sim, "set-up")) return 1;
  const int n = (int)std::ceil((sim.MaximumZ - sim.MinimumZ) / sim.conventionaldz);
  for (int i = 1; i <= n; ++i) {
    sim.MultisliceStepConv(i);
    if (!ok(sim, "MultisliceStepConv")) return 1;
  }
  sim.GetExitWave();
  if (!ok(sim, "GetExitWave")) return 1;
  sim.AllocHostImage();
  sim.SimulateImage(dose);  // electrons per pixel
  if (!ok(sim, "SimulateImage")) return 1;

  std::vector<float> img((size_t)N * N);
  for (size_t i = 0; i < img.size(); ++i) img[i] = sim.GetDQEImValue((int)(i % N), (int)(i / N));
  FILE* f = std::fopen((prefix + "_ew.bin").c_str(), "wb");
  FILE* g = std::fopen((prefix + "_image.bin").c_str(), "wb");
  const bool wrote = f && g && std::fwrite(sim.EW, sizeof(msl_float2), (size_t)N * N, f) == (size_t)N * N &&
                     std::fwrite(img.data(), sizeof(float), img.size(), g) == img.size();
  if (f) std::fclose(f);
  if (g) std::fclose(g);
  if (!wrote) {
    std::fprintf(stderr, "could not write %s_*.bin\n", prefix.c_str());
    return 1;
  }
  std::printf("slices %d dz %g |EW| max %g min %g at (10,20) %g image contrast %g read (10,20) %g\n", n,
              sim.conventionaldz, sim.maxposEW, sim.minposEW, sim.GetEWValueAbs(10, 20), sim.imagecontrast,
              sim.GetImValue(10, 20));
  sim.FreeExitWave();
  return 0;
}
