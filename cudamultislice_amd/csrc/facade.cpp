// facade.cpp -- UnmanagedMultisliceSimulation (U.cuh:62-199) over the C-ABI.
#include "../../include/UnmanagedMultisliceSimulation.h"

#include <cfloat>
#include <cmath>
#include <cstring>

UnmanagedMultisliceSimulation::UnmanagedMultisliceSimulation()
    : MaximumX(0), MaximumY(0), MaximumZ(0), MinimumX(0), MinimumY(0), MinimumZ(0), numberOfSlices(0), kmax(0),
      conventionaldz(1), maxposIm(0), maxposEW(0), minposEW(0), minposIm(0), EW(nullptr), TDS(nullptr),
      TDS2(nullptr), SimulatedImage(nullptr), imagecontrast(0), sim_(nullptr), status_(MSL_OK), resX_(0), resY_(0),
      reads_(0) {
  status_ = msl_create(&sim_, 0);
  // SimulateImage runs the reference's detector model (seeded noise)
  if (status_ == MSL_OK) status_ = msl_set_detector(sim_, 1, 0x4D534CULL);
}

UnmanagedMultisliceSimulation::~UnmanagedMultisliceSimulation() {
  delete[] EW;
  delete[] TDS;
  delete[] TDS2;
  delete[] SimulatedImage;
  msl_destroy(sim_);
}

const char* UnmanagedMultisliceSimulation::GetLastError() const {
  return sim_ ? msl_last_error(sim_) : "no device";
}

void UnmanagedMultisliceSimulation::Sync(msl_status st) {
  status_ = st;
  msl_info in;
  if (sim_ && msl_get_info(sim_, &in) == MSL_OK) {
    MaximumX = in.MaximumX; MaximumY = in.MaximumY; MaximumZ = in.MaximumZ;
    MinimumX = in.MinimumX; MinimumY = in.MinimumY; MinimumZ = in.MinimumZ;
    numberOfSlices = in.numberOfSlices; kmax = in.kmax; conventionaldz = in.conventionaldz;
    maxposIm = in.maxposIm; maxposEW = in.maxposEW; minposEW = in.minposEW; minposIm = in.minposIm;
    imagecontrast = in.imagecontrast;
    resX_ = in.resX; resY_ = in.resY;
  }
}

void UnmanagedMultisliceSimulation::LoadAtomFile(std::string p) { Sync(msl_load_atom_file(sim_, p.c_str())); }
void UnmanagedMultisliceSimulation::UploadBinnedAtoms() { Sync(msl_upload_binned_atoms(sim_)); }
void UnmanagedMultisliceSimulation::UploadBinnedAtomsTDS() { Sync(msl_upload_binned_atoms_tds(sim_)); }
// the reference reads ./fparams.dat (U.cu:1720); fall back to the embedded table
void UnmanagedMultisliceSimulation::GetParameterisation() {
  msl_status st = msl_get_parameterisation(sim_, "fparams.dat");
  if (st == MSL_ERR_IO) st = msl_get_parameterisation(sim_, nullptr);
  Sync(st);
}
void UnmanagedMultisliceSimulation::ApplyMicroscopeParameters(float V, float df, float m2, float a2, float cs, float B,
                                                              float D, float obj) {
  Sync(msl_apply_microscope_parameters(sim_, V, df, m2, a2, cs, B, D, obj));
}
void UnmanagedMultisliceSimulation::SetCalculationVariables(float ps, float df, int rx, int ry, float sx, float sy,
                                                            float sz, float ox, float oy) {
  Sync(msl_set_calculation_variables(sim_, ps, df, rx, ry, sx, sy, sz, ox, oy));
}
void UnmanagedMultisliceSimulation::STEMSetCalculationVariables(float ps, float df, int rx, int ry, float sx, float sy,
                                                                float sz) {
  Sync(msl_stem_set_calculation_variables(sim_, ps, df, rx, ry, sx, sy, sz));
}
void UnmanagedMultisliceSimulation::PreCalculateFrequencies() { Sync(msl_precalculate_frequencies(sim_)); }
void UnmanagedMultisliceSimulation::InitialiseWavefunctions() { Sync(msl_initialise_wavefunctions(sim_)); }
void UnmanagedMultisliceSimulation::InitialiseWavefunctions(float dz) {
  Sync(msl_initialise_wavefunctions_dz(sim_, dz));
}
void UnmanagedMultisliceSimulation::STEMInitialiseWavefunctions(int px, int py) {
  Sync(msl_stem_initialise_wavefunctions(sim_, px, py));
}
void UnmanagedMultisliceSimulation::AllocTDS() {
  Sync(msl_alloc_tds(sim_));
  delete[] TDS;
  delete[] TDS2;
  TDS = new float[(size_t)resX_ * resY_]();
  TDS2 = new float[(size_t)resX_ * resY_]();
}
void UnmanagedMultisliceSimulation::MultisliceStep(int it) { Sync(msl_multislice_step(sim_, it)); }
void UnmanagedMultisliceSimulation::MultisliceStepConv(int it) { Sync(msl_multislice_step_conv(sim_, it)); }
void UnmanagedMultisliceSimulation::MultisliceStepRS(int it) { Sync(msl_multislice_step_rs(sim_, it)); }
void UnmanagedMultisliceSimulation::AllocHostImage() {
  delete[] SimulatedImage;
  SimulatedImage = new msl_float2[(size_t)resX_ * resY_]();
}
// CTF image followed by the detector model (U.cu:2322-2428): SimulatedImage
// holds the complex detector image; with the model switched off it holds
// (I, 0) of the CTF intensity
void UnmanagedMultisliceSimulation::SimulateImage(float dose) {
  msl_status st = msl_simulate_image(sim_, dose);
  if (st == MSL_OK) {
    if (!SimulatedImage) AllocHostImage();
    st = msl_get_detector_image(sim_, reinterpret_cast<float*>(SimulatedImage));
    if (st == MSL_ERR_STATE) {  // detector model off
      float* tmp = new float[(size_t)resX_ * resY_];
      st = msl_get_image(sim_, tmp);
      for (size_t i = 0; i < (size_t)resX_ * resY_; ++i) SimulatedImage[i] = msl_float2{tmp[i], 0.f};
      delete[] tmp;
    }
  }
  Sync(st);
}
void UnmanagedMultisliceSimulation::GetExitWave() {
  delete[] EW;
  EW = new msl_float2[(size_t)resX_ * resY_];
  Sync(msl_get_exit_wave(sim_, reinterpret_cast<float*>(EW)));
}
void UnmanagedMultisliceSimulation::GetSTEMExitWave() {
  delete[] EW;
  EW = new msl_float2[(size_t)resX_ * resY_];
  Sync(msl_get_stem_exit_wave(sim_, reinterpret_cast<float*>(EW)));
}
void UnmanagedMultisliceSimulation::FreeExitWave() {
  delete[] EW;
  EW = nullptr;
}
void UnmanagedMultisliceSimulation::AddTDSWaves() {
  msl_status st = msl_add_tds_waves(sim_);
  if (st == MSL_OK && TDS) st = msl_get_tds(sim_, TDS);
  Sync(st);
}
float UnmanagedMultisliceSimulation::GetEWValueAbs(int x, int y) {
  const msl_float2 v = EW[x + y * resX_];
  return std::sqrt(v.x * v.x + v.y * v.y);
}
float UnmanagedMultisliceSimulation::GetEWValueRe(int x, int y) { return EW[x + y * resX_].x; }
float UnmanagedMultisliceSimulation::GetEWValueIm(int x, int y) { return EW[x + y * resX_].y; }
// |SimulatedImage| plus readout noise, a fresh draw per call (U.cu:2497-2508)
float UnmanagedMultisliceSimulation::GetImValue(int x, int y) {
  float noise = 4.f;
  msl_readout_noise(sim_, reads_++, &noise);
  return noise + GetDQEImValue(x, y);
}
// |SimulatedImage| (U.cu:2510-2512)
float UnmanagedMultisliceSimulation::GetDQEImValue(int x, int y) {
  const msl_float2 v = SimulatedImage[x + y * resX_];
  return std::sqrt(v.x * v.x + v.y * v.y);
}
float UnmanagedMultisliceSimulation::GetTDSMin() {
  float m = FLT_MAX;
  for (size_t j = 0; j < (size_t)resX_ * resY_; ++j) m = TDS[j] < m ? TDS[j] : m;
  return m;
}
float UnmanagedMultisliceSimulation::GetTDSMax() {
  float m = FLT_MIN;  // as U.cu:2780
  for (size_t j = 0; j < (size_t)resX_ * resY_; ++j) m = TDS[j] > m ? TDS[j] : m;
  return m;
}
float UnmanagedMultisliceSimulation::GetTDSVal(int x, int y) { return TDS[x + resX_ * y]; }
// declared but never defined in the reference (U.cu:2519-2547); returns the
// contrast computed by SimulateImage, as ManagedMultislice does (MM.cpp:205-210)
float UnmanagedMultisliceSimulation::GetImageContrast(int, int, int, int) { return imagecontrast; }
