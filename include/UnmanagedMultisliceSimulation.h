// UnmanagedMultisliceSimulation.h -- drop-in C++ facade over libmsl.
//
// Same class name, public methods, signatures and public fields as the
// reference's UnmanagedMultisliceSimulation ("GUI Simulation/
// unmanagedGUISimulation.cuh", U.cuh:62-199), so a caller written against the
// reference (its C++/CLI bridge, ManagedMultislice.cpp) recompiles unchanged.
// No CUDA or HIP types appear here: device state lives behind msl_sim.
// Methods keep the reference's void signatures; the status of the last call
// is available from GetLastStatus() / GetLastError() instead of exit(1).
#pragma once
#include <string>

#include "msl.h"

#ifndef MSL_FLOAT2_DEFINED
#define MSL_FLOAT2_DEFINED
// layout-compatible with CUDA's / HIP's float2
struct msl_float2 {
  float x, y;
};
#endif

struct Atom {  // U.cuh:11-15
  int atomicNumber;
  float x, y, z;
};

class UnmanagedMultisliceSimulation {
 public:
  UnmanagedMultisliceSimulation();   // U.cu:2551-2556
  ~UnmanagedMultisliceSimulation();  // U.cu:2558-2561
  UnmanagedMultisliceSimulation(const UnmanagedMultisliceSimulation&) = delete;
  UnmanagedMultisliceSimulation& operator=(const UnmanagedMultisliceSimulation&) = delete;

  // public fields, U.cuh:127-150 and 197
  float MaximumX, MaximumY, MaximumZ, MinimumX, MinimumY, MinimumZ;
  int numberOfSlices;
  float kmax;
  float conventionaldz;
  float maxposIm, maxposEW, minposEW, minposIm;
  msl_float2* EW;
  float* TDS;
  float* TDS2;
  msl_float2* SimulatedImage;
  float imagecontrast;

  // methods, U.cuh:154-195
  void LoadAtomFile(std::string filepath);
  void UploadBinnedAtoms();
  void UploadBinnedAtomsTDS();
  void GetParameterisation();
  void ApplyMicroscopeParameters(float Voltage, float defocus, float Mod2f, float Arg2f, float spherical, float B,
                                 float D, float obj);
  void SetCalculationVariables(float PixelScaleIn, float defocus, int resX, int resY, float SizeX, float SizeY,
                               float SizeZ, float blockoffsetx, float blockoffsety);
  void STEMSetCalculationVariables(float PixelScaleIn, float defocus, int resX, int resY, float SizeX, float SizeY,
                                   float SizeZ);
  void PreCalculateFrequencies();
  void InitialiseWavefunctions();
  void InitialiseWavefunctions(float dz);
  void STEMInitialiseWavefunctions(int posx, int posy);
  void AllocTDS();
  void MultisliceStep(int iteration);
  void MultisliceStepConv(int iteration);
  void MultisliceStepRS(int iteration);
  void AllocHostImage();
  void SimulateImage(float doseperpix);
  void GetExitWave();
  void GetSTEMExitWave();
  void FreeExitWave();
  void AddTDSWaves();
  float GetEWValueAbs(int xpos, int ypos);
  float GetEWValueRe(int xpos, int ypos);
  float GetEWValueIm(int xpos, int ypos);
  float GetImValue(int xpos, int ypos);
  float GetTDSMin();
  float GetTDSMax();
  float GetTDSVal(int xpos, int ypos);
  float GetDQEImValue(int xpos, int ypos);
  float GetImageContrast(int t, int b, int l, int r);

  // additions (no counterpart in the reference)
  msl_status GetLastStatus() const { return status_; }
  const char* GetLastError() const;
  msl_sim* Handle() { return sim_; }  // e.g. msl_set_detector(Handle(), 0, seed) for the CTF image only

 private:
  void Sync(msl_status st);
  msl_sim* sim_;
  msl_status status_;
  int resX_, resY_;
  long long reads_;  // readout-noise draws taken by GetImValue
};
