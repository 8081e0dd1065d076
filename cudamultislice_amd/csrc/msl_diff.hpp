// msl_diff.hpp -- launch interface of the adjoint kernels (k_diff.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace msl {

struct AdjArgs {
  const float2* G;   // gradient w.r.t. the row-pass output a_i, (kx, y) domain
  const float2* Z;   // real_z: w_i = psi_{i-1} in real space; else z_{i-1} in the (kx, y) domain
  float2* Gout;      // gradient w.r.t. z_{i-1} (may alias G)
  const float2* t;   // transmission exp(i sigma V_i) of slice i, [ny][nx]
  float sigma;
  long long bstride; // elements between probes
  int nx, ny, nbatch;
  int per_line;      // probes each line processes
  int ngroups;       // workgroups per row
  float* partial;    // [ngroups][ny][nx] gradient partial sums
  // ngroups == 1: the gradient rows go straight to grad ([ny][nx]), added to
  // what it holds when accumulate != 0; partial is unused
  float* grad;
  int accumulate;
  const float2* tw;  // twiddles for nx
  int prune;         // band-limited G / z (only x < N/4, x >= 3N/4 non-zero)
};

int adj_lines(int nx, bool real_z);
hipError_t launch_row_adjoint(const AdjArgs& a, bool real_z, hipStream_t s);

}  // namespace msl
