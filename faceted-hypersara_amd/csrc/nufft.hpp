// Scaled-FFT plan shared by the NUFFT and data-fidelity translation units.
#pragma once

#include "common.hpp"

struct fhs_nufft {
  int device = 0;
  int Ny = 0, Nx = 0, Kty = 0, Ktx = 0;
  float* sy = nullptr;   // [Ny] scaling factors
  float* sx = nullptr;   // [Nx]
  float2* twy = nullptr; // [Kty] exp(-2 pi i k / Kty)
  float2* twx = nullptr; // [Ktx]
};
