// fieldgen.h -- generated gauge fields, clover inverse and sources (device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qws_device.h"

namespace qws {

enum { TAG_LINKS = 1, TAG_SOURCE = 2 };

// 2x2 chirality blocks of sigma_{mu nu} = (i/2)[gamma_mu, gamma_nu] for the six
// planes (mu<nu), chirality i, spin pair p, p', (re, im).
struct SigmaChiral {
  double v[6][2][2][2][2];
};

SigmaChiral sigma_chiral_blocks();  // built on the host from QWS's gamma matrices

void launch_gen_links(const Geo& g, double eps, uint64_t seed, double* Ud, hipStream_t st);
void launch_links_d2s(const Geo& g, int nl, const double* Ud, float* Us, hipStream_t st);
void launch_clover(const Geo& g, const double* Ud, double kcsw, const SigmaChiral& sig, double* Cd, hipStream_t st);
void launch_clover_d2s(const Geo& g, const double* Cd, float* Cs, hipStream_t st);
void launch_gaussian_spinor(const Geo& g, uint64_t seed, double* out, hipStream_t st);
void launch_point_spinor(const Geo& g, const int x[4], int spin, int colour, double* out, hipStream_t st);

}  // namespace qws
