// Host-side physics helpers of the engine (no device code).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace cvf {

// Distance-shell table of gpu_water3D.cu:865-915 for an L^3 periodic lattice:
// distinct float minimum-image distances from cell 0 below the cutoff, in
// discovery order (cell index order), with multiplicity (count/2)*N
// (integer division, so odd classes lose a pair for L < 12).
void shell_table(int L, float r_cutoff, std::vector<float>& dist, std::vector<int32_t>& freq);

struct Thresholds {
  double Jeff, beta;
  uint64_t sigma[33];  // accept iff u32 < sigma[(eta*dhb + 1)*11 + dsig + 5]
  uint64_t eta[9];     // accept iff u32 < eta[dN + 4]
  uint64_t sw_hb, sw_s;
};

// Integer acceptance thresholds: p -> floor(p * 2^32), p >= 1 -> 2^32.
//   sigma:  p = exp(-beta*dE), dE = -Jeff*h - J_sig*d      (gpu_water3D.cu:1975-2000)
//   eta:    p = exp(beta*(Jeff*dN))                         (3416-3438)
//   SW:     p = 1 - exp(-beta*|J|)                          (992-993)
// with Jeff = J_hb - P*v_hb evaluated in float (391) and beta = 1/(double)T.
Thresholds make_thresholds(float P, float T, float J_hb, float J_sig, float v_hb);

// Sets the message cvf_last_error() returns (printf format) and returns code.
int set_error(int code, const char* fmt, ...);

}  // namespace cvf
