// seeded.hpp -- counter-based synthetic spectra (host side, FP64).
//
// Every value is a function of (seed, field, radial index k, rj mode j)
// only, so any subset of a field can be generated in any order and the
// same numbers come out at every resolution that holds the mode:
//   key    = h(h(h(h(seed) ^ field) ^ k) ^ j)
//   u1     = (h(key) >> 11 + 1/2) 2^-53          in (0, 1)
//   u2     = (h(key ^ 0xD1B54A32D192ED03) >> 11) 2^-53   in [0, 1)
//   gauss  = sqrt(-2 ln u1) cos(2 pi u2)          (Box & Muller 1958)
// with h the SplitMix64 output function of a state x (Steele, Lea & Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014):
//   z = x + 0x9E3779B97F4A7C15; z = (z ^ z>>30) 0xBF58476D1CE4E5B9;
//   z = (z ^ z>>27) 0x94D049BB133111EB; h = z ^ z>>31.
// A spectrum multiplies the Gaussian by amplitude (l+1)^-power and, when
// asked, by the radial envelope ((r - r_in)(r_out - r))^2 / (d/2)^4
// (d = r_out - r_in), which vanishes with its first derivative at both
// walls, so the fields meet rigid, insulating and fixed-value boundary
// conditions before the first solve.
#pragma once

#include <cmath>
#include <cstdint>

namespace calypso {

inline uint64_t splitmix64_out(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double seeded_gauss(uint64_t seed, uint64_t field, uint64_t k, uint64_t j)
{
    const uint64_t key = splitmix64_out(
        splitmix64_out(splitmix64_out(splitmix64_out(seed) ^ field) ^ k) ^ j);
    const double u1 = ((double)(splitmix64_out(key) >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = (double)(splitmix64_out(key ^ 0xD1B54A32D192ED03ull) >> 11) * 0x1.0p-53;
    const double two_pi = 8.0 * std::atan(1.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2);
}

// out[j + J k] for J = (L+1)^2 modes j = l(l+1)+m and nr radii; modes with
// l < l_min are zero.
inline void seeded_spectrum(uint64_t seed, int field, int L, int nr, const double *radius,
                            double amplitude, double power, bool envelope, int l_min,
                            double *out)
{
    const int J = (L + 1) * (L + 1);
    const double r_in = radius[0], r_out = radius[nr - 1];
    const double h = 0.5 * (r_out - r_in), h4 = h * h * h * h;
    for (int k = 0; k < nr; ++k) {
        const double r = radius[k];
        const double e = envelope ? ((r - r_in) * (r_out - r)) * ((r - r_in) * (r_out - r)) / h4
                                  : 1.0;
        for (int l = 0; l <= L; ++l) {
            const double a = amplitude * std::pow((double)(l + 1), -power) * e;
            for (int m = -l; m <= l; ++m) {
                const int j = l * (l + 1) + m;
                out[j + (size_t)J * k] =
                    l < l_min ? 0.0
                              : a * seeded_gauss(seed, (uint64_t)field, (uint64_t)k, (uint64_t)j);
            }
        }
    }
}

}  // namespace calypso
