// rng.hpp -- the reference's random stream: Park-Miller minimal standard
// generator (Schrage's method, u = seed / (2^31 - 1)) and the Marsaglia polar
// Gaussian that caches its second deviate (src/random.cpp:7-68).
#pragma once
#include <cmath>

namespace mapp {

struct ParkMiller {
    int seed;
    bool have_spare = false;
    double spare = 0.0;
    explicit ParkMiller(int s) : seed(s) {}
    double uniform()
    {
        const int q = seed / 127773, r = seed % 127773;
        const int t = 16807 * r - 2836 * q;
        seed = t > 0 ? t : t + 2147483647;
        return seed * 1.0 / 2147483647;
    }
    double gaussian()
    {
        if (have_spare) {
            have_spare = false;
            return spare;
        }
        have_spare = true;
        double u = 0.0, v = 0.0, s = 1.0;
        while (s >= 1.0) {
            u = 2.0 * uniform() - 1.0;
            v = 2.0 * uniform() - 1.0;
            s = u * u + v * v;
        }
        const double f = std::sqrt(-2.0 * std::log(s) / s);
        spare = v * f;
        return u * f;
    }
};

}  // namespace mapp
