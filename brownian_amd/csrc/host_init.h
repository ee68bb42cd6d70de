// Host-side start-up of libbrownian (reference generator and initialisation).
#pragma once

namespace bw {

struct RefRng {
    int seed, idum, iy, iv[33];
    void reset(int s);
    double next();
    void get(int *out36) const;
};

struct InitParams {
    int L, density, n_colloid;
    double kbt, mass_fl, mass_colloid, sig_colloid, sigma;
    int vel_law, orient_law;
};

void init_colloids(const InitParams &p, RefRng &g, double *pos, double *vel, double *omega);
void init_fluid(const InitParams &p, RefRng &g, const double *cpos, double *pos, double *vel);
void init_orientation(const InitParams &p, RefRng &g, const double *cpos, double *ra);

}  // namespace bw
