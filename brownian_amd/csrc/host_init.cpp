// Host-side start-up of libbrownian: the reference generator and the
// reference's initialisation routines (colloid placement, fluid placement,
// velocities, orientations).  The reference runs these on the host too
// (cpp/initialize.cpp, cpp/run.cpp:3-8); they are done once per run.
#include "host_init.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace bw {

// cpp/parameters.hpp:16-42: combined L'Ecuyer generator with a 32-entry
// Bays-Durham table.  A positive seed starts from idum = 123456789, iy = 0 and
// an all-zero table; a non-positive seed fills the table first.
void RefRng::reset(int s) {
    seed = s;
    idum = 123456789;
    iy = 0;
    memset(iv, 0, sizeof(iv));
}

double RefRng::next() {
    const int im1 = 2147483563, im2 = 2147483399, ia1 = 40014, ia2 = 40692, iq1 = 53668, iq2 = 52774;
    const int imm = im1 - 1, ir1 = 12211, ir2 = 3791, ntab = 32, ndiv = 1 + imm / ntab;
    const double rnmx = 1 - 1.2e-7, am = 1.0 / im1;
    int j, k;
    if (seed <= 0) {
        seed = std::max(-seed, 1);
        idum = seed;
        for (j = ntab + 8; j >= 1; j--) {
            k = seed / iq1;
            seed = ia1 * (seed - k * iq1) - k * ir1;
            if (seed < 0) seed += im1;
            if (j <= ntab) iv[j] = seed;
        }
        iy = iv[1];
    }
    k = seed / iq1;
    seed = ia1 * (seed - k * iq1) - k * ir1;
    if (seed < 0) seed += im1;
    k = idum / iq2;
    idum = ia2 * (idum - k * iq2) - k * ir2;
    if (idum < 0) idum += im2;
    j = 1 + iy / ndiv;
    iy = iv[j] - idum;
    iv[j] = seed;
    if (iy < 1) iy += imm;
    return std::min(am * iy, rnmx);
}

void RefRng::get(int *o) const {
    o[0] = seed; o[1] = idum; o[2] = iy;
    for (int i = 1; i <= 32; i++) o[2 + i] = iv[i];
    o[35] = 0;
}

namespace {
struct P3 {
    double x, y, z;
};
// img(a,b) per component, cpp/parameters.hpp:74,76 (lround: half away from zero)
inline double img1(double a, double L) { return a - L * (double)lround(a / L); }
inline double dist(const P3 &t, const P3 &c, double L) {
    double x = img1(t.x - c.x, L), y = img1(t.y - c.y, L), z = img1(t.z - c.z, L);
    return sqrt((x * x + y * y) + z * z);
}
inline P3 rnd3(RefRng &g) {
    double x = g.next(), y = g.next(), z = g.next();
    return {x, y, z};
}
// Box-Muller normal from two generator draws (velocity law BW_VEL_MAXWELL)
inline double gauss(RefRng &g) {
    double u1 = g.next(), u2 = g.next();
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}
}  // namespace

// cpp/initialize.cpp:31-60 (+ Maxwell-Boltzmann option)
void init_colloids(const InitParams &p, RefRng &g, double *pos, double *vel, double *omega) {
    const double L = p.L;
    const int nc = p.n_colloid;
    double space_limit = 1.3 * p.sig_colloid, I = 0.1 * p.mass_colloid * p.sigma * p.sigma;
    double ang_vscale = sqrt(12.0 * p.kbt / I), vscale = sqrt(12.0 * p.kbt / p.mass_colloid);
    std::vector<P3> c(nc);
    int counter = 0;
    long long attempts = 0;
    while (counter < nc) {
        P3 r = rnd3(g);
        P3 t = {r.x * L - 0, r.y * L - 0, r.z * L - 0};
        int check = 1;
        for (int j = 0; j < counter; j++)
            if (dist(t, c[j], L) < space_limit) check = 0;
        if (check) c[counter++] = t;
        if (++attempts > 100000000LL) break;  // unplaceable: leave the rest at 0
    }
    std::vector<P3> v(nc);
    P3 avr = {0, 0, 0};
    for (int j = 0; j < nc; j++) {
        if (p.vel_law == 1) {
            double s = sqrt(p.kbt / p.mass_colloid);
            v[j] = {gauss(g) * s, gauss(g) * s, gauss(g) * s};
        } else {
            P3 r = rnd3(g);
            v[j] = {(r.x * 1 - 0.5) * vscale, (r.y * 1 - 0.5) * vscale, (r.z * 1 - 0.5) * vscale};
        }
        avr = {avr.x + v[j].x, avr.y + v[j].y, avr.z + v[j].z};
    }
    avr = {avr.x / (double)nc, avr.y / (double)nc, avr.z / (double)nc};
    for (int j = 0; j < nc; j++) {
        v[j] = {v[j].x - avr.x, v[j].y - avr.y, v[j].z - avr.z};
        P3 w;
        if (p.vel_law == 1) {
            double s = sqrt(p.kbt / I);
            w = {gauss(g) * s, gauss(g) * s, gauss(g) * s};
        } else {
            P3 r = rnd3(g);
            w = {(r.x * 1 - 0.5) * ang_vscale, (r.y * 1 - 0.5) * ang_vscale, (r.z * 1 - 0.5) * ang_vscale};
        }
        if (pos) { pos[3 * j] = c[j].x; pos[3 * j + 1] = c[j].y; pos[3 * j + 2] = c[j].z; }
        if (vel) { vel[3 * j] = v[j].x; vel[3 * j + 1] = v[j].y; vel[3 * j + 2] = v[j].z; }
        if (omega) { omega[3 * j] = w.x; omega[3 * j + 1] = w.y; omega[3 * j + 2] = w.z; }
    }
}

// cpp/initialize.cpp:62-83 (+ Maxwell-Boltzmann option).  The exclusion test
// only visits colloids in neighbouring bins of a coarse grid; every colloid it
// skips is farther than sigma/2, so the accept/reject decisions, and with them
// the generator stream, are those of the reference's all-colloid loop.
void init_fluid(const InitParams &p, RefRng &g, const double *cpos, double *pos, double *vel) {
    const double L = p.L, rad = p.sigma * 0.5;
    const long long nf = (long long)p.L * p.L * p.L * p.density;
    const int nc = p.n_colloid;
    int G = std::max(1, (int)floor(L / std::max(rad + 1.0, 1.0)));
    G = std::min(G, 64);
    double h = L / G;
    std::vector<std::vector<int>> bins((size_t)G * G * G);
    auto binof = [&](double x) {
        int b = (int)floor(x / h);
        if (b >= G) b = G - 1;
        if (b < 0) b = 0;
        return b;
    };
    for (int j = 0; j < nc; j++)
        bins[(size_t)binof(cpos[3 * j]) + G * ((size_t)binof(cpos[3 * j + 1]) + G * binof(cpos[3 * j + 2]))].push_back(j);
    const bool small_grid = G < 3;
    long long counter = 0;
    while (counter < nf) {
        P3 r = rnd3(g);
        P3 t = {r.x * L - 0, r.y * L - 0, r.z * L - 0};
        int check = 1;
        if (small_grid) {
            for (int j = 0; j < nc && check; j++)
                if (dist(t, {cpos[3 * j], cpos[3 * j + 1], cpos[3 * j + 2]}, L) < rad) check = 0;
        } else {
            int bx = binof(t.x), by = binof(t.y), bz = binof(t.z);
            for (int dz = -1; dz <= 1 && check; dz++)
                for (int dy = -1; dy <= 1 && check; dy++)
                    for (int dx = -1; dx <= 1 && check; dx++) {
                        int X = (bx + dx + G) % G, Y = (by + dy + G) % G, Z = (bz + dz + G) % G;
                        for (int j : bins[(size_t)X + G * ((size_t)Y + G * Z)])
                            if (dist(t, {cpos[3 * j], cpos[3 * j + 1], cpos[3 * j + 2]}, L) < rad) {
                                check = 0;
                                break;
                            }
                    }
        }
        if (check) {
            pos[3 * counter] = t.x; pos[3 * counter + 1] = t.y; pos[3 * counter + 2] = t.z;
            counter++;
        }
    }
    double vscale = sqrt(12.0 * p.kbt / p.mass_fl);
    double ax = 0, ay = 0, az = 0;
    for (long long j = 0; j < nf; j++) {
        double x, y, z;
        if (p.vel_law == 1) {
            double s = sqrt(p.kbt / p.mass_fl);
            x = gauss(g) * s; y = gauss(g) * s; z = gauss(g) * s;
        } else {
            P3 r = rnd3(g);
            x = (r.x * 1 - 0.5) * vscale; y = (r.y * 1 - 0.5) * vscale; z = (r.z * 1 - 0.5) * vscale;
        }
        vel[3 * j] = x; vel[3 * j + 1] = y; vel[3 * j + 2] = z;
        ax = ax + x; ay = ay + y; az = az + z;
    }
    ax = ax / (double)nf; ay = ay / (double)nf; az = az / (double)nf;
    for (long long j = 0; j < nf; j++) {
        vel[3 * j] = vel[3 * j] - ax;
        vel[3 * j + 1] = vel[3 * j + 1] - ay;
        vel[3 * j + 2] = vel[3 * j + 2] - az;
    }
}

// cpp/run.cpp:3-8 (+ uniform-on-sphere option)
void init_orientation(const InitParams &p, RefRng &g, const double *cpos, double *ra) {
    const double L = p.L;
    for (int i = 0; i < p.n_colloid; i++) {
        double x, y, z;
        if (p.orient_law == 1) {
            double u = 2.0 * g.next() - 1.0, ph = 2.0 * M_PI * g.next();
            double s = sqrt(1 - u * u);
            x = s * cos(ph); y = s * sin(ph); z = u;
        } else {
            P3 r = rnd3(g);
            P3 t = {r.x * L - 0, r.y * L - 0, r.z * L - 0};
            x = img1(cpos[3 * i] - t.x, L);
            y = img1(cpos[3 * i + 1] - t.y, L);
            z = img1(cpos[3 * i + 2] - t.z, L);
            double n = sqrt((x * x + y * y) + z * z);
            x = x / n; y = y / n; z = z / n;
        }
        ra[3 * i] = x; ra[3 * i + 1] = y; ra[3 * i + 2] = z;
    }
}

}  // namespace bw
