// ic.cpp -- seeded initial conditions for the benchmark systems (host code).
//
// The reference draws unseeded uniform boxes (Renderer.swift:441-493); these generators
// replace it with reproducible self-gravitating systems in G = 1 units.  All randomness comes
// from one SplitMix64 stream per call (Steele, Lea & Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014), so a (generator, n, seed) triple always yields the same
// bytes on one platform.  Every particle gets type 0 (a massive body), radius 0.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/astro/astro.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform_open() {  // (0,1)
        double u;
        do u = uniform();
        while (u <= 0.0);
        return u;
    }
    double normal() {  // Box-Muller, one value per call
        double u1 = uniform_open(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

void isotropic(SplitMix64& r, double len, double* v) {
    double z = 2.0 * r.uniform() - 1.0;
    double phi = 2.0 * M_PI * r.uniform();
    double s = std::sqrt(std::max(0.0, 1.0 - z * z));
    v[0] = len * s * std::cos(phi);
    v[1] = len * s * std::sin(phi);
    v[2] = len * z;
}

void store(uint32_t i, const double* x, const double* v, double m, PositionMass* pos,
           VelocityRadius* vel, ColorType* col) {
    std::memset(&pos[i], 0, sizeof(PositionMass));
    for (int k = 0; k < 3; k++) pos[i].position[k] = (float)x[k];
    pos[i].mass = (float)m;
    if (vel) {
        std::memset(&vel[i], 0, sizeof(VelocityRadius));
        for (int k = 0; k < 3; k++) vel[i].velocity[k] = (float)v[k];
    }
    if (col) {
        std::memset(&col[i], 0, sizeof(ColorType));
        col[i].color[3] = 1.0f;
        col[i].type = 0;
    }
}

// Plummer sphere with scale radius a and total mass M (Aarseth, Henon & Wielen 1974): radius
// from the cumulative mass, speed by von Neumann rejection on q^2 (1-q^2)^(7/2) times the
// local escape speed.  Radii beyond 10 a are redrawn.
void plummer_body(SplitMix64& r, double a, double M, double* x, double* v) {
    double rad;
    do {
        double X = r.uniform_open();
        rad = a / std::sqrt(std::pow(X, -2.0 / 3.0) - 1.0);
    } while (!(rad < 10.0 * a));
    isotropic(r, rad, x);
    double q, g;
    do {
        q = r.uniform();
        g = 0.1 * r.uniform();
    } while (g > q * q * std::pow(1.0 - q * q, 3.5));
    double vesc = std::sqrt(2.0 * M / a) * std::pow(1.0 + rad * rad / (a * a), -0.25);
    isotropic(r, q * vesc, v);
}

void recentre(std::vector<double>& X, std::vector<double>& V, const std::vector<double>& m,
              size_t b, size_t e) {
    double cx[3] = {0, 0, 0}, cv[3] = {0, 0, 0}, mt = 0;
    for (size_t i = b; i < e; i++) {
        mt += m[i];
        for (int k = 0; k < 3; k++) {
            cx[k] += m[i] * X[3 * i + k];
            cv[k] += m[i] * V[3 * i + k];
        }
    }
    if (mt <= 0) return;
    for (size_t i = b; i < e; i++)
        for (int k = 0; k < 3; k++) {
            X[3 * i + k] -= cx[k] / mt;
            V[3 * i + k] -= cv[k] / mt;
        }
}

}  // namespace

extern "C" int astro_ic_plummer(uint32_t n, uint64_t seed, PositionMass* pos,
                                VelocityRadius* vel, ColorType* col) {
    if (!pos || n == 0) return ASTRO_ERR_ARG;
    SplitMix64 r(seed);
    std::vector<double> X(3ull * n), V(3ull * n), m(n, 1.0 / n);
    for (uint32_t i = 0; i < n; i++) plummer_body(r, 1.0, 1.0, &X[3ull * i], &V[3ull * i]);
    recentre(X, V, m, 0, n);
    for (uint32_t i = 0; i < n; i++) store(i, &X[3ull * i], &V[3ull * i], m[i], pos, vel, col);
    return ASTRO_OK;
}

// Two exponential disks (scale length 1, exponential vertical profile of height 0.1), masses
// log-normal with sigma 0.5 normalised to 1/2 per disk, circular speeds from the enclosed
// mass (softened by eps), on a parabolic orbit: separation `sep`, pericentre `peri`.  Disk 2
// is inclined by 45 degrees about the x axis.
extern "C" int astro_ic_two_disks(uint32_t n_per_disk, uint64_t seed, float sep, float peri,
                                  PositionMass* pos, VelocityRadius* vel, ColorType* col) {
    if (!pos || n_per_disk == 0 || !(sep > 0.0f) || !(peri > 0.0f) || peri > sep)
        return ASTRO_ERR_ARG;
    SplitMix64 r(seed);
    const uint64_t n = 2ull * n_per_disk;
    const double eps = 0.1, h = 0.1;
    std::vector<double> X(3 * n), V(3 * n), m(n);
    for (int d = 0; d < 2; d++) {
        const size_t b = (size_t)d * n_per_disk, e = b + n_per_disk;
        double msum = 0;
        std::vector<double> R(n_per_disk);
        for (size_t i = b; i < e; i++) {
            m[i] = std::exp(0.5 * r.normal());
            msum += m[i];
            double rad = -std::log(r.uniform_open() * r.uniform_open());  // Gamma(2): R e^-R
            double phi = 2.0 * M_PI * r.uniform();
            double z = -h * std::log(r.uniform_open()) * (r.uniform() < 0.5 ? -1.0 : 1.0);
            X[3 * i + 0] = rad * std::cos(phi);
            X[3 * i + 1] = rad * std::sin(phi);
            X[3 * i + 2] = z;
            R[i - b] = rad;
        }
        for (size_t i = b; i < e; i++) m[i] *= 0.5 / msum;
        // enclosed mass by radius: sort particle radii once
        std::vector<uint32_t> ord(n_per_disk);
        for (uint32_t k = 0; k < n_per_disk; k++) ord[k] = k;
        std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return R[x] < R[y]; });
        double menc = 0;
        for (uint32_t k = 0; k < n_per_disk; k++) {
            size_t i = b + ord[k];
            menc += m[i];
            double rad = R[ord[k]];
            double vc = std::sqrt(menc * rad * rad / std::pow(rad * rad + eps * eps, 1.5));
            double phi = std::atan2(X[3 * i + 1], X[3 * i + 0]);
            V[3 * i + 0] = -vc * std::sin(phi);
            V[3 * i + 1] = vc * std::cos(phi);
            V[3 * i + 2] = 0.0;
        }
        if (d == 1) {
            const double c = std::cos(M_PI / 4), s = std::sin(M_PI / 4);
            for (size_t i = b; i < e; i++) {
                double y = X[3 * i + 1], z = X[3 * i + 2];
                X[3 * i + 1] = c * y - s * z;
                X[3 * i + 2] = s * y + c * z;
                double vy = V[3 * i + 1], vz = V[3 * i + 2];
                V[3 * i + 1] = c * vy - s * vz;
                V[3 * i + 2] = s * vy + c * vz;
            }
        }
        recentre(X, V, m, b, e);
    }
    // parabolic relative orbit of two unit-half masses: v^2 = 2 G M / d, L = sqrt(2 G M q)
    const double M = 1.0, d = sep;
    const double vrel = std::sqrt(2.0 * M / d);
    const double vt = std::sqrt(2.0 * M * peri) / d;
    const double vr = -std::sqrt(std::max(0.0, vrel * vrel - vt * vt));
    for (int dsk = 0; dsk < 2; dsk++) {
        const double sgn = dsk == 0 ? 0.5 : -0.5;
        for (size_t i = (size_t)dsk * n_per_disk; i < (size_t)(dsk + 1) * n_per_disk; i++) {
            X[3 * i + 0] += sgn * d;
            V[3 * i + 0] += sgn * vr;
            V[3 * i + 1] += sgn * vt;
        }
    }
    for (uint64_t i = 0; i < n; i++)
        store((uint32_t)i, &X[3 * i], &V[3 * i], m[i], pos, vel, col);
    return ASTRO_OK;
}

// Plummer sphere plus a uniform background cube (fraction `bg` of the particles, side
// `cube`, at rest), masses from a power law m^-slope on [mlo, mhi], total mass 1.
extern "C" int astro_ic_plummer_background(uint32_t n, uint64_t seed, float bg, float cube,
                                           float slope, float mlo, float mhi,
                                           PositionMass* pos, VelocityRadius* vel,
                                           ColorType* col) {
    if (!pos || n == 0 || bg < 0.0f || bg > 1.0f || !(mhi > mlo) || !(mlo > 0.0f))
        return ASTRO_ERR_ARG;
    SplitMix64 r(seed);
    const uint32_t nbg = (uint32_t)std::floor((double)n * bg + 0.5);
    const uint32_t npl = n - nbg;
    std::vector<double> X(3ull * n), V(3ull * n, 0.0), m(n);
    const double k = 1.0 - (double)slope;
    const double a = std::pow((double)mlo, k), b = std::pow((double)mhi, k);
    double mt = 0;
    for (uint32_t i = 0; i < n; i++) {
        double u = r.uniform();
        m[i] = std::fabs(k) < 1e-12 ? mlo * std::pow((double)mhi / mlo, u)
                                    : std::pow(a + u * (b - a), 1.0 / k);
        mt += m[i];
    }
    for (uint32_t i = 0; i < n; i++) m[i] /= mt;
    for (uint32_t i = 0; i < npl; i++) plummer_body(r, 1.0, 1.0, &X[3ull * i], &V[3ull * i]);
    recentre(X, V, m, 0, npl);
    for (uint32_t i = npl; i < n; i++)
        for (int q = 0; q < 3; q++) X[3ull * i + q] = ((double)cube) * (r.uniform() - 0.5);
    for (uint32_t i = 0; i < n; i++) store(i, &X[3ull * i], &V[3ull * i], m[i], pos, vel, col);
    return ASTRO_OK;
}

// nsys sizes drawn log-uniform in [nmin, nmax] (rounded to the nearest integer).
extern "C" int astro_ic_batch_sizes(uint32_t nsys, uint32_t nmin, uint32_t nmax, uint64_t seed,
                                    uint32_t* sizes) {
    if (!sizes || nmin == 0 || nmax < nmin) return ASTRO_ERR_ARG;
    SplitMix64 r(seed);
    const double lo = std::log((double)nmin), hi = std::log((double)nmax);
    for (uint32_t s = 0; s < nsys; s++) {
        double v = std::exp(lo + r.uniform() * (hi - lo));
        uint32_t n = (uint32_t)std::floor(v + 0.5);
        sizes[s] = n < nmin ? nmin : (n > nmax ? nmax : n);
    }
    return ASTRO_OK;
}
