// synth.cpp -- seeded synthetic ALS-like tiles (the benchmark workload).
//
// A tile of edge `side` metres: a fraction `ground_frac` of the points lie on
// a terrain surface made of three sinusoids (amplitudes 2.5, 1.5 and 1.0 m,
// wavelengths 50-200 m, random directions and phases); the rest of the
// non-degenerate points fall in Gaussian vegetation clusters and box-shaped
// buildings up to 20 m tall (one cluster per 2,500 m^2).  Gaussian noise of
// `noise` metres is added to x, y and z.  Optional exactly planar roof
// patches (constant quantized z) and exactly collinear wires (axis-aligned or
// 45 degrees on the quantum lattice, constant z) add degenerate
// neighborhoods.  Coordinates are quantized like LAS records:
// v = offset + round((v - offset) / quantum) * quantum, stored as FP64, and
// the point order is shuffled.  The random stream is xoshiro256** seeded
// through splitmix64 from the tile seed; tile i of a run uses
// splitmix64(0x5EED ^ i) (the reference's synthetic default seed,
// include/pdg/SyntheticCloud.hpp:21).
#include "gpupdal_amd.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace {

uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t t = seed;
        for (auto& v : s) v = splitmix64(t);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double u() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    double normal()
    {
        double a = 1.0 - u(), b = u();
        return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
    }
    uint64_t below(uint64_t m) { return m ? next() % m : 0; }
};

struct Terrain {
    double amp[3] = {2.5, 1.5, 1.0};
    double kx[3], ky[3], ph[3];
    explicit Terrain(Rng& r)
    {
        for (int j = 0; j < 3; ++j) {
            double lambda = 50.0 + 150.0 * r.u();
            double th = 6.283185307179586 * r.u();
            kx[j] = 6.283185307179586 * std::cos(th) / lambda;
            ky[j] = 6.283185307179586 * std::sin(th) / lambda;
            ph[j] = 6.283185307179586 * r.u();
        }
    }
    double at(double x, double y) const
    {
        double z = 0.0;
        for (int j = 0; j < 3; ++j) z += amp[j] * std::sin(kx[j] * x + ky[j] * y + ph[j]);
        return z;
    }
};

struct Cluster {
    double cx, cy, sigma, height;
    bool building;
};

} // namespace

extern "C" {

void gf_synth_defaults(gf_synth* p)
{
    p->seed = gf_tile_seed(0x5EED, 0);
    p->n = 250000;
    p->side = 500.0;
    p->ground_frac = 0.7;
    p->roof_frac = 0.0;
    p->wire_frac = 0.0;
    p->noise = 0.05;
    p->quantum = 0.01;
    p->offset_x = 500000.0;
    p->offset_y = 4000000.0;
}

uint64_t gf_tile_seed(uint64_t base, uint64_t tile)
{
    uint64_t s = base ^ tile;
    return splitmix64(s);
}

int gf_synth_cloud(const gf_synth* p, double* x, double* y, double* z)
{
    if (!p || (p->n && (!x || !y || !z)) || !(p->side > 0) || !(p->quantum > 0))
        return GF_ERR_ARG;
    Rng r(p->seed);
    const double L = p->side, q = p->quantum;
    Terrain ter(r);
    const uint64_t n = p->n;
    const uint64_t nroof = static_cast<uint64_t>(std::floor(n * p->roof_frac));
    const uint64_t nwire = static_cast<uint64_t>(std::floor(n * p->wire_frac));
    const uint64_t rest = n - nroof - nwire;
    const uint64_t nground = static_cast<uint64_t>(std::floor(rest * p->ground_frac));
    const uint64_t ncl_pts = rest - nground;

    const uint64_t ncl = std::max<uint64_t>(1, static_cast<uint64_t>(std::llround(L * L / 2500.0)));
    std::vector<Cluster> cl(ncl);
    for (auto& c : cl) {
        c.cx = r.u() * L;
        c.cy = r.u() * L;
        c.sigma = 2.0 + 3.0 * r.u();
        c.height = 3.0 + 17.0 * r.u();
        c.building = r.u() < 0.4;
    }
    auto clampL = [&](double v) { return v < 0 ? 0.0 : (v >= L ? std::nextafter(L, 0.0) : v); };
    auto qz = [&](double v) { return std::llround(v / q) * q; };
    auto put = [&](uint64_t i, double lx, double ly, double lz) {
        x[i] = p->offset_x + std::llround(lx / q) * q;
        y[i] = p->offset_y + std::llround(ly / q) * q;
        z[i] = qz(lz);
    };
    uint64_t i = 0;
    for (uint64_t t = 0; t < nground; ++t, ++i) {
        double lx = r.u() * L, ly = r.u() * L;
        double lz = ter.at(lx, ly);
        put(i, clampL(lx + p->noise * r.normal()), clampL(ly + p->noise * r.normal()),
            lz + p->noise * r.normal());
    }
    for (uint64_t t = 0; t < ncl_pts; ++t, ++i) {
        const Cluster& c = cl[r.below(ncl)];
        double lx, ly, lz;
        if (c.building) {
            lx = c.cx + c.sigma * (2.0 * r.u() - 1.0);
            ly = c.cy + c.sigma * (2.0 * r.u() - 1.0);
            double base = ter.at(c.cx, c.cy);
            lz = base + (r.u() < 0.8 ? c.height : c.height * r.u());
        } else {
            lx = c.cx + c.sigma * r.normal();
            ly = c.cy + c.sigma * r.normal();
            lx = clampL(lx);
            ly = clampL(ly);
            lz = ter.at(lx, ly) + c.height * r.u();
        }
        put(i, clampL(lx + p->noise * r.normal()), clampL(ly + p->noise * r.normal()),
            lz + p->noise * r.normal());
    }
    if (nroof) {
        const uint64_t np = std::max<uint64_t>(1, static_cast<uint64_t>(std::llround(L * L / 10000.0)));
        std::vector<double> px(np), py(np), pe(np), pz(np);
        for (uint64_t k = 0; k < np; ++k) {
            px[k] = r.u() * L;
            py[k] = r.u() * L;
            pe[k] = 10.0 + 10.0 * r.u();
            pz[k] = qz(ter.at(px[k], py[k]) + 5.0 + 10.0 * r.u());
        }
        for (uint64_t t = 0; t < nroof; ++t, ++i) {
            uint64_t k = r.below(np);
            double lx = clampL(px[k] + pe[k] * (r.u() - 0.5));
            double ly = clampL(py[k] + pe[k] * (r.u() - 0.5));
            x[i] = p->offset_x + std::llround(lx / q) * q;
            y[i] = p->offset_y + std::llround(ly / q) * q;
            z[i] = pz[k];
        }
    }
    if (nwire) {
        const uint64_t nw = std::max<uint64_t>(1, static_cast<uint64_t>(std::llround(L / 100.0)));
        std::vector<int64_t> wx(nw), wy(nw);
        std::vector<int> wd(nw);
        std::vector<double> wz(nw);
        const int64_t steps = static_cast<int64_t>(std::floor(std::min(50.0, L / 2) / q));
        const int64_t lim = static_cast<int64_t>(std::floor(L / q)) - steps - 1;
        for (uint64_t k = 0; k < nw; ++k) {
            wx[k] = static_cast<int64_t>(r.below(static_cast<uint64_t>(std::max<int64_t>(lim, 1))));
            wy[k] = static_cast<int64_t>(r.below(static_cast<uint64_t>(std::max<int64_t>(lim, 1))));
            wd[k] = static_cast<int>(r.below(3));  // 0: along x, 1: along y, 2: 45 degrees
            wz[k] = qz(ter.at(wx[k] * q, wy[k] * q) + 8.0 + 4.0 * r.u());
        }
        for (uint64_t t = 0; t < nwire; ++t, ++i) {
            uint64_t k = r.below(nw);
            int64_t m = static_cast<int64_t>(r.below(static_cast<uint64_t>(steps + 1)));
            int64_t ix = wx[k] + (wd[k] != 1 ? m : 0);
            int64_t iy = wy[k] + (wd[k] != 0 ? m : 0);
            x[i] = p->offset_x + ix * q;
            y[i] = p->offset_y + iy * q;
            z[i] = wz[k];
        }
    }
    // Shuffle (Fisher-Yates) so the input order carries no spatial locality.
    for (uint64_t t = n; t > 1; --t) {
        uint64_t j = r.below(t);
        std::swap(x[t - 1], x[j]);
        std::swap(y[t - 1], y[j]);
        std::swap(z[t - 1], z[j]);
    }
    return GF_OK;
}

} // extern "C"
