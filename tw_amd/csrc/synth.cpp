// synth.cpp -- deterministic synthetic speckle pairs (host, OpenMP).
// Spot rendering is bucketed by row band so every output pixel is written by
// exactly one thread, and the summation order per pixel is the spot order:
// results are identical for any thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "twdic_synth.h"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
};

inline uint64_t mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based standard normal for pixel i (Box-Muller on two hashed uniforms).
inline double normal_at(uint64_t seed, uint64_t i)
{
    uint64_t a = mix(seed + 0x9E3779B97F4A7C15ull * (2 * i + 1));
    uint64_t b = mix(seed + 0x9E3779B97F4A7C15ull * (2 * i + 2));
    double u1 = ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);  // (0, 1)
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

struct Spot { double cx, cy, sig, amp; };

std::vector<Spot> make_spots(const twdic_speckle& sp)
{
    Rng r{sp.seed};
    std::vector<Spot> v;
    v.reserve(sp.nspots);
    for (int i = 0; i < sp.nspots; ++i) {
        Spot s;
        s.cx = r.uniform() * sp.w;
        s.cy = r.uniform() * sp.h;
        s.sig = sp.sigma_min + r.uniform() * (sp.sigma_max - sp.sigma_min);
        s.amp = sp.amp_min + r.uniform() * (sp.amp_max - sp.amp_min);
        bool drop = false;
        for (int b = 0; b < sp.nboxes && !drop; ++b) {
            const int32_t* bx = sp.boxes + 4 * b;
            drop = s.cx >= bx[0] && s.cx < bx[2] && s.cy >= bx[1] && s.cy < bx[3];
        }
        if (!drop) v.push_back(s);
    }
    return v;
}

inline uint8_t quantise(double x)
{
    x = std::min(255.0, std::max(0.0, x));
    return (uint8_t)std::floor(x + 0.5);
}

// Inverse of the forward map x' = phi(X) = X + u(X).
struct InverseMap {
    const twdic_deform& d;
    double a00, a01, a10, a11;  // inverse of [[1+exx, exy], [exy, 1+eyy]]
    explicit InverseMap(const twdic_deform& dd) : d(dd)
    {
        double m00 = 1 + d.exx, m01 = d.exy, m10 = d.exy, m11 = 1 + d.eyy;
        double det = m00 * m11 - m01 * m10;
        a00 = m11 / det; a01 = -m01 / det; a10 = -m10 / det; a11 = m00 / det;
    }
    void operator()(double xp, double yp, double& X, double& Y) const
    {
        if (d.kind == TWDIC_DEFORM_AFFINE) {
            double bx = xp - d.xc - d.u0, by = yp - d.yc - d.v0;
            X = d.xc + a00 * bx + a01 * by;
            Y = d.yc + a10 * bx + a11 * by;
        } else {  // sinusoid: separable, fixed-point iteration (|du/dx| << 1)
            X = xp; Y = yp;
            for (int k = 0; k < 30; ++k) {
                X = xp - d.amp_u * std::sin(2.0 * M_PI * X / d.per_u);
                Y = yp - d.amp_v * std::cos(2.0 * M_PI * Y / d.per_v);
            }
        }
    }
    void forward(double X, double Y, double& xp, double& yp) const
    {
        double o[6];
        twdic_synth_displacement(&d, X, Y, o);
        xp = X + o[0];
        yp = Y + o[3];
    }
};

// Accumulate the spot field at positions given by `pos(x, y) -> (X, Y)` for
// output pixels; spot footprints are mapped to output space through fwd().
template <class Pos, class Fwd>
void render(const std::vector<Spot>& spots, int w, int h, double grow, Pos pos, Fwd fwd, std::vector<double>& acc)
{
    acc.assign((size_t)w * h, 0.0);
    const int band = 16;
    const int nb = (h + band - 1) / band;
    // bucket spots by output-space row band of their footprint
    std::vector<std::vector<int>> buckets(nb);
    for (size_t i = 0; i < spots.size(); ++i) {
        double xp, yp;
        fwd(spots[i].cx, spots[i].cy, xp, yp);
        double r = 5.0 * spots[i].sig * grow + 2.0;
        int b0 = std::max(0, (int)std::floor((yp - r) / band));
        int b1 = std::min(nb - 1, (int)std::floor((yp + r) / band));
        for (int b = b0; b <= b1; ++b) buckets[b].push_back((int)i);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b) {
        const int y0 = b * band, y1 = std::min(h, y0 + band);
        for (int idx : buckets[b]) {
            const Spot& s = spots[idx];
            double xp, yp;
            fwd(s.cx, s.cy, xp, yp);
            double r = 5.0 * s.sig * grow + 2.0;
            int ya = std::max(y0, (int)std::floor(yp - r)), yb = std::min(y1 - 1, (int)std::ceil(yp + r));
            int xa = std::max(0, (int)std::floor(xp - r)), xb = std::min(w - 1, (int)std::ceil(xp + r));
            const double k = -0.5 / (s.sig * s.sig);
            for (int y = ya; y <= yb; ++y)
                for (int x = xa; x <= xb; ++x) {
                    double X, Y;
                    pos(x, y, X, Y);
                    double dx = X - s.cx, dy = Y - s.cy;
                    acc[(size_t)y * w + x] += s.amp * std::exp(k * (dx * dx + dy * dy));
                }
        }
    }
}

inline void cr_weights(double t, double w[4])
{
    double t2 = t * t, t3 = t2 * t;
    w[0] = 0.5 * (-t3 + 2 * t2 - t);
    w[1] = 0.5 * (3 * t3 - 5 * t2 + 2);
    w[2] = 0.5 * (-3 * t3 + 4 * t2 + t);
    w[3] = 0.5 * (t3 - t2);
}

}  // namespace

extern "C" {

void twdic_synth_displacement(const twdic_deform* d, double x, double y, double o[6])
{
    if (d->kind == TWDIC_DEFORM_SINUSOID) {
        double ax = 2.0 * M_PI / d->per_u, ay = 2.0 * M_PI / d->per_v;
        o[0] = d->amp_u * std::sin(ax * x);
        o[1] = d->amp_u * ax * std::cos(ax * x);
        o[2] = 0.0;
        o[3] = d->amp_v * std::cos(ay * y);
        o[4] = 0.0;
        o[5] = -d->amp_v * ay * std::sin(ay * y);
    } else if (d->kind == TWDIC_DEFORM_RESAMPLE) {
        o[0] = d->u0; o[1] = 0; o[2] = 0; o[3] = d->v0; o[4] = 0; o[5] = 0;
    } else {
        o[0] = d->u0 + d->exx * (x - d->xc) + d->exy * (y - d->yc);
        o[1] = d->exx;
        o[2] = d->exy;
        o[3] = d->v0 + d->exy * (x - d->xc) + d->eyy * (y - d->yc);
        o[4] = d->exy;
        o[5] = d->eyy;
    }
}

int twdic_synth_pair(const twdic_speckle* sp, const twdic_deform* d, uint8_t* ref, uint8_t* tar, int stride)
{
    if (!sp || sp->w <= 0 || sp->h <= 0 || stride < sp->w || sp->nspots < 0) return -1;
    if (tar && !d) return -1;
    const int w = sp->w, h = sp->h;
    std::vector<Spot> spots = make_spots(*sp);
    std::vector<double> acc;
    auto ident_pos = [](int x, int y, double& X, double& Y) { X = x; Y = y; };
    auto ident_fwd = [](double X, double Y, double& xp, double& yp) { xp = X; yp = Y; };
    std::vector<uint8_t> refbuf;
    const bool need_ref = ref || (tar && d->kind == TWDIC_DEFORM_RESAMPLE);
    if (need_ref) {
        render(spots, w, h, 1.0, ident_pos, ident_fwd, acc);
        refbuf.resize((size_t)w * h);
#pragma omp parallel for
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) refbuf[(size_t)y * w + x] = quantise(acc[(size_t)y * w + x]);
        if (ref)
            for (int y = 0; y < h; ++y) std::memcpy(ref + (size_t)y * stride, &refbuf[(size_t)y * w], w);
    }
    if (!tar) return 0;
    if (d->kind == TWDIC_DEFORM_RESAMPLE) {
        // Catmull-Rom resample of the rounded reference at (x - u0, y - v0),
        // edge pixels clamped
        acc.assign((size_t)w * h, 0.0);
#pragma omp parallel for
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                double X = x - d->u0, Y = y - d->v0;
                int ix = (int)std::floor(X), iy = (int)std::floor(Y);
                double wx[4], wy[4];
                cr_weights(X - ix, wx);
                cr_weights(Y - iy, wy);
                double s = 0;
                for (int r = 0; r < 4; ++r) {
                    int yy = std::min(h - 1, std::max(0, iy - 1 + r));
                    double row = 0;
                    for (int c = 0; c < 4; ++c) {
                        int xx = std::min(w - 1, std::max(0, ix - 1 + c));
                        row += wx[c] * refbuf[(size_t)yy * w + xx];
                    }
                    s += wy[r] * row;
                }
                acc[(size_t)y * w + x] = s;
            }
    } else {
        InverseMap inv(*d);
        double grow = 1.0 + std::fabs(d->exx) + std::fabs(d->eyy) + 2 * std::fabs(d->exy);
        if (d->kind == TWDIC_DEFORM_SINUSOID)
            grow = 1.0 + 2.0 * M_PI * (std::fabs(d->amp_u) / d->per_u + std::fabs(d->amp_v) / d->per_v);
        // the inverse map of the sinusoid is separable: tabulate it per column/row
        std::vector<double> invx, invy;
        if (d->kind == TWDIC_DEFORM_SINUSOID) {
            invx.resize(w); invy.resize(h);
            for (int x = 0; x < w; ++x) { double X, Y; inv((double)x, 0.0, X, Y); invx[x] = X; }
            for (int y = 0; y < h; ++y) { double X, Y; inv(0.0, (double)y, X, Y); invy[y] = Y; }
            auto pos = [&](int x, int y, double& X, double& Y) { X = invx[x]; Y = invy[y]; };
            auto fwd = [&](double X, double Y, double& xp, double& yp) { inv.forward(X, Y, xp, yp); };
            render(spots, w, h, grow, pos, fwd, acc);
        } else {
            auto pos = [&](int x, int y, double& X, double& Y) { inv((double)x, (double)y, X, Y); };
            auto fwd = [&](double X, double Y, double& xp, double& yp) { inv.forward(X, Y, xp, yp); };
            render(spots, w, h, grow, pos, fwd, acc);
        }
    }
#pragma omp parallel for
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            double vv = acc[(size_t)y * w + x];
            if (sp->noise_sigma > 0) vv += sp->noise_sigma * normal_at(sp->noise_seed, (uint64_t)y * w + x);
            tar[(size_t)y * stride + x] = quantise(vv);
        }
    return 0;
}

}  // extern "C"
