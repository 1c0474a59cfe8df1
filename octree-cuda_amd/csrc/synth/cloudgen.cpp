// Synthetic point-cloud workloads and the caller-side bounding-box step.
//
// The reference ships no datasets (its example clouds are external downloads), so benchmarks and
// tests use seeded synthetic clouds of the same sizes and record layouts. The generator is
// counter based: point i draws its random numbers from stream i of the seed, so any sub-range of
// a cloud can be generated independently and the result does not depend on the thread count.
//
// Cloud shape: a fraction of the points lies on a 2.5D heightfield z = Lz/2 + sum of four random
// sinusoids (amplitudes summing to at most 0.45 Lz), the rest in K isotropic Gaussian clusters
// (Box-Muller normals). Colours follow a height colour map plus uniform integer noise in
// [-16, 16]. Records are packed AoS: 3 x float32 + 3 x uint8 (15 bytes) or 3 x float64 + 3 x uint8
// (27 bytes), in index order (random spatial order).
//
// The bounding-box helpers restate the caller's host computation (src/boundingbox.h:10-52):
// the real BB with fmin/fmax over the promoted coordinates, then the cube
// min - (side - dim) / 2, max = min + side.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z)  // SplitMix64 finalizer (Steele, Lea, Flood 2014)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// uniform double in [0, 1) from (stream, draw)
inline double unit(uint64_t base, uint64_t stream, uint64_t draw)
{
    uint64_t z = mix64(base + stream * 0x9E3779B97F4A7C15ULL + draw * 0xD1B54A32D192ED03ULL);
    return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
}

constexpr double kTwoPi = 6.283185307179586476925286766559;

struct Params
{
    uint64_t points;
    uint32_t cloud_type;
    double box[3];
    uint32_t clusters;
    uint64_t seed;
    double offset;
    double heightfield_fraction;
    double sigma;
    uint32_t threads;
};

struct Shape
{
    double amp[4], kx[4], ky[4], phase[4];
    std::vector<double> centers;  // 3 * K
};

Shape makeShape(const Params& p, uint64_t base)
{
    Shape s;
    // cloud-level parameters come from stream 2^63 + j (disjoint from the point streams)
    const uint64_t paramStream = 0x8000000000000000ULL;
    for (int j = 0; j < 4; ++j)
    {
        s.amp[j]   = p.box[2] * (0.04 + 0.0725 * unit(base, paramStream + j, 0));
        s.kx[j]    = 0.5 + 3.5 * unit(base, paramStream + j, 1);
        s.ky[j]    = 0.5 + 3.5 * unit(base, paramStream + j, 2);
        s.phase[j] = kTwoPi * unit(base, paramStream + j, 3);
    }
    s.centers.resize(3 * static_cast<size_t>(p.clusters));
    for (uint32_t k = 0; k < p.clusters; ++k)
        for (int a = 0; a < 3; ++a)
            s.centers[3 * k + a] = p.box[a] * unit(base, paramStream + 16 + k, static_cast<uint64_t>(a));
    return s;
}

inline double clampd(double v, double lo, double hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

inline uint8_t clampColor(int v)
{
    return static_cast<uint8_t>(v < 0 ? 0 : (v > 255 ? 255 : v));
}

// height colour map: blue -> cyan -> green -> yellow -> red
inline void colorMap(double h, int& r, int& g, int& b)
{
    h = clampd(h, 0.0, 1.0);
    double seg = h * 4.0;
    int i      = std::min(3, static_cast<int>(seg));
    double f   = seg - i;
    switch (i)
    {
    case 0: r = 0; g = static_cast<int>(255 * f); b = 255; break;
    case 1: r = 0; g = 255; b = static_cast<int>(255 * (1 - f)); break;
    case 2: r = static_cast<int>(255 * f); g = 255; b = 0; break;
    default: r = 255; g = static_cast<int>(255 * (1 - f)); b = 0; break;
    }
}

void generateRange(const Params& p, const Shape& s, uint64_t base, uint64_t b, uint64_t e, uint8_t* dst)
{
    const uint32_t stride = p.cloud_type == 0 ? 15 : 27;
    for (uint64_t i = b; i < e; ++i)
    {
        double x, y, z;
        if (p.clusters == 0 || unit(base, i, 0) < p.heightfield_fraction)
        {
            x = p.box[0] * unit(base, i, 2);
            y = p.box[1] * unit(base, i, 3);
            z = p.box[2] / 2;
            for (int j = 0; j < 4; ++j)
                z += s.amp[j] * std::sin(kTwoPi * (s.kx[j] * x / p.box[0] + s.ky[j] * y / p.box[1]) + s.phase[j]);
        }
        else
        {
            uint32_t k   = std::min(p.clusters - 1, static_cast<uint32_t>(unit(base, i, 1) * p.clusters));
            double r1    = std::sqrt(-2.0 * std::log(1.0 - unit(base, i, 4)));
            double a1    = kTwoPi * unit(base, i, 5);
            double r2    = std::sqrt(-2.0 * std::log(1.0 - unit(base, i, 6)));
            double a2    = kTwoPi * unit(base, i, 7);
            const double* c = &s.centers[3 * k];
            x = c[0] + p.sigma * r1 * std::cos(a1);
            y = c[1] + p.sigma * r1 * std::sin(a1);
            z = c[2] + p.sigma * r2 * std::cos(a2);
        }
        x = clampd(x, 0.0, p.box[0]);
        y = clampd(y, 0.0, p.box[1]);
        z = clampd(z, 0.0, p.box[2]);

        int r, g, bl;
        colorMap(z / p.box[2], r, g, bl);
        r += static_cast<int>(unit(base, i, 8) * 33.0) - 16;
        g += static_cast<int>(unit(base, i, 9) * 33.0) - 16;
        bl += static_cast<int>(unit(base, i, 10) * 33.0) - 16;

        uint8_t* rec = dst + i * stride;
        if (p.cloud_type == 0)
        {
            float v[3] = {static_cast<float>(x + p.offset), static_cast<float>(y + p.offset),
                          static_cast<float>(z + p.offset)};
            std::memcpy(rec, v, 12);
            rec += 12;
        }
        else
        {
            double v[3] = {x + p.offset, y + p.offset, z + p.offset};
            std::memcpy(rec, v, 24);
            rec += 24;
        }
        rec[0] = clampColor(r);
        rec[1] = clampColor(g);
        rec[2] = clampColor(bl);
    }
}

unsigned threadsFor(uint32_t requested)
{
    if (requested)
        return requested;
    unsigned hc = std::thread::hardware_concurrency();
    return hc ? hc : 1;
}

void runParallel(uint64_t n, unsigned threads, const std::function<void(uint64_t, uint64_t, unsigned)>& fn)
{
    threads = static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(threads, n)));
    std::vector<std::thread> pool;
    uint64_t chunk = (n + threads - 1) / threads;
    for (unsigned t = 0; t < threads; ++t)
    {
        uint64_t b = t * chunk, e = std::min<uint64_t>(n, b + chunk);
        if (b >= e)
            break;
        pool.emplace_back(fn, b, e, t);
    }
    for (auto& th : pool)
        th.join();
}

}  // namespace

extern "C" {

typedef Params ocs_params;

uint32_t ocs_record_stride(uint32_t cloudType)
{
    return cloudType == 0 ? 15 : 27;
}

// Writes points [first, first + count) of the cloud described by p into dst (count * stride bytes).
int ocs_generate_range(const ocs_params* p, uint64_t first, uint64_t count, uint8_t* dst)
{
    if (!p || !dst || p->cloud_type > 1 || p->box[0] <= 0 || p->box[1] <= 0 || p->box[2] <= 0)
        return -1;
    uint64_t base = mix64(p->seed ^ 0x6A09E667F3BCC909ULL);
    Shape s       = makeShape(*p, base);
    const uint32_t stride = ocs_record_stride(p->cloud_type);
    runParallel(count, threadsFor(p->threads), [&](uint64_t b, uint64_t e, unsigned) {
        generateRange(*p, s, base, first + b, first + e, dst - first * stride);
    });
    return 0;
}

int ocs_generate(const ocs_params* p, uint8_t* dst)
{
    return ocs_generate_range(p, 0, p->points, dst);
}

// calculateRealBB<T> (src/boundingbox.h:10-33) and calculateCubicBB (:35-52)
int ocs_bounding_box(const uint8_t* cloud, uint32_t n, uint32_t stride, uint32_t cloudType, uint32_t threads,
                     double* real6, double* cubic6)
{
    if (!cloud || n == 0 || cloudType > 1)
        return -1;
    unsigned nt = threadsFor(threads);
    std::vector<double> part(6 * static_cast<size_t>(nt));
    for (unsigned t = 0; t < nt; ++t)
        for (int a = 0; a < 3; ++a)
        {
            part[6 * t + a]     = INFINITY;
            part[6 * t + 3 + a] = -INFINITY;
        }
    runParallel(n, nt, [&](uint64_t b, uint64_t e, unsigned t) {
        double* bb = &part[6 * t];
        for (uint64_t i = b; i < e; ++i)
        {
            const uint8_t* rec = cloud + i * stride;
            double v[3];
            if (cloudType == 0)
            {
                float f[3];
                std::memcpy(f, rec, 12);
                v[0] = f[0];
                v[1] = f[1];
                v[2] = f[2];
            }
            else
            {
                std::memcpy(v, rec, 24);
            }
            for (int a = 0; a < 3; ++a)
            {
                bb[a]     = std::fmin(bb[a], v[a]);
                bb[3 + a] = std::fmax(bb[3 + a], v[a]);
            }
        }
    });
    double bb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (unsigned t = 0; t < nt; ++t)
        for (int a = 0; a < 3; ++a)
        {
            bb[a]     = std::fmin(bb[a], part[6 * t + a]);
            bb[3 + a] = std::fmax(bb[3 + a], part[6 * t + 3 + a]);
        }
    for (int k = 0; k < 6; ++k)
        real6[k] = bb[k];
    double dimX = bb[3] - bb[0], dimY = bb[4] - bb[1], dimZ = bb[5] - bb[2];
    double side = std::fmax(std::fmax(dimX, dimY), dimZ);
    cubic6[0]   = bb[0] - ((side - dimX) / 2.0f);
    cubic6[1]   = bb[1] - ((side - dimY) / 2.0f);
    cubic6[2]   = bb[2] - ((side - dimZ) / 2.0f);
    cubic6[3]   = cubic6[0] + side;
    cubic6[4]   = cubic6[1] + side;
    cubic6[5]   = cubic6[2] + side;
    return 0;
}

}  // extern "C"
