// bsf_oracle.cpp -- CPU oracle (TEST INFRASTRUCTURE, never the product path).
//
// A scalar restatement of the reference's Java terrain semantics, written to be
// read side by side with the cited Java methods.  It is deliberately naive: whole
// arrays, Java loop order, Java operand order.  It exists only to check the HIP
// worker; tests/, the smoke entry and bench.py's CPU-baseline leg are its only users.
//
// Build flags (see oracle/Makefile): -O2 -ffp-contract=off -fno-fast-math.
#include "bsf_oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// java.util.Random (JDK 17 Javadoc algorithm; call sites
// NoiseGeneratorPerlin173.java:63-78, NoiseGenerator2173.java:27-48,
// BetaTerrain173.java:84, WorldChunkManager173.java:25-27).
// ---------------------------------------------------------------------------
struct JRandom {
    uint64_t s = 0;
    explicit JRandom(int64_t seed) { setSeed(seed); }
    void setSeed(int64_t seed) { s = (static_cast<uint64_t>(seed) ^ 0x5DEECE66DULL) & ((1ULL << 48) - 1); }
    int32_t next(int bits) {
        s = (s * 0x5DEECE66DULL + 0xBULL) & ((1ULL << 48) - 1);
        return static_cast<int32_t>(static_cast<int64_t>(s) >> (48 - bits));
    }
    int32_t nextInt(int32_t bound) {
        // Random.nextInt(int bound), JDK 17 form.
        int32_t r = next(31);
        int32_t m = bound - 1;
        if ((bound & m) == 0) {
            r = static_cast<int32_t>((static_cast<int64_t>(bound) * static_cast<int64_t>(r)) >> 31);
        } else {
            for (int32_t u = r; ; u = next(31)) {
                r = u % bound;
                // u - r + m < 0 in Java int arithmetic (wrapping)
                uint32_t w = static_cast<uint32_t>(u) - static_cast<uint32_t>(r) + static_cast<uint32_t>(m);
                if (static_cast<int32_t>(w) >= 0) break;
            }
        }
        return r;
    }
    double nextDouble() {
        int64_t hi = static_cast<int64_t>(next(26));
        int64_t lo = static_cast<int64_t>(next(27));
        return static_cast<double>((hi << 27) + lo) * 0x1.0p-53;
    }
};

// Java (int) cast of a double: truncation, NaN -> 0, saturating.
inline int32_t jd2i(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return -2147483647 - 1;
    return static_cast<int32_t>(v);
}

// ---------------------------------------------------------------------------
// NoiseGeneratorPerlin173 (src/beta173/noise/NoiseGeneratorPerlin173.java)
// ---------------------------------------------------------------------------
struct Perlin {
    int d[512];
    double a, b, c;

    // reseed / constructor, :63-78
    void reseed(JRandom& r) {
        a = r.nextDouble() * 256.0;
        b = r.nextDouble() * 256.0;
        c = r.nextDouble() * 256.0;
        for (int i = 0; i < 256; ++i) d[i] = i;
        for (int i = 0; i < 256; ++i) {
            int j = r.nextInt(256 - i) + i;
            int k = d[i];
            d[i] = d[j];
            d[j] = k;
            d[i + 256] = d[i];
        }
    }
    // consumeConstructorRandom, :81-88
    static void consume(JRandom& r) {
        r.nextDouble(); r.nextDouble(); r.nextDouble();
        for (int i = 0; i < 256; ++i) r.nextInt(256 - i);
    }

    static double lerp(double t, double x, double y) { return x + t * (y - x); }  // b(...), :119-121

    // a(int, double, double) -- the 2-argument legacy gradient, :123-143
    static double grad2(int h, double x, double z) {
        double zeroX = 0.0 * x;
        switch (h & 15) {
            case 0: return x + 0.0;
            case 1: return -x + 0.0;
            case 2: return x + -0.0;
            case 3: return -x + -0.0;
            case 4: return x + z;
            case 5: return -x + z;
            case 6: return x + -z;
            case 7: return -x + -z;
            case 8: return zeroX + z;
            case 9: return -zeroX + z;
            case 10: return zeroX + -z;
            case 11: return -zeroX + -z;
            case 12: return zeroX + x;
            case 13: return -zeroX + z;
            case 14: return zeroX + -x;
            default: return -zeroX + -z;
        }
    }
    // a(int, double, double, double) -- 3D gradient, :145-165
    static double grad3(int h, double x, double y, double z) {
        switch (h & 15) {
            case 0: return x + y;
            case 1: return -x + y;
            case 2: return x + -y;
            case 3: return -x + -y;
            case 4: return x + z;
            case 5: return -x + z;
            case 6: return x + -z;
            case 7: return -x + -z;
            case 8: return y + z;
            case 9: return -y + z;
            case 10: return y + -z;
            case 11: return -y + -z;
            case 12: return y + x;
            case 13: return -y + z;
            case 14: return y + -x;
            default: return -y + -z;
        }
    }

    // Single point, a(double,double,double), :90-113
    double point(double x0, double y0, double z0) const {
        double x = x0 + a, y = y0 + b, z = z0 + c;
        int fx = jd2i(x), fy = jd2i(y), fz = jd2i(z);
        if (x < static_cast<double>(fx)) --fx;
        if (y < static_cast<double>(fy)) --fy;
        if (z < static_cast<double>(fz)) --fz;
        int xi = fx & 255, yi = fy & 255, zi = fz & 255;
        x -= static_cast<double>(fx);
        y -= static_cast<double>(fy);
        z -= static_cast<double>(fz);
        double u = x * x * x * (x * (x * 6.0 - 15.0) + 10.0);
        double v = y * y * y * (y * (y * 6.0 - 15.0) + 10.0);
        double w = z * z * z * (z * (z * 6.0 - 15.0) + 10.0);
        int p25 = d[xi] + yi;
        int p26 = d[p25] + zi;
        int p27 = d[p25 + 1] + zi;
        int p28 = d[xi + 1] + yi;
        int p29 = d[p28] + zi;
        int p30 = d[p28 + 1] + zi;
        return lerp(w,
                lerp(v, lerp(u, grad3(d[p26], x, y, z), grad3(d[p29], x - 1.0, y, z)),
                        lerp(u, grad3(d[p27], x, y - 1.0, z), grad3(d[p30], x - 1.0, y - 1.0, z))),
                lerp(v, lerp(u, grad3(d[p26 + 1], x, y, z - 1.0), grad3(d[p29 + 1], x - 1.0, y, z - 1.0)),
                        lerp(u, grad3(d[p27 + 1], x, y - 1.0, z - 1.0), grad3(d[p30 + 1], x - 1.0, y - 1.0, z - 1.0))));
    }

    // The legacy bulk method a(double[], ...), :172-284.  out[idx] += value * (1/amp).
    void add(double* out, double fromX, double fromY, double fromZ, int xLen, int yLen, int zLen,
             double sx, double sy, double sz, double amp) const {
        if (yLen == 1) {
            double inv = 1.0 / amp;
            int k = 0;
            for (int ix = 0; ix < xLen; ++ix) {
                double x = (fromX + static_cast<double>(ix)) * sx + a;
                int fx = jd2i(x);
                if (x < static_cast<double>(fx)) --fx;
                int xi = fx & 255;
                x -= static_cast<double>(fx);
                double u = x * x * x * (x * (x * 6.0 - 15.0) + 10.0);
                for (int iz = 0; iz < zLen; ++iz) {
                    double z = (fromZ + static_cast<double>(iz)) * sz + c;
                    int fz = jd2i(z);
                    if (z < static_cast<double>(fz)) --fz;
                    int zi = fz & 255;
                    z -= static_cast<double>(fz);
                    double w = z * z * z * (z * (z * 6.0 - 15.0) + 10.0);
                    int p19 = d[xi] + 0;
                    int p47 = d[p19] + zi;
                    int p48 = d[xi + 1] + 0;
                    int p20 = d[p48] + zi;
                    double q0 = lerp(u, grad2(d[p47], x, z), grad3(d[p20], x - 1.0, 0.0, z));
                    double q1 = lerp(u, grad3(d[p47 + 1], x, 0.0, z - 1.0), grad3(d[p20 + 1], x - 1.0, 0.0, z - 1.0));
                    double val = lerp(w, q0, q1);
                    out[k++] += val * inv;
                }
            }
            return;
        }
        int k = 0;
        double inv = 1.0 / amp;
        int prevYi = -1;
        double q00 = 0, q01 = 0, q10 = 0, q11 = 0;
        for (int ix = 0; ix < xLen; ++ix) {
            double x = (fromX + static_cast<double>(ix)) * sx + a;
            int fx = jd2i(x);
            if (x < static_cast<double>(fx)) --fx;
            int xi = fx & 255;
            x -= static_cast<double>(fx);
            double u = x * x * x * (x * (x * 6.0 - 15.0) + 10.0);
            for (int iz = 0; iz < zLen; ++iz) {
                double z = (fromZ + static_cast<double>(iz)) * sz + c;
                int fz = jd2i(z);
                if (z < static_cast<double>(fz)) --fz;
                int zi = fz & 255;
                z -= static_cast<double>(fz);
                double w = z * z * z * (z * (z * 6.0 - 15.0) + 10.0);
                for (int iy = 0; iy < yLen; ++iy) {
                    double y = (fromY + static_cast<double>(iy)) * sy + b;
                    int fy = jd2i(y);
                    if (y < static_cast<double>(fy)) --fy;
                    int yi = fy & 255;
                    y -= static_cast<double>(fy);
                    double v = y * y * y * (y * (y * 6.0 - 15.0) + 10.0);
                    if (iy == 0 || yi != prevYi) {
                        prevYi = yi;
                        int p60 = d[xi] + yi;
                        int p61 = d[p60] + zi;
                        int p62 = d[p60 + 1] + zi;
                        int p63 = d[xi + 1] + yi;
                        int p33 = d[p63] + zi;
                        int p64 = d[p63 + 1] + zi;
                        q00 = lerp(u, grad3(d[p61], x, y, z), grad3(d[p33], x - 1.0, y, z));
                        q01 = lerp(u, grad3(d[p62], x, y - 1.0, z), grad3(d[p64], x - 1.0, y - 1.0, z));
                        q10 = lerp(u, grad3(d[p61 + 1], x, y, z - 1.0), grad3(d[p33 + 1], x - 1.0, y, z - 1.0));
                        q11 = lerp(u, grad3(d[p62 + 1], x, y - 1.0, z - 1.0), grad3(d[p64 + 1], x - 1.0, y - 1.0, z - 1.0));
                    }
                    double r0 = lerp(v, q00, q01);
                    double r1 = lerp(v, q10, q11);
                    double val = lerp(w, r0, r1);
                    out[k++] += val * inv;
                }
            }
        }
    }
};

// NoiseGeneratorOctaves173 (NoiseGeneratorOctaves173.java)
struct Octaves {
    std::vector<Perlin> g;
    void build(JRandom& r, int n) {
        g.resize(n);
        for (int i = 0; i < n; ++i) g[i].reseed(r);
    }
    // generateNoise(double[], double x3, int lens, double scales), :46-83
    void noise(double* out, double fx, double fy, double fz, int xl, int yl, int zl,
               double sx, double sy, double sz) const {
        for (int i = 0; i < xl * yl * zl; ++i) out[i] = 0.0;
        double amp = 1.0;
        for (size_t o = 0; o < g.size(); ++o) {
            g[o].add(out, fx, fy, fz, xl, yl, zl, sx * amp, sy * amp, sz * amp, amp);
            amp /= 2.0;
        }
    }
    // generateNoise(double[], int, int, int, int, double, double, double), :85-90
    void noise2d(double* out, int fx, int fz, int xl, int zl, double sx, double sz) const {
        noise(out, static_cast<double>(fx), 10.0, static_cast<double>(fz), xl, 1, zl, sx, 1.0, sz);
    }
};

// ---------------------------------------------------------------------------
// NoiseGenerator2173 (simplex), NoiseGenerator2173.java
// ---------------------------------------------------------------------------
const double kSimplexF = 0.5 * (std::sqrt(3.0) - 1.0);   // :13
const double kSimplexG = (3.0 - std::sqrt(3.0)) / 6.0;   // :14

struct Simplex {
    int e[512];
    double a, b, c;
    void reseed(JRandom& r) {  // :27-48
        a = r.nextDouble() * 256.0;
        b = r.nextDouble() * 256.0;
        c = r.nextDouble() * 256.0;
        for (int i = 0; i < 256; ++i) e[i] = i;
        for (int i = 0; i < 256; ++i) {
            int j = r.nextInt(256 - i) + i;
            int k = e[i];
            e[i] = e[j];
            e[j] = k;
            e[i + 256] = e[i];
        }
    }
    static int ffloor(double v) { return v > 0.0 ? jd2i(v) : jd2i(v) - 1; }  // a(double), :50-52
    static double grad(int g, double x, double z) {  // gradient2, :58-72
        switch (g) {
            case 0: return 1.0 * x + 1.0 * z;
            case 1: return -1.0 * x + 1.0 * z;
            case 2: return 1.0 * x + -1.0 * z;
            case 3: return -1.0 * x + -1.0 * z;
            case 4:
            case 6: return 1.0 * x + 0.0 * z;
            case 5:
            case 7: return -1.0 * x + 0.0 * z;
            case 8:
            case 10: return 0.0 * x + 1.0 * z;
            default: return 0.0 * x + -1.0 * z;
        }
    }
    // one sample at already-offset coordinates d5 (x) / d6 (z), :117-172
    double sample(double d5, double d6) const {
        const double f = kSimplexF, g = kSimplexG;
        double d7 = (d5 + d6) * f;
        int j1 = ffloor(d5 + d7);
        int k1 = ffloor(d6 + d7);
        double d8 = static_cast<double>(j1 + k1) * g;
        double d9 = static_cast<double>(j1) - d8;
        double d10 = static_cast<double>(k1) - d8;
        double d11 = d5 - d9;
        double d12 = d6 - d10;
        int b0, b1;
        if (d11 > d12) { b0 = 1; b1 = 0; } else { b0 = 0; b1 = 1; }
        double d13 = d11 - static_cast<double>(b0) + g;
        double d14 = d12 - static_cast<double>(b1) + g;
        double d15 = d11 - 1.0 + 2.0 * g;
        double d16 = d12 - 1.0 + 2.0 * g;
        int l1 = j1 & 255;
        int i2 = k1 & 255;
        int j2 = e[l1 + e[i2]] % 12;
        int k2 = e[l1 + b0 + e[i2 + b1]] % 12;
        int l2 = e[l1 + 1 + e[i2 + 1]] % 12;
        double d17 = 0.5 - d11 * d11 - d12 * d12, d18;
        if (d17 < 0.0) d18 = 0.0; else { d17 *= d17; d18 = d17 * d17 * grad(j2, d11, d12); }
        double d19 = 0.5 - d13 * d13 - d14 * d14, d20;
        if (d19 < 0.0) d20 = 0.0; else { d19 *= d19; d20 = d19 * d19 * grad(k2, d13, d14); }
        double d21 = 0.5 - d15 * d15 - d16 * d16, d22;
        if (d21 < 0.0) d22 = 0.0; else { d21 *= d21; d22 = d21 * d21 * grad(l2, d15, d16); }
        return 70.0 * (d18 + d20 + d22);
    }
    // aStrided, :86-176
    void addStrided(double* out, double d0, double d1, int ni, int nj, int strideX, int strideZ,
                    double d2, double d3, double d4) const {
        int k = 0;
        for (int l = 0; l < ni; ++l) {
            double d5 = (d0 + static_cast<double>(l * strideX)) * d2 + a;
            for (int i1 = 0; i1 < nj; ++i1) {
                double d6 = (d1 + static_cast<double>(i1 * strideZ)) * d3 + b;
                out[k++] += sample(d5, d6) * d4;
            }
        }
    }
};

// NoiseGeneratorOctaves2173.aStrided, NoiseGeneratorOctaves2173.java:66-104
struct SimplexOctaves {
    std::vector<Simplex> g;
    void build(JRandom& r, int n) { g.resize(n); for (int i = 0; i < n; ++i) g[i].reseed(r); }
    void noise(double* out, double d0, double d1, int ni, int nj, int strideX, int strideZ,
               double d2, double d3, double d4, double d5 = 0.5) const {
        d2 /= 1.5;
        d3 /= 1.5;
        for (int k = 0; k < ni * nj; ++k) out[k] = 0.0;
        double d6 = 1.0, d7 = 1.0;
        for (size_t l = 0; l < g.size(); ++l) {
            g[l].addStrided(out, d0, d1, ni, nj, strideX, strideZ, d2 * d7, d3 * d7, 0.55 / d6);
            d7 *= d4;
            d6 *= d5;
        }
    }
};

// WorldChunkManager173 (WorldChunkManager173.java)
struct Climate {
    SimplexOctaves e, f, g;
    explicit Climate(int64_t seed) {
        JRandom re(static_cast<int64_t>(static_cast<uint64_t>(seed) * 9871ULL));
        JRandom rf(static_cast<int64_t>(static_cast<uint64_t>(seed) * 39811ULL));
        JRandom rg(static_cast<int64_t>(static_cast<uint64_t>(seed) * 543321ULL));
        e.build(re, 4);
        f.build(rf, 4);
        g.build(rg, 2);
    }
    // getBiomeNoise (:85-130) generalised to getClimateNoiseStrided (:140-219): both run the same
    // per-point arithmetic; a dense call is stride 1.
    void sample(int startX, int startZ, int sizeX, int sizeZ, int strideX, int strideZ,
                double* temp, double* rain) const {
        const int total = sizeX * sizeZ;
        std::vector<double> cc(total);
        e.noise(temp, startX, startZ, sizeX, sizeZ, strideX, strideZ, 0.02500000037252903, 0.02500000037252903, 0.25);
        f.noise(rain, startX, startZ, sizeX, sizeZ, strideX, strideZ, 0.05000000074505806, 0.05000000074505806, 0.3333333333333333);
        g.noise(cc.data(), startX, startZ, sizeX, sizeZ, strideX, strideZ, 0.25, 0.25, 0.5882352941176471);
        for (int i = 0; i < total; ++i) {
            double d0 = cc[i] * 1.1 + 0.5;
            double d1 = 0.01;
            double d2 = 1.0 - d1;
            double d3 = (temp[i] * 0.15 + 0.7) * d2 + d0 * d1;
            d1 = 0.0020;
            d2 = 1.0 - d1;
            double d4 = (rain[i] * 0.15 + 0.5) * d2 + d0 * d1;
            d3 = 1.0 - (1.0 - d3) * (1.0 - d3);
            if (d3 < 0.0) d3 = 0.0;
            if (d4 < 0.0) d4 = 0.0;
            if (d3 > 1.0) d3 = 1.0;
            if (d4 > 1.0) d4 = 1.0;
            temp[i] = d3;
            rain[i] = d4;
        }
    }
};

// ---------------------------------------------------------------------------
// BetaTerrain173 (src/beta173/BetaTerrain173.java)
// ---------------------------------------------------------------------------
struct Terrain {
    Climate climate;
    Octaves n2, n3, n1, n4, n5;
    // constructor order, :77-93
    explicit Terrain(int64_t seed) : climate(seed) {
        JRandom r(seed);
        n2.build(r, 16);
        n3.build(r, 16);
        n1.build(r, 8);
        for (int i = 0; i < 8; ++i) Perlin::consume(r);  // two 4-octave generators, never read
        n4.build(r, 10);
        n5.build(r, 16);
    }

    // generateTerrainNoise, :667-834.  temps/rain are the 16x16 dense chunk climate arrays
    // (index x*16+z) when climateStride==3 style per-chunk sampling is used, or the
    // per-column climate (index column) when perColumn is set.
    void terrainNoise(double* out, int fromX, int fromY, int fromZ, int xLen, int yLen, int zLen,
                      const double* temps, const double* rain, bool perColumn,
                      double* o1 = nullptr, double* o2 = nullptr, double* o3 = nullptr,
                      double* o4 = nullptr, double* o5 = nullptr) const {
        const double d0 = 684.412, d1 = 684.412;
        std::vector<double> t4(xLen * zLen), t5(xLen * zLen), t1(xLen * yLen * zLen), t2(xLen * yLen * zLen),
                t3(xLen * yLen * zLen);
        n4.noise2d(t4.data(), fromX, fromZ, xLen, zLen, 1.121, 1.121);
        n5.noise2d(t5.data(), fromX, fromZ, xLen, zLen, 200.0, 200.0);
        n1.noise(t1.data(), fromX, fromY, fromZ, xLen, yLen, zLen, d0 / 80.0, d1 / 160.0, d0 / 80.0);
        n2.noise(t2.data(), fromX, fromY, fromZ, xLen, yLen, zLen, d0, d1, d0);
        n3.noise(t3.data(), fromX, fromY, fromZ, xLen, yLen, zLen, d0, d1, d0);
        if (o1) std::copy(t1.begin(), t1.end(), o1);
        if (o2) std::copy(t2.begin(), t2.end(), o2);
        if (o3) std::copy(t3.begin(), t3.end(), o3);
        if (o4) std::copy(t4.begin(), t4.end(), o4);
        if (o5) std::copy(t5.begin(), t5.end(), o5);
        int k1 = 0, l1 = 0;
        int i2 = 16 / xLen;
        for (int j2 = 0; j2 < xLen; ++j2) {
            int k2 = j2 * i2 + i2 / 2;
            for (int l2 = 0; l2 < zLen; ++l2) {
                int i3 = l2 * i2 + i2 / 2;
                double dd2 = perColumn ? temps[l1] : temps[k2 * 16 + i3];
                double dd3 = (perColumn ? rain[l1] : rain[k2 * 16 + i3]) * dd2;
                double d4 = 1.0 - dd3;
                d4 *= d4;
                d4 *= d4;
                d4 = 1.0 - d4;
                double d5 = (t4[l1] + 256.0) / 512.0;
                d5 *= d4;
                if (d5 > 1.0) d5 = 1.0;
                double d6 = t5[l1] / 8000.0;
                if (d6 < 0.0) d6 = -d6 * 0.3;
                d6 = d6 * 3.0 - 2.0;
                if (d6 < 0.0) {
                    d6 /= 2.0;
                    if (d6 < -1.0) d6 = -1.0;
                    d6 /= 1.4;
                    d6 /= 2.0;
                    d5 = 0.0;
                } else {
                    if (d6 > 1.0) d6 = 1.0;
                    d6 /= 8.0;
                }
                if (d5 < 0.0) d5 = 0.0;
                d5 += 0.5;
                d6 = d6 * static_cast<double>(yLen) / 16.0;
                double d7 = static_cast<double>(yLen) / 2.0 + d6 * 4.0;
                ++l1;
                for (int j3 = 0; j3 < yLen; ++j3) {
                    double d8;
                    double d9 = (static_cast<double>(j3) - d7) * 12.0 / d5;
                    if (d9 < 0.0) d9 *= 4.0;
                    double d10 = t2[k1] / 512.0;
                    double d11 = t3[k1] / 512.0;
                    double d12 = (t1[k1] / 10.0 + 1.0) / 2.0;
                    if (d12 < 0.0) d8 = d10;
                    else if (d12 > 1.0) d8 = d11;
                    else d8 = d10 + (d11 - d10) * d12;
                    d8 -= d9;
                    if (j3 > yLen - 4) {
                        double d13 = static_cast<double>(static_cast<float>(j3 - (yLen - 4)) / 3.0F);
                        d8 = d8 * (1.0 - d13) + -10.0 * d13;
                    }
                    out[k1] = d8;
                    ++k1;
                }
            }
        }
    }

    void chunkClimate(int cx, int cz, double* temp, double* rain) const {
        climate.sample(cx * 16, cz * 16, 16, 16, 1, 1, temp, rain);  // getBiomeNoise(.., 16, 16)
    }

    void chunkDensity(int cx, int cz, double* out) const {
        double temp[256], rain[256];
        chunkClimate(cx, cz, temp, rain);
        terrainNoise(out, cx * 4, 0, cz * 4, 5, 17, 5, temp, rain, false);
    }
};

// generateBareTerrain incremental upsampler, BetaTerrain173.java:582-665 (stone only;
// WATER/ICE never affect the stone mask).
template <class Sink>
void upsample(const double* noise, Sink&& sink) {
    const int l = 5, b2 = 17;
    for (int i1 = 0; i1 < 4; ++i1) {
        for (int j1 = 0; j1 < 4; ++j1) {
            for (int k1 = 0; k1 < 16; ++k1) {
                double d0 = 0.125;
                double d1 = noise[((i1 + 0) * l + j1 + 0) * b2 + k1 + 0];
                double d2 = noise[((i1 + 0) * l + j1 + 1) * b2 + k1 + 0];
                double d3 = noise[((i1 + 1) * l + j1 + 0) * b2 + k1 + 0];
                double d4 = noise[((i1 + 1) * l + j1 + 1) * b2 + k1 + 0];
                double d5 = (noise[((i1 + 0) * l + j1 + 0) * b2 + k1 + 1] - d1) * d0;
                double d6 = (noise[((i1 + 0) * l + j1 + 1) * b2 + k1 + 1] - d2) * d0;
                double d7 = (noise[((i1 + 1) * l + j1 + 0) * b2 + k1 + 1] - d3) * d0;
                double d8 = (noise[((i1 + 1) * l + j1 + 1) * b2 + k1 + 1] - d4) * d0;
                for (int l1 = 0; l1 < 8; ++l1) {
                    double d9 = 0.25;
                    double d10 = d1, d11 = d2;
                    double d12 = (d3 - d1) * d9;
                    double d13 = (d4 - d2) * d9;
                    for (int i2 = 0; i2 < 4; ++i2) {
                        int x = i2 + i1 * 4;
                        int y = k1 * 8 + l1;
                        double d14 = 0.25;
                        double d15 = d10;
                        double d16 = (d11 - d10) * d14;
                        for (int k2 = 0; k2 < 4; ++k2) {
                            int z = j1 * 4 + k2;
                            sink((x * 16 + z) * 128 + y, d15);
                            d15 += d16;
                        }
                        d10 += d12;
                        d11 += d13;
                    }
                    d1 += d5;
                    d2 += d6;
                    d3 += d7;
                    d4 += d8;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// IslandSearchEngine exact scan (src/IslandSearchEngine.java)
// ---------------------------------------------------------------------------
constexpr int kIslandFloorY = 68;  // ISLAND_FLOOR_Y, :74

struct Component {
    int blocks = 0, columns = 0;
    int minWorldX = 2147483647, maxWorldX = -2147483647 - 1;
    int minWorldZ = 2147483647, maxWorldZ = -2147483647 - 1;
    int minY = 2147483647, maxY = -2147483647 - 1;
    int64_t sumWorldX = 0, sumWorldZ = 0;
    int centerX = 0, centerZ = 0;
    bool touchesBottom = false, touchesSide = false;
};

// Java Math.round(double) for the small magnitudes here: floor(a + 1/2) evaluated exactly.
int64_t javaRound(double a) {
    double f = std::floor(a);
    int64_t r = static_cast<int64_t>(f);
    if (a - f >= 0.5) ++r;
    return r;
}

// floodFillFlat, :4096-4202; index3 = (x*sizeY + y)*sizeZ + z, :4225-4227
Component floodFill(const uint8_t* stone, uint8_t* visited, std::vector<int>& queue, int* columnSeen, int mark,
                    int sx, int sy, int sz, int sizeX, int sizeY, int sizeZ, int minWX, int minWZ, int floodMinY) {
    const int yzArea = sizeY * sizeZ;
    const int offX = yzArea, offY = sizeZ, offZ = 1;
    int head = 0, tail = 0;
    int start = (sx * sizeY + sy) * sizeZ + sz;
    queue[tail++] = start;
    visited[start] = 1;
    Component c;
    auto tryAdd = [&](int idx) {
        if (visited[idx] != 0) return;
        if (stone[idx] == 0) return;
        visited[idx] = 1;
        queue[tail++] = idx;
    };
    while (head < tail) {
        int index = queue[head++];
        int x = index / yzArea;
        int rem = index - x * yzArea;
        int y = rem / sizeZ;
        int z = rem - y * sizeZ;
        int wx = minWX + x, wz = minWZ + z;
        c.blocks++;
        int col = x * sizeZ + z;
        if (columnSeen[col] != mark) {
            columnSeen[col] = mark;
            c.columns++;
            c.sumWorldX += wx;
            c.sumWorldZ += wz;
        }
        c.minWorldX = std::min(c.minWorldX, wx);
        c.maxWorldX = std::max(c.maxWorldX, wx);
        c.minWorldZ = std::min(c.minWorldZ, wz);
        c.maxWorldZ = std::max(c.maxWorldZ, wz);
        c.minY = std::min(c.minY, y);
        c.maxY = std::max(c.maxY, y);
        if (y == 0) c.touchesBottom = true;
        if (x == 0 || x == sizeX - 1 || z == 0 || z == sizeZ - 1) c.touchesSide = true;
        if (x > 0) tryAdd(index - offX);
        if (x < sizeX - 1) tryAdd(index + offX);
        if (y > floodMinY) tryAdd(index - offY);
        else if (y > 0 && stone[index - offY] != 0) c.touchesBottom = true;
        if (y < sizeY - 1) tryAdd(index + offY);
        if (z > 0) tryAdd(index - offZ);
        if (z < sizeZ - 1) tryAdd(index + offZ);
    }
    if (c.columns > 0) {
        c.centerX = static_cast<int>(javaRound(static_cast<double>(c.sumWorldX) / c.columns));
        c.centerZ = static_cast<int>(javaRound(static_cast<double>(c.sumWorldZ) / c.columns));
    }
    return c;
}

// isMatch, :3972-4003
bool isMatch(const Component& c, const bsfo_island_params& p) {
    if (c.blocks < p.min_blocks) return false;
    if (c.columns < p.min_columns) return false;
    if (c.maxY < p.min_y_for_match) return false;
    if (c.touchesBottom) return false;
    int width = c.maxWorldX - c.minWorldX + 1;
    int depth = c.maxWorldZ - c.minWorldZ + 1;
    if (width < p.min_width) return false;
    if (depth < p.min_depth) return false;
    double avg = c.blocks / static_cast<double>(std::max(1, c.columns));
    return avg >= p.min_avg_thickness;
}

// compareComponents, :4079-4094
int compareComponents(const Component& a, const Component& b) {
    if (a.blocks != b.blocks) return a.blocks < b.blocks ? -1 : 1;
    if (a.columns != b.columns) return a.columns < b.columns ? -1 : 1;
    int areaA = (a.maxWorldX - a.minWorldX + 1) * (a.maxWorldZ - a.minWorldZ + 1);
    int areaB = (b.maxWorldX - b.minWorldX + 1) * (b.maxWorldZ - b.minWorldZ + 1);
    return areaA < areaB ? -1 : (areaA == areaB ? 0 : 1);
}

inline bool isSampleCandidateColumn(int x, int z) { return (x == 4 || x == 12) && (z == 4 || z == 12); }  // :4005-4007

int floorDiv16(int v) { return v >= 0 ? v / 16 : -((-v + 15) / 16); }

int islandFromStone(int64_t seed, const uint8_t* stone, int sizeX, int sizeZ, int minWX, int minWZ,
                    const int* cand, int nCand, const bsfo_island_params& p, bsfo_island_hit* hit) {
    const int sizeY = 128;
    std::vector<uint8_t> visited(static_cast<size_t>(sizeX) * sizeY * sizeZ, 0);
    std::vector<int> queue(static_cast<size_t>(sizeX) * sizeY * sizeZ);
    std::vector<int> columnSeen(static_cast<size_t>(sizeX) * sizeZ, 0);
    int mark = 1;
    bool have = false;
    Component best;
    const int yzArea = sizeY * sizeZ;
    for (int i = 0; i < nCand; ++i) {
        int index = cand[i];
        if (stone[index] == 0 || visited[index] != 0) continue;
        int x = index / yzArea;
        int rem = index - x * yzArea;
        int y = rem / sizeZ;
        int z = rem - y * sizeZ;
        Component c = floodFill(stone, visited.data(), queue, columnSeen.data(), mark++, x, y, z, sizeX, sizeY,
                                sizeZ, minWX, minWZ, kIslandFloorY);
        if (!isMatch(c, p)) continue;
        if (!have || compareComponents(c, best) > 0) {
            best = c;
            have = true;
        }
    }
    if (!have) return 0;
    if (hit) {
        std::memset(hit, 0, sizeof(*hit));
        hit->seed = seed;
        hit->chunk_x = floorDiv16(best.centerX);
        hit->chunk_z = floorDiv16(best.centerZ);
        hit->min_x = best.minWorldX;
        hit->max_x = best.maxWorldX;
        hit->min_z = best.minWorldZ;
        hit->max_z = best.maxWorldZ;
        hit->min_y = best.minY;
        hit->max_y = best.maxY;
        hit->blocks = best.blocks;
        hit->columns = best.columns;
        hit->center_x = best.centerX;
        hit->center_z = best.centerZ;
        hit->flags = best.touchesSide ? 1 : 0;
    }
    return 1;
}

// findFloatingIslandInSeed exact block scan (:2861-2977) over an arbitrary chunk window.
int islandScan(int64_t seed, const bsfo_island_params& p, bsfo_island_hit* hit) {
    Terrain t(seed);
    const int sizeX = p.chunks_x * 16, sizeZ = p.chunks_z * 16, sizeY = 128;
    const int minWX = p.chunk_x0 * 16, minWZ = p.chunk_z0 * 16;
    std::vector<uint8_t> stone(static_cast<size_t>(sizeX) * sizeY * sizeZ, 0);
    std::vector<int> cand;
    const int copyStartY = kIslandFloorY - 1;
    std::vector<double> lattice(425);
    std::vector<uint8_t> blocks(32768);
    for (int cx = p.chunk_x0; cx < p.chunk_x0 + p.chunks_x; ++cx) {
        for (int cz = p.chunk_z0; cz < p.chunk_z0 + p.chunks_z; ++cz) {
            t.chunkDensity(cx, cz, lattice.data());
            upsample(lattice.data(), [&](int idx, double d) { blocks[idx] = d > 0.0 ? 1 : 0; });
            int baseX = (cx - p.chunk_x0) * 16, baseZ = (cz - p.chunk_z0) * 16;
            for (int x = 0; x < 16; ++x) {
                for (int z = 0; z < 16; ++z) {
                    bool allow = isSampleCandidateColumn(x, z);
                    for (int y = copyStartY; y < sizeY; ++y) {
                        if (blocks[(x * 16 + z) * 128 + y]) {
                            int flat = ((baseX + x) * sizeY + y) * sizeZ + (baseZ + z);
                            stone[flat] = 1;
                            if (allow && y >= kIslandFloorY && !blocks[(x * 16 + z) * 128 + y - 1]) cand.push_back(flat);
                        }
                    }
                }
            }
        }
    }
    return islandFromStone(seed, stone.data(), sizeX, sizeZ, minWX, minWZ, cand.data(),
                           static_cast<int>(cand.size()), p, hit);
}

int islandHasCandidate(int64_t seed, const bsfo_island_params& p) {
    Terrain t(seed);
    std::vector<double> lattice(425);
    std::vector<uint8_t> blocks(32768);
    for (int cx = p.chunk_x0; cx < p.chunk_x0 + p.chunks_x; ++cx) {
        for (int cz = p.chunk_z0; cz < p.chunk_z0 + p.chunks_z; ++cz) {
            t.chunkDensity(cx, cz, lattice.data());
            upsample(lattice.data(), [&](int idx, double d) { blocks[idx] = d > 0.0 ? 1 : 0; });
            for (int x : {4, 12})
                for (int z : {4, 12})
                    for (int y = kIslandFloorY; y < 128; ++y)
                        if (blocks[(x * 16 + z) * 128 + y] && !blocks[(x * 16 + z) * 128 + y - 1]) return 1;
        }
    }
    return 0;
}

// The ground gate's predicate: 1 iff some chunk of the window has a candidate start (sample
// column, y >= 68 over non-stone; :2909-2925, :4005-4007) in its stone that is not grounded within
// the chunk -- not 6-connected, above y = 67 and inside the chunk, to a block at y = 68 over stone
// at y = 67 (floodFillFlat's touchesBottom, IslandSearchEngine.java:4170-4175) -- and the stone
// connected to such starts inside the chunk either touches a face the chunk shares with another
// chunk of the window or holds min_blocks blocks (else each start's component lies inside the
// chunk and is too small for isMatch, :3972-4003).  Breadth-first from the contacts, then from the
// floating starts, per chunk.
int islandGroundGate(int64_t seed, const bsfo_island_params& p) {
    Terrain t(seed);
    std::vector<double> lattice(425);
    std::vector<uint8_t> blocks(32768), grounded(32768), reached(32768);
    std::vector<int> queue;
    auto at = [](int x, int y, int z) { return (x * 16 + z) * 128 + y; };
    for (int cx = p.chunk_x0; cx < p.chunk_x0 + p.chunks_x; ++cx) {
        for (int cz = p.chunk_z0; cz < p.chunk_z0 + p.chunks_z; ++cz) {
            t.chunkDensity(cx, cz, lattice.data());
            upsample(lattice.data(), [&](int idx, double d) { blocks[idx] = d > 0.0 ? 1 : 0; });
            std::fill(grounded.begin(), grounded.end(), 0);
            queue.clear();
            for (int x = 0; x < 16; ++x)
                for (int z = 0; z < 16; ++z)
                    if (blocks[at(x, kIslandFloorY, z)] && blocks[at(x, kIslandFloorY - 1, z)]) {
                        grounded[at(x, kIslandFloorY, z)] = 1;
                        queue.push_back(at(x, kIslandFloorY, z));
                    }
            for (size_t head = 0; head < queue.size(); ++head) {
                const int i = queue[head];
                const int x = i / (16 * 128), z = (i / 128) % 16, y = i % 128;
                const int nb[6][3] = {{x - 1, y, z}, {x + 1, y, z}, {x, y - 1, z},
                                      {x, y + 1, z}, {x, y, z - 1}, {x, y, z + 1}};
                for (const auto& n : nb) {
                    if (n[0] < 0 || n[0] > 15 || n[2] < 0 || n[2] > 15) continue;
                    if (n[1] < kIslandFloorY || n[1] > 127) continue;
                    const int j = at(n[0], n[1], n[2]);
                    if (blocks[j] && !grounded[j]) {
                        grounded[j] = 1;
                        queue.push_back(j);
                    }
                }
            }
            // the floating starts, and the stone connected to them that is not grounded
            std::fill(reached.begin(), reached.end(), 0);
            queue.clear();
            for (int sx : {4, 12})
                for (int sz : {4, 12})
                    for (int sy = kIslandFloorY; sy < 128; ++sy) {
                        const int i = at(sx, sy, sz);
                        if (blocks[i] && !blocks[at(sx, sy - 1, sz)] && !grounded[i] && !reached[i]) {
                            reached[i] = 1;
                            queue.push_back(i);
                        }
                    }
            if (queue.empty()) continue;
            const bool lo_x = cx > p.chunk_x0, hi_x = cx < p.chunk_x0 + p.chunks_x - 1;
            const bool lo_z = cz > p.chunk_z0, hi_z = cz < p.chunk_z0 + p.chunks_z - 1;
            bool face = false;
            for (size_t head = 0; head < queue.size(); ++head) {
                const int i = queue[head];
                const int x = i / (16 * 128), z = (i / 128) % 16, y = i % 128;
                face = face || (x == 0 && lo_x) || (x == 15 && hi_x) || (z == 0 && lo_z) || (z == 15 && hi_z);
                const int nb[6][3] = {{x - 1, y, z}, {x + 1, y, z}, {x, y - 1, z},
                                      {x, y + 1, z}, {x, y, z - 1}, {x, y, z + 1}};
                for (const auto& n : nb) {
                    if (n[0] < 0 || n[0] > 15 || n[2] < 0 || n[2] > 15) continue;
                    if (n[1] < kIslandFloorY || n[1] > 127) continue;
                    const int j = at(n[0], n[1], n[2]);
                    if (blocks[j] && !grounded[j] && !reached[j]) {
                        reached[j] = 1;
                        queue.push_back(j);
                    }
                }
            }
            if (face || static_cast<int>(queue.size()) >= p.min_blocks) return 1;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Stage-0 scout statistic (config 2): noise2 octave 0 only, chunk (0,0) 5x17x5.
// ---------------------------------------------------------------------------
double scoutStatistic(int64_t seed) {
    JRandom r(seed);
    Perlin p;
    p.reseed(r);
    double out[425];
    for (double& v : out) v = 0.0;
    p.add(out, 0.0, 0.0, 0.0, 5, 17, 5, 684.412, 684.412, 684.412, 1.0);
    double s = 0.0;
    for (int i = 0; i < 425; ++i) s += out[i];
    return s;
}

// ---------------------------------------------------------------------------
// Stage0 lattice (coarse semantics): 16x16 columns at coarse -28+4i (+offset), 17 Y,
// climate at coarse*4+2.  BetaTerrain173.java:1534-1580, 2524-2552 with the cached
// strided noise paths (NoiseGeneratorPerlin173.java:401-507), which evaluate the same
// per-point arithmetic as the legacy bulk method at coordinates (from + i*step).
// ---------------------------------------------------------------------------
void stage0Lattice(const Terrain& t, int offX, int offZ, double* out /* 256*17 */) {
    const int n = 16;
    std::vector<double> temp(n * n), rain(n * n);
    const int fromX = -28 + offX, fromZ = -28 + offZ;
    t.climate.sample(fromX * 4 + 2, fromZ * 4 + 2, n, n, 16, 16, temp.data(), rain.data());
    // Evaluate each column with the legacy bulk method on a 1x17x1 box at the exact coordinate.
    std::vector<double> one(17), c1(17), c2(17), c3(17);
    for (int ix = 0; ix < n; ++ix) {
        for (int iz = 0; iz < n; ++iz) {
            double cx = static_cast<double>(fromX + 4 * ix);
            double cz = static_cast<double>(fromZ + 4 * iz);
            double col4 = 0, col5 = 0;
            auto oct3 = [&](const Octaves& o, double s1, double s2, double* dst) {
                for (int y = 0; y < 17; ++y) dst[y] = 0.0;
                double amp = 1.0;
                for (size_t k = 0; k < o.g.size(); ++k) {
                    o.g[k].add(dst, cx, 0.0, cz, 1, 17, 1, s1 * amp, s2 * amp, s1 * amp, amp);
                    amp /= 2.0;
                }
            };
            auto oct2 = [&](const Octaves& o, double s) {
                double v = 0.0;
                double amp = 1.0;
                for (size_t k = 0; k < o.g.size(); ++k) {
                    o.g[k].add(&v, cx, 10.0, cz, 1, 1, 1, s * amp, amp, s * amp, amp);
                    amp /= 2.0;
                }
                return v;
            };
            col4 = oct2(t.n4, 1.121);
            col5 = oct2(t.n5, 200.0);
            oct3(t.n1, 684.412 / 80.0, 684.412 / 160.0, c1.data());
            oct3(t.n2, 684.412, 684.412, c2.data());
            oct3(t.n3, 684.412, 684.412, c3.data());
            // combine exactly as generateTerrainNoise for one column
            double dd2 = temp[ix * n + iz];
            double dd3 = rain[ix * n + iz] * dd2;
            double d4 = 1.0 - dd3;
            d4 *= d4;
            d4 *= d4;
            d4 = 1.0 - d4;
            double d5 = (col4 + 256.0) / 512.0;
            d5 *= d4;
            if (d5 > 1.0) d5 = 1.0;
            double d6 = col5 / 8000.0;
            if (d6 < 0.0) d6 = -d6 * 0.3;
            d6 = d6 * 3.0 - 2.0;
            if (d6 < 0.0) {
                d6 /= 2.0;
                if (d6 < -1.0) d6 = -1.0;
                d6 /= 1.4;
                d6 /= 2.0;
                d5 = 0.0;
            } else {
                if (d6 > 1.0) d6 = 1.0;
                d6 /= 8.0;
            }
            if (d5 < 0.0) d5 = 0.0;
            d5 += 0.5;
            d6 = d6 * 17.0 / 16.0;
            double d7 = 17.0 / 2.0 + d6 * 4.0;
            for (int y = 0; y < 17; ++y) {
                double d9 = (static_cast<double>(y) - d7) * 12.0 / d5;
                if (d9 < 0.0) d9 *= 4.0;
                double d10 = c2[y] / 512.0, d11 = c3[y] / 512.0, d12 = (c1[y] / 10.0 + 1.0) / 2.0, d8;
                if (d12 < 0.0) d8 = d10;
                else if (d12 > 1.0) d8 = d11;
                else d8 = d10 + (d11 - d10) * d12;
                d8 -= d9;
                if (y > 13) {
                    double d13 = static_cast<double>(static_cast<float>(y - 13) / 3.0F);
                    d8 = d8 * (1.0 - d13) + -10.0 * d13;
                }
                out[(ix * n + iz) * 17 + y] = d8;
            }
        }
    }
}

bool isP20Axis(int i) { return i == 0 || i == 2 || i == 5 || i == 7 || i == 8 || i == 10 || i == 13 || i == 15; }

// highestSparseReentryYIndex, BetaTerrain173.java:2486-2504 (positive -> gap -> positive), on
// a column's 17-level positive mask (bit y = density > 0 at level y).
int highestReentryMask(uint32_t mask) {
    bool sawPositive = false, sawGap = false;
    int highest = -1;
    for (int y = 0; y < 17; ++y) {
        bool pos = (mask >> y) & 1u;
        if (pos) {
            if (sawPositive && sawGap) highest = y;
            sawPositive = true;
        } else if (sawPositive) {
            sawGap = true;
        }
    }
    return highest;
}

int highestReentry(const double* col) {
    uint32_t mask = 0;
    for (int y = 0; y < 17; ++y)
        if (col[y] > 0.0) mask |= 1u << y;
    return highestReentryMask(mask);
}

// ---------------------------------------------------------------------------
// P19 monster gate features (gpu_p20_benchmark.cpp:1022-1114,
// BetaTerrain173.java:1926-2064)
// ---------------------------------------------------------------------------
void shapeStats(const uint8_t* grid, int size, int* largest, int* width, int* depth, int* touches) {
    std::vector<uint8_t> visited(size * size, 0);
    std::vector<int> queue(size * size);
    int big = 0, minX = size, maxX = -1, minZ = size, maxZ = -1;
    bool border = false;
    for (int x = 0; x < size; ++x) {
        for (int z = 0; z < size; ++z) {
            int idx = x * size + z;
            if (!grid[idx]) continue;
            minX = std::min(minX, x);
            maxX = std::max(maxX, x);
            minZ = std::min(minZ, z);
            maxZ = std::max(maxZ, z);
            if (x == 0 || z == 0 || x == size - 1 || z == size - 1) border = true;
            if (visited[idx]) continue;
            int head = 0, tail = 0, cluster = 0;
            visited[idx] = 1;
            queue[tail++] = idx;
            while (head < tail) {
                int cur = queue[head++];
                ++cluster;
                int cx = cur / size, cz = cur % size;
                const int nx[4] = {cx - 1, cx + 1, cx, cx};
                const int nz[4] = {cz, cz, cz - 1, cz + 1};
                for (int k = 0; k < 4; ++k) {
                    if (nx[k] < 0 || nz[k] < 0 || nx[k] >= size || nz[k] >= size) continue;
                    int nb = nx[k] * size + nz[k];
                    if (!grid[nb] || visited[nb]) continue;
                    visited[nb] = 1;
                    queue[tail++] = nb;
                }
            }
            big = std::max(big, cluster);
        }
    }
    *largest = big;
    *width = big == 0 ? 0 : maxX - minX + 1;
    *depth = big == 0 ? 0 : maxZ - minZ + 1;
    *touches = border ? 1 : 0;
}

}  // namespace

// P19 tree ensemble: generated table (oracle/p19_trees.inc), see that file's header.
#include "p19_trees.inc"

namespace {

double p19Feature(int kind, int scout, const bsfo_p19_features& p) {
    const double EPS = 1.0E-6;
    switch (kind) {
        case 0: return static_cast<double>(p.full_y88) / (scout + EPS);
        case 1: return p.full_y88;
        case 2: return p.full_y96;
        case 3: return p.y88_width + p.y88_depth + p.y96_width + p.y96_depth;
        case 4: return p.y96_width * p.y96_depth;
        case 5: return scout;
        case 6: return p.y88_width * p.y88_depth;
        case 7: return static_cast<double>(p.y88_largest) / (p.full_y88 + EPS);
        case 8: return p.y96_width;
        case 9: return p.full_y96 + p.full_y104 + p.full_y112;
        case 10: return p.y88_depth;
        case 11: return p.y88_width;
        case 12: return p.y96_depth;
        case 13: return static_cast<double>(p.y96_largest) / (p.full_y96 + EPS);
        case 14: return static_cast<double>(p.full_y104) / (p.full_y96 + EPS);
        case 15: return static_cast<double>(p.full_y96) / (p.full_y88 + EPS);
        case 16: return p.y96_largest;
        case 17: return p.full_y112;
        case 18: return p.y88_largest + p.y96_largest;
        case 19: return p.full_y104;
        case 20: return p.y88_largest;
        default: return 0.0;
    }
}

double p19Tree(int tree, int scout, const bsfo_p19_features& p) {
    int node = kP19TreeRoot[tree];
    for (;;) {
        const P19Node& n = kP19Nodes[node];
        if (n.feature < 0) return n.value;
        node = p19Feature(n.feature, scout, p) <= n.value ? n.left : n.right;
    }
}

// coarse 61x61x17 grid, coarse_exact_core.hpp:11-16, 83-169, 218-386
constexpr int kCoarseSize = 61, kCoarseY = 17;

int coarseScoreSigns(const uint8_t* signs) {
    const int SIZE = kCoarseSize, YL = kCoarseY, COLUMNS = SIZE * SIZE, CELLS = COLUMNS * YL;
    std::vector<int> labels(CELLS, 0), queue(CELLS), columnSeen(COLUMNS, 0), columnMinY(COLUMNS), compCols(COLUMNS);
    auto index3 = [&](int x, int y, int z) { return (x * SIZE + z) * YL + y; };
    int best = 0, nextId = 1;
    for (int x = 0; x < SIZE; ++x)
        for (int z = 0; z < SIZE; ++z)
            for (int y = 0; y < YL; ++y) {
                int start = index3(x, y, z);
                if (labels[start] != 0 || signs[start] == 0) continue;
                int id = nextId++;
                int head = 0, tail = 0;
                queue[tail++] = start;
                labels[start] = id;
                int cells = 0, maxY = 0, colCount = 0;
                bool bottom = false, side = false;
                while (head < tail) {
                    int idx = queue[head++];
                    int cy = idx % YL, tmp = idx / YL, cz = tmp % SIZE, cx = tmp / SIZE;
                    ++cells;
                    if (cy > maxY) maxY = cy;
                    if (cy == 0) bottom = true;
                    if (cx == 0 || cz == 0 || cx == SIZE - 1 || cz == SIZE - 1) side = true;
                    int col = cx * SIZE + cz;
                    if (columnSeen[col] != id) {
                        columnSeen[col] = id;
                        compCols[colCount++] = col;
                        columnMinY[col] = cy;
                    } else if (cy < columnMinY[col]) {
                        columnMinY[col] = cy;
                    }
                    auto enq = [&](int ni) {
                        if (labels[ni] == 0 && signs[ni] != 0) {
                            labels[ni] = id;
                            queue[tail++] = ni;
                        }
                    };
                    if (cx + 1 < SIZE) enq(index3(cx + 1, cy, cz));
                    if (cx > 0) enq(index3(cx - 1, cy, cz));
                    if (cy + 1 < YL) enq(index3(cx, cy + 1, cz));
                    if (cy > 0) enq(index3(cx, cy - 1, cz));
                    if (cz + 1 < SIZE) enq(index3(cx, cy, cz + 1));
                    if (cz > 0) enq(index3(cx, cy, cz - 1));
                }
                if (cells <= best || maxY < 8 || bottom || side) continue;
                bool reentry = false;
                for (int i = 0; i < colCount && !reentry; ++i) {
                    int col = compCols[i];
                    int cx = col / SIZE, cz = col - cx * SIZE;
                    int minY = columnMinY[col];
                    if (minY <= 0 || minY >= YL) continue;
                    for (int ly = minY - 1; ly >= 0; --ly) {
                        int idx = index3(cx, ly, cz);
                        if (signs[idx] != 0 && labels[idx] != id) {
                            if (minY - ly - 1 >= 1) reentry = true;
                            break;
                        }
                    }
                }
                if (reentry) best = cells;
            }
    return best;
}

}  // namespace

template <class F>
void parallelFor(int64_t count, int threads, F&& f) {
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    int64_t per = (count + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        int64_t lo = t * per, hi = std::min<int64_t>(count, lo + per);
        if (lo >= hi) break;
        pool.emplace_back([=, &f]() { f(lo, hi); });
    }
    for (auto& th : pool) th.join();
}

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

void bsfo_random_doubles(int64_t seed, int n, double* out) {
    JRandom r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.nextDouble();
}
void bsfo_random_ints(int64_t seed, int n, int32_t bound, int32_t* out) {
    JRandom r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.nextInt(bound);
}
void bsfo_random_next_bits(int64_t seed, int n, int32_t bits, int32_t* out) {
    JRandom r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next(bits);
}

void bsfo_random_mixed(int64_t seed, int n, const int32_t* ops, int64_t* out) {
    JRandom r(seed);
    for (int i = 0; i < n; ++i) {
        if (ops[i] == -1) {
            double d = r.nextDouble();
            std::memcpy(&out[i], &d, 8);
        } else if (ops[i] == -2) {
            out[i] = r.next(31);
        } else {
            out[i] = r.nextInt(ops[i]);
        }
    }
}

void bsfo_perlin_state(int64_t seed, int generator_index, uint8_t perm256[256], double abc[3]) {
    JRandom r(seed);
    Perlin p;
    for (int i = 0; i <= generator_index; ++i) p.reseed(r);
    for (int i = 0; i < 256; ++i) perm256[i] = static_cast<uint8_t>(p.d[i]);
    abc[0] = p.a; abc[1] = p.b; abc[2] = p.c;
}

void bsfo_simplex_state(int64_t seed, int kind, int octave, uint8_t perm256[256], double abc[3]) {
    const uint64_t mul[3] = {9871ULL, 39811ULL, 543321ULL};
    JRandom r(static_cast<int64_t>(static_cast<uint64_t>(seed) * mul[kind]));
    Simplex s;
    for (int i = 0; i <= octave; ++i) s.reseed(r);
    for (int i = 0; i < 256; ++i) perm256[i] = static_cast<uint8_t>(s.e[i]);
    abc[0] = s.a; abc[1] = s.b; abc[2] = s.c;
}

static Perlin perlinFrom(const uint8_t* perm, const double* abc) {
    Perlin p;
    for (int i = 0; i < 256; ++i) p.d[i] = p.d[i + 256] = perm[i];
    p.a = abc[0]; p.b = abc[1]; p.c = abc[2];
    return p;
}

double bsfo_perlin3_point(const uint8_t perm256[256], const double abc[3], double x, double y, double z) {
    return perlinFrom(perm256, abc).point(x, y, z);
}

void bsfo_perlin3_column(const uint8_t perm256[256], const double abc[3], double x, double sy, double z,
                         int y_len, double* out) {
    Perlin p = perlinFrom(perm256, abc);
    for (int i = 0; i < y_len; ++i) out[i] = 0.0;
    // xCoord/zCoord are already scaled: (from + 0) * 1.0 keeps them exact.
    p.add(out, x, 0.0, z, 1, y_len, 1, 1.0, sy, 1.0, 1.0);
}

double bsfo_perlin2_point(const uint8_t perm256[256], const double abc[3], double x, double z) {
    Perlin p = perlinFrom(perm256, abc);
    double v = 0.0;
    p.add(&v, x, 10.0, z, 1, 1, 1, 1.0, 1.0, 1.0, 1.0);
    return v;
}

double bsfo_simplex2_point(const uint8_t perm256[256], const double abc[3], double x, double z) {
    Simplex s;
    for (int i = 0; i < 256; ++i) s.e[i] = s.e[i + 256] = perm256[i];
    s.a = abc[0]; s.b = abc[1]; s.c = abc[2];
    return s.sample(x + s.a, z + s.b);
}

void bsfo_chunk_density(int64_t seed, int32_t cx, int32_t cz, double* out425) {
    Terrain t(seed);
    t.chunkDensity(cx, cz, out425);
}

void bsfo_chunk_climate(int64_t seed, int32_t cx, int32_t cz, double* temp256, double* rain256) {
    Climate c(seed);
    c.sample(cx * 16, cz * 16, 16, 16, 1, 1, temp256, rain256);
}

void bsfo_chunk_noise(int64_t seed, int32_t cx, int32_t cz, double* n1, double* n2, double* n3, double* n4,
                      double* n5) {
    Terrain t(seed);
    double temp[256], rain[256], out[425];
    t.chunkClimate(cx, cz, temp, rain);
    t.terrainNoise(out, cx * 4, 0, cz * 4, 5, 17, 5, temp, rain, false, n1, n2, n3, n4, n5);
}

void bsfo_chunk_stone(int64_t seed, int32_t cx, int32_t cz, uint8_t* out) {
    double lattice[425];
    bsfo_chunk_density(seed, cx, cz, lattice);
    upsample(lattice, [&](int idx, double d) { out[idx] = d > 0.0 ? 1 : 0; });
}

void bsfo_upsample_stone(const double* lattice425, uint8_t* out) {
    upsample(lattice425, [&](int idx, double d) { out[idx] = d > 0.0 ? 1 : 0; });
}

void bsfo_upsample_density(const double* lattice425, double* out) {
    upsample(lattice425, [&](int idx, double d) { out[idx] = d; });
}

double bsfo_scout_statistic(int64_t seed) { return scoutStatistic(seed); }

// Java: z = sequenceSeed + attempt * 0x9E3779B97F4A7C15L, two xor-shift-multiply rounds
// (>>> 30, >>> 27) and a final xor-shift (>>> 31), all in wrapping 64-bit arithmetic.
uint64_t bsfo_splitmix_seed(uint64_t sequence_seed, uint64_t attempt) {
    uint64_t z = sequence_seed + attempt * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

int bsfo_island_scan(int64_t seed, const bsfo_island_params* p, bsfo_island_hit* hit) {
    return islandScan(seed, *p, hit);
}

int bsfo_island_from_stone(int64_t seed, const uint8_t* stone, int32_t size_x, int32_t size_z, int32_t min_wx,
                           int32_t min_wz, const int32_t* cand, int32_t n_cand, const bsfo_island_params* p,
                           bsfo_island_hit* hit) {
    return islandFromStone(seed, stone, size_x, size_z, min_wx, min_wz, cand, n_cand, *p, hit);
}

int bsfo_island_has_candidate(int64_t seed, const bsfo_island_params* p) { return islandHasCandidate(seed, *p); }

int bsfo_island_ground_gate(int64_t seed, const bsfo_island_params* p) { return islandGroundGate(seed, *p); }

uint64_t bsfo_density_digest(const double* d) {
    uint64_t h = 1469598103934665603ULL;
    for (int i = 0; i < 425; ++i) {
        uint64_t bits;
        std::memcpy(&bits, &d[i], 8);
        for (int b = 0; b < 8; ++b) {
            h ^= (bits >> (8 * b)) & 0xFFu;
            h *= 1099511628211ULL;
        }
    }
    return h;
}

uint64_t bsfo_run_density(int64_t first, int64_t count, int threads, uint64_t* digests) {
    std::atomic<uint64_t> combined{0};
    parallelFor(count, threads, [&](int64_t lo, int64_t hi) {
        double d[425];
        uint64_t acc = 0;
        for (int64_t i = lo; i < hi; ++i) {
            bsfo_chunk_density(first + i, 0, 0, d);
            uint64_t h = bsfo_density_digest(d);
            if (digests) digests[i] = h;
            acc ^= h * (0x9E3779B97F4A7C15ULL ^ static_cast<uint64_t>(i));
        }
        combined.fetch_xor(acc);
    });
    return combined.load();
}

int64_t bsfo_run_scout(int64_t first, int64_t count, double threshold, int threads, int64_t* out, int64_t cap) {
    std::vector<std::vector<int64_t>> parts(std::max(1, threads));
    std::atomic<int> slot{0};
    parallelFor(count, threads, [&](int64_t lo, int64_t hi) {
        int me = slot.fetch_add(1);
        auto& v = parts[me];
        for (int64_t i = lo; i < hi; ++i)
            if (scoutStatistic(first + i) > threshold) v.push_back(first + i);
    });
    std::vector<int64_t> all;
    for (auto& v : parts) all.insert(all.end(), v.begin(), v.end());
    std::sort(all.begin(), all.end());
    int64_t n = static_cast<int64_t>(all.size());
    for (int64_t i = 0; i < std::min(n, cap); ++i) out[i] = all[i];
    return n;
}

int64_t bsfo_run_island(int64_t first, int64_t count, const bsfo_island_params* p, int threads,
                        bsfo_island_hit* out, int64_t cap) {
    std::vector<std::vector<bsfo_island_hit>> parts(std::max(1, threads));
    std::atomic<int> slot{0};
    parallelFor(count, threads, [&](int64_t lo, int64_t hi) {
        int me = slot.fetch_add(1);
        auto& v = parts[me];
        for (int64_t i = lo; i < hi; ++i) {
            bsfo_island_hit h;
            if (islandScan(first + i, *p, &h)) v.push_back(h);
        }
    });
    std::vector<bsfo_island_hit> all;
    for (auto& v : parts) all.insert(all.end(), v.begin(), v.end());
    std::sort(all.begin(), all.end(), [](const bsfo_island_hit& a, const bsfo_island_hit& b) { return a.seed < b.seed; });
    int64_t n = static_cast<int64_t>(all.size());
    for (int64_t i = 0; i < std::min(n, cap); ++i) out[i] = all[i];
    return n;
}

void bsfo_stage0_lattice(int64_t seed, int32_t ox, int32_t oz, double* out) {
    Terrain t(seed);
    stage0Lattice(t, ox, oz, out);
}

void bsfo_stage0(int64_t seed, int32_t ox, int32_t oz, int32_t p20_min, int32_t upper_min, bsfo_stage0_result* r) {
    std::vector<double> lat(256 * 17);
    bsfo_stage0_lattice(seed, ox, oz, lat.data());
    std::memset(r, 0, sizeof(*r));
    std::memset(r->highest, 0xFF, 256);
    int p20 = 0, upper = 0;
    for (int c = 0; c < 256; ++c) {
        uint8_t m = 0;
        for (int k = 0; k < 6; ++k)
            if (lat[c * 17 + 11 + k] > 0.0) m |= static_cast<uint8_t>(1u << k);
        r->upper_mask[c] = m;
        if (m) {
            ++upper;
            if (isP20Axis(c >> 4) && isP20Axis(c & 15)) ++p20;
        }
    }
    r->p20 = p20;
    // applyP20MinimumGateKernel + upper/lower kernels: a failing P20 gate zeroes the later stages
    const int gated = p20 >= p20_min ? p20 : 0;
    if (gated <= 0) return;
    r->upper = upper;
    if (upper < upper_min) return;
    int high = 0;
    for (int c = 0; c < 256; ++c) {
        if (!r->upper_mask[c]) continue;
        int h = highestReentry(&lat[c * 17]);
        r->highest[c] = h < 0 ? 0xFF : static_cast<uint8_t>(h);
        if (h >= 11) ++high;
    }
    r->high = high;
}

void bsfo_p19_features_from_highest(const uint8_t highest[256], bsfo_p19_features* f) {
    std::memset(f, 0, sizeof(*f));
    uint8_t y88[256] = {0}, y96[256] = {0};
    for (int c = 0; c < 256; ++c) {
        if (highest[c] == 0xFF) continue;
        int h = highest[c];
        if (h >= 11) { ++f->full_y88; y88[c] = 1; }
        if (h >= 12) { ++f->full_y96; y96[c] = 1; }
        if (h >= 13) ++f->full_y104;
        if (h >= 14) ++f->full_y112;
    }
    shapeStats(y88, 16, &f->y88_largest, &f->y88_width, &f->y88_depth, &f->y88_touches);
    shapeStats(y96, 16, &f->y96_largest, &f->y96_width, &f->y96_depth, &f->y96_touches);
}

double bsfo_p19_score(int32_t scout, const bsfo_p19_features* f) {
    double sum = 0.0;
    for (int t = 0; t < 16; ++t) sum = sum + p19Tree(t, scout, *f);
    return sum / 16;
}

int bsfo_p19_extreme(const bsfo_p19_features* f) { return f->y88_largest >= 14 || f->y96_largest >= 14; }

int bsfo_mega_topology_reject(double score, int extreme, const bsfo_p19_features* f) {
    if (extreme) return 0;
    if (!(score < 5.0)) return 0;
    return f->full_y112 == 0 || f->y96_largest < 4;
}

int bsfo_coarse_score_signs(const uint8_t* signs) { return coarseScoreSigns(signs); }

// Exact scan of the window of chunk radius r around a centre chunk (findFloatingIslandInSeed
// with hunterMode off, :2861-2977).
static int regionScan(int64_t seed, int cx, int cz, int r, const bsfo_island_params& s, bsfo_island_hit* hit) {
    bsfo_island_params p = s;
    p.chunk_x0 = cx - r;
    p.chunk_z0 = cz - r;
    p.chunks_x = 2 * r + 1;
    p.chunks_z = 2 * r + 1;
    return islandScan(seed, p, hit);
}

// compareComponents(a, b) > 0 on two hits (:4079-4094): blocks, then columns, then the
// bounding-box area.
static bool hitBetter(const bsfo_island_hit& a, const bsfo_island_hit& b) {
    if (a.blocks != b.blocks) return a.blocks > b.blocks;
    if (a.columns != b.columns) return a.columns > b.columns;
    const int areaA = (a.max_x - a.min_x + 1) * (a.max_z - a.min_z + 1);
    const int areaB = (b.max_x - b.min_x + 1) * (b.max_z - b.min_z + 1);
    return areaA > areaB;
}

static const int kBorderVerifyRadii[5] = {10, 13, 16, 20, 24};  // :106-107

// verifySideBorderMatch (:3188-3259): widen through the checkpoint radii above the result's
// own until the component no longer touches the side; a wider scan without a match drops it.
static bool borderVerify(int64_t seed, int cx, int cz, const bsfo_island_params& s, bsfo_island_hit* best,
                         int* radius) {
    for (int r : kBorderVerifyRadii) {
        if (r <= *radius) continue;
        bsfo_island_hit e;
        if (!regionScan(seed, cx, cz, r, s, &e)) return false;
        *best = e;
        *radius = r;
        if (!(e.flags & 1)) return true;
    }
    return true;
}

void bsfo_region_attempt(int64_t seed, int32_t cx, int32_t cz, int32_t coarse, const bsfo_island_params* s,
                         int32_t radius, bsfo_region_result* out) {
    std::memset(out, 0, sizeof(*out));
    bsfo_island_hit normal;
    int normal_r = radius;
    bool have_normal = regionScan(seed, cx, cz, radius, *s, &normal) != 0;
    out->normal_found = have_normal ? 1 : 0;
    if (have_normal && (normal.flags & 1) && radius < 24) have_normal = borderVerify(seed, cx, cz, *s, &normal, &normal_r);
    if (have_normal) {
        out->normal_hit = normal;
        out->normal_kept = 1;
        out->normal_radius = normal_r;
    }
    bool have = have_normal;
    bsfo_island_hit match = normal;
    int match_r = normal_r;
    if (coarse >= 120 && radius < 16) {  // shouldRunHotSeedScan, :3131-3136
        out->hot_scan_ran = 1;
        bsfo_island_hit hot;
        int hot_r = 16;
        bool have_hot = regionScan(seed, cx, cz, 16, *s, &hot) != 0;
        if (have_hot && (hot.flags & 1) && hot_r < 24) have_hot = borderVerify(seed, cx, cz, *s, &hot, &hot_r);
        if (have_hot) {
            out->hot_hit = hot;
            out->hot_found = 1;
            out->hot_radius = hot_r;
        }
        if (have_hot && (!have || hitBetter(hot, match))) {
            match = hot;
            match_r = hot_r;
            have = true;
            out->hot_won = 1;
        }
    }
    out->found = have ? 1 : 0;
    if (have) {
        out->hit = match;
        out->radius_used = match_r;
    }
}

// HuntProfileConfig (gpu_p20_benchmark.cpp:7270-7290) and the production centre chunk offsets
// (:7316-7317).
struct Stage0Profile {
    int p20_min, upper_min, high_min;
    double p19_min;
    int coarse_min, mega;
};
static const Stage0Profile kStage0Profiles[4] = {
    {1, 5, 5, 4.02, 85, 0}, {1, 8, 6, 4.02, 85, 1}, {3, 19, 12, 6.70, 95, 0}, {3, 19, 20, 6.70, 700, 0}};
static const int kStage0CenterX[8] = {0, 15, 0, 15, -15, -15, 0, 15};
static const int kStage0CenterZ[8] = {0, 0, 15, 15, 0, 15, -15, -15};

int bsfo_highest_reentry_mask(uint32_t mask17) { return highestReentryMask(mask17 & 0x1FFFFu); }

void bsfo_stage0_profile(int32_t profile, int32_t* p20_min, int32_t* upper_min, int32_t* high_min, double* p19_min,
                         int32_t* coarse_min, int32_t* mega) {
    const Stage0Profile& pr = kStage0Profiles[profile & 3];
    *p20_min = pr.p20_min;
    *upper_min = pr.upper_min;
    *high_min = pr.high_min;
    *p19_min = pr.p19_min;
    *coarse_min = pr.coarse_min;
    *mega = pr.mega;
}

void bsfo_stage0_records(const int64_t* seeds, int64_t count, int32_t profile, int32_t research, int32_t centers,
                         int threads, uint8_t* out, uint64_t* counters) {
    const Stage0Profile& pr = kStage0Profiles[profile];
    const size_t rb = research ? 34 : 16;
    std::atomic<uint64_t> sums[11];
    for (auto& v : sums) v = 0;
    parallelFor(count * centers, threads, [&](int64_t lo, int64_t hi) {
        uint64_t c[11] = {0};
        for (int64_t item = lo; item < hi; ++item) {
            const int64_t seed = seeds[item / centers];
            const int ci = static_cast<int>(item % centers);
            const int ox = kStage0CenterX[ci] * 4, oz = kStage0CenterZ[ci] * 4;
            bsfo_stage0_result r;
            bsfo_stage0(seed, ox, oz, pr.p20_min, pr.upper_min, &r);
            bsfo_p19_features f;
            std::memset(&f, 0, sizeof(f));
            double score = std::numeric_limits<double>::quiet_NaN();
            int pass = 0, extreme = 0, mega = 0, coarse = 0;
            const bool evaluated = r.p20 >= pr.p20_min && r.upper >= pr.upper_min && r.high >= pr.high_min;
            if (evaluated) {
                bsfo_p19_features_from_highest(r.highest, &f);
                score = bsfo_p19_score(r.upper, &f);
                extreme = bsfo_p19_extreme(&f);
                pass = extreme || score >= pr.p19_min;
                mega = pass && pr.mega && bsfo_mega_topology_reject(score, extreme, &f);
                if (pass && !mega) coarse = bsfo_coarse_score(seed, ox, oz, nullptr);
            }
            uint8_t* o = out + static_cast<size_t>(item) * rb;
            auto u8 = [](int v) { return static_cast<uint8_t>(v < 0 ? 0 : (v > 255 ? 255 : v)); };
            auto u16 = [&](size_t at, int v) {
                const unsigned u = static_cast<unsigned>(v < 0 ? 0 : (v > 65535 ? 65535 : v));
                o[at] = static_cast<uint8_t>(u & 0xFF);
                o[at + 1] = static_cast<uint8_t>(u >> 8);
            };
            o[0] = u8(r.p20);
            o[1] = u8(r.upper);
            o[2] = u8(r.high);
            o[3] = static_cast<uint8_t>(pass);
            o[4] = static_cast<uint8_t>(coarse & 0xFF);
            o[5] = static_cast<uint8_t>((coarse >> 8) & 0xFF);
            o[6] = static_cast<uint8_t>(extreme);
            o[7] = static_cast<uint8_t>(mega);
            uint64_t bits;
            std::memcpy(&bits, &score, 8);
            for (int k = 0; k < 8; ++k) o[8 + k] = static_cast<uint8_t>(bits >> (8 * k));
            if (research) {
                u16(16, f.full_y88);
                u16(18, f.full_y96);
                u16(20, f.full_y104);
                u16(22, f.full_y112);
                u16(24, f.y88_largest);
                o[26] = static_cast<uint8_t>(f.y88_width);
                o[27] = static_cast<uint8_t>(f.y88_depth);
                o[28] = static_cast<uint8_t>(f.y88_touches ? 1 : 0);
                u16(29, f.y96_largest);
                o[31] = static_cast<uint8_t>(f.y96_width);
                o[32] = static_cast<uint8_t>(f.y96_depth);
                o[33] = static_cast<uint8_t>(f.y96_touches ? 1 : 0);
            }
            // closing-summary counters, gpu_p20_benchmark.cpp:7555-7571
            if (r.p20 >= pr.p20_min) ++c[2];
            if (r.p20 >= pr.p20_min && r.upper >= pr.upper_min) ++c[3];
            if (evaluated) {
                ++c[4];
                ++c[5];
                if (pass) {
                    ++c[6];
                    if (mega) {
                        ++c[8];
                    } else {
                        ++c[9];
                        if (coarse >= pr.coarse_min) ++c[10];
                    }
                } else {
                    ++c[7];
                }
            }
        }
        for (int k = 2; k < 11; ++k) sums[k] += c[k];
    });
    sums[0] += static_cast<uint64_t>(count);
    sums[1] += static_cast<uint64_t>(count * centers);
    for (int k = 0; k < 11; ++k) counters[k] += sums[k].load();
}

void bsfo_coarse_arrays(int64_t seed, int32_t ox, int32_t oz, double* n1, double* n2, double* n3, double* n4,
                        double* n5, double* temp, double* rain, uint8_t* signs) {
    Terrain t(seed);
    const int SIZE = kCoarseSize, YL = kCoarseY;
    const int fromX = -28 + ox, fromZ = -28 + oz;
    // coarse climate at block coarse*4+2 (BetaTerrain173.java:2562-2584, strided sampler)
    t.climate.sample(fromX * 4 + 2, fromZ * 4 + 2, SIZE, SIZE, 4, 4, temp, rain);
    std::vector<double> dens(SIZE * YL * SIZE);
    t.terrainNoise(dens.data(), fromX, 0, fromZ, SIZE, YL, SIZE, temp, rain, true, n1, n2, n3, n4, n5);
    for (size_t i = 0; i < dens.size(); ++i) signs[i] = dens[i] > 0.0 ? 1 : 0;
}

int bsfo_coarse_score(int64_t seed, int32_t ox, int32_t oz, uint8_t* signs_out) {
    const int SIZE = kCoarseSize, YL = kCoarseY, CELLS = SIZE * SIZE * YL, COLS = SIZE * SIZE;
    std::vector<double> n1(CELLS), n2(CELLS), n3(CELLS), n4(COLS), n5(COLS), temp(COLS), rain(COLS);
    std::vector<uint8_t> signs(CELLS);
    bsfo_coarse_arrays(seed, ox, oz, n1.data(), n2.data(), n3.data(), n4.data(), n5.data(), temp.data(),
                       rain.data(), signs.data());
    if (signs_out) std::memcpy(signs_out, signs.data(), signs.size());
    return coarseScoreSigns(signs.data());
}

}  // extern "C"

// ---------------------------------------------------------------------------
// FP64 operation count of island Stage 2: the work the staged pipeline does for one Stage-1
// survivor of a window, written once over a number type T.  With T = double it computes the
// window's lattice densities at levels 8..15 (checked bit for bit against bsfo_chunk_density);
// with T = CountedDouble it tallies the FP64 operations that computation executes:
//  - per 3D octave (noise2, noise3: 16 each; noise1: 8): the octave's scales and weight, the
//    y axis of levels 8..15 once (with the walk back to the start of level 8's cell), then per
//    lattice column the x and z axes, the corner gradients and x-lerps of every Y cell
//    (generateNoise's legacy bulk method, NoiseGeneratorPerlin173.java:224-283, which recomputes
//    them only when the masked y index changes), and per level the y / z lerps, the weight and
//    the accumulation;
//  - per 2D octave (noise4: 10, noise5: 16) and lattice column the yLen == 1 branch (:188-223);
//  - per density column (25 per chunk, climate at block (3i+1, 3l+1), BetaTerrain173.java:741-749)
//    the 10 simplex octaves (NoiseGenerator2173.java:117-172), WorldChunkManager173's climate
//    post-processing (:100-127), the column shape and the densities of levels 8..15 (:741-831).
// A Java gradient is one add (the sign flips are negations, exact and free); comparisons,
// int <-> double conversions, adds, subtractions, multiplications and divisions are counted.
// ---------------------------------------------------------------------------
namespace {

struct OpTally {
    uint64_t add = 0, mul = 0, div = 0, cmp = 0, cvt = 0;
};
thread_local OpTally* t_tally = nullptr;

struct CountedDouble {
    double v = 0.0;
    CountedDouble() = default;
    CountedDouble(double x) : v(x) {}  // constants: no operation
};
using CD = CountedDouble;
inline CD operator+(CD a, CD b) { ++t_tally->add; return CD(a.v + b.v); }
inline CD operator-(CD a, CD b) { ++t_tally->add; return CD(a.v - b.v); }
inline CD operator-(CD a) { return CD(-a.v); }
inline CD operator*(CD a, CD b) { ++t_tally->mul; return CD(a.v * b.v); }
inline CD operator/(CD a, CD b) { ++t_tally->div; return CD(a.v / b.v); }
inline bool operator<(CD a, CD b) { ++t_tally->cmp; return a.v < b.v; }
inline bool operator>(CD a, CD b) { ++t_tally->cmp; return a.v > b.v; }
inline bool operator<=(CD a, CD b) { ++t_tally->cmp; return a.v <= b.v; }

template <class T> T i2d(int i);
template <> inline double i2d<double>(int i) { return static_cast<double>(i); }
template <> inline CD i2d<CD>(int i) { ++t_tally->cvt; return CD(static_cast<double>(i)); }
inline int d2i(double v) { return jd2i(v); }
inline int d2i(CD v) { ++t_tally->cvt; return jd2i(v.v); }
inline double raw(double v) { return v; }
inline double raw(CD v) { return v.v; }

template <class T> T fadeT(T v) { return v * v * v * (v * (v * T(6.0) - T(15.0)) + T(10.0)); }
template <class T> T lerpT(T t, T a, T b) { return a + t * (b - a); }
template <class T> int floorT(T v) {
    int i = d2i(v);
    if (v < i2d<T>(i)) --i;
    return i;
}
template <class T> struct AxisT {
    int i;
    T f, fm1, fd;
};
template <class T> AxisT<T> axisT(T v) {
    AxisT<T> a;
    const int fl = floorT(v);
    a.i = fl & 255;
    a.f = v - i2d<T>(fl);
    a.fm1 = a.f - T(1.0);
    a.fd = fadeT(a.f);
    return a;
}
template <class T> T grad3T(int h, T x, T y, T z) {  // :145-165
    switch (h & 15) {
        case 0: return x + y;
        case 1: return -x + y;
        case 2: return x + -y;
        case 3: return -x + -y;
        case 4: return x + z;
        case 5: return -x + z;
        case 6: return x + -z;
        case 7: return -x + -z;
        case 8: return y + z;
        case 9: return -y + z;
        case 10: return y + -z;
        case 11: return -y + -z;
        case 12: return y + x;
        case 13: return -y + z;
        case 14: return y + -x;
        default: return -y + -z;
    }
}
template <class T> T grad2T(int h, T x, T z) {  // :123-143
    const T zeroX = T(0.0) * x;
    switch (h & 15) {
        case 0: return x + T(0.0);
        case 1: return -x + T(0.0);
        case 2: return x + T(-0.0);
        case 3: return -x + T(-0.0);
        case 4: return x + z;
        case 5: return -x + z;
        case 6: return x + -z;
        case 7: return -x + -z;
        case 8: return zeroX + z;
        case 9: return -zeroX + z;
        case 10: return zeroX + -z;
        case 11: return -zeroX + -z;
        case 12: return zeroX + x;
        case 13: return -zeroX + z;
        case 14: return zeroX + -x;
        default: return -zeroX + -z;
    }
}

constexpr int kS2Y0 = 8, kS2Levels = 8;

// One 3D octave group over the window's lattice columns, levels 8..15; acc[col][k].
template <class T>
void stage2Octaves3d(const Octaves& oct, double sxz, double sy0, int x0, int z0, int nx, int nz,
                     std::vector<std::array<T, kS2Levels>>& acc) {
    T amp = T(1.0);
    for (size_t o = 0; o < oct.g.size(); ++o) {
        const Perlin& p = oct.g[o];
        const T sx = T(sxz) * amp, sy = T(sy0) * amp, w = T(1.0) / amp;
        // the y axis of levels 8..15, once per octave; a level starts a cell when its masked
        // index differs from the level below (level 8's cell may start below level 8: its
        // corners use that level's fraction)
        int yi[kS2Levels];
        bool fresh[kS2Levels];
        T cell_f[kS2Levels], fd[kS2Levels];
        for (int k = 0; k < kS2Levels; ++k) {
            const T v = i2d<T>(kS2Y0 + k) * sy + T(p.b);
            const int fl = floorT(v);
            yi[k] = fl & 255;
            const T f = v - i2d<T>(fl);
            fd[k] = fadeT(f);
            fresh[k] = k == 0 || yi[k] != yi[k - 1];
            cell_f[k] = fresh[k] ? f : cell_f[k - 1];
        }
        for (int y = kS2Y0 - 1; y >= 0; --y) {
            const T v = i2d<T>(y) * sy + T(p.b);
            const int fl = floorT(v);
            if ((fl & 255) != yi[0]) break;
            cell_f[0] = v - i2d<T>(fl);
        }
        T cell_fm1[kS2Levels];
        for (int k = 0; k < kS2Levels; ++k)
            if (fresh[k]) cell_fm1[k] = cell_f[k] - T(1.0);
            else cell_fm1[k] = cell_fm1[k - 1];
        for (int u = 0; u < nx; ++u) {
            const AxisT<T> ax = axisT(i2d<T>(x0 + u) * sx + T(p.a));
            for (int v = 0; v < nz; ++v) {
                const AxisT<T> az = axisT(i2d<T>(z0 + v) * sx + T(p.c));
                auto& out = acc[u * nz + v];
                T q00 = T(0.0), q01 = T(0.0), q10 = T(0.0), q11 = T(0.0);
                for (int k = 0; k < kS2Levels; ++k) {
                    if (fresh[k]) {
                        const int p60 = p.d[ax.i] + yi[k], p61 = p.d[p60] + az.i, p62 = p.d[p60 + 1] + az.i;
                        const int p63 = p.d[ax.i + 1] + yi[k], p33 = p.d[p63] + az.i, p64 = p.d[p63 + 1] + az.i;
                        const T yf = cell_f[k], yfm1 = cell_fm1[k];
                        q00 = lerpT(ax.fd, grad3T(p.d[p61], ax.f, yf, az.f), grad3T(p.d[p33], ax.fm1, yf, az.f));
                        q01 = lerpT(ax.fd, grad3T(p.d[p62], ax.f, yfm1, az.f), grad3T(p.d[p64], ax.fm1, yfm1, az.f));
                        q10 = lerpT(ax.fd, grad3T(p.d[p61 + 1], ax.f, yf, az.fm1),
                                    grad3T(p.d[p33 + 1], ax.fm1, yf, az.fm1));
                        q11 = lerpT(ax.fd, grad3T(p.d[p62 + 1], ax.f, yfm1, az.fm1),
                                    grad3T(p.d[p64 + 1], ax.fm1, yfm1, az.fm1));
                    }
                    const T r0 = lerpT(fd[k], q00, q01), r1 = lerpT(fd[k], q10, q11);
                    const T val = lerpT(az.fd, r0, r1) * w;
                    out[k] = o == 0 ? T(0.0) + val : out[k] + val;
                }
            }
        }
        amp = amp / T(2.0);
    }
}

// One 2D octave group (generateNoise's yLen == 1 overload) over the window's lattice columns.
template <class T>
void stage2Octaves2d(const Octaves& oct, double sc, int x0, int z0, int nx, int nz, std::vector<T>& acc) {
    T amp = T(1.0);
    for (size_t o = 0; o < oct.g.size(); ++o) {
        const Perlin& p = oct.g[o];
        const T sx = T(sc) * amp, w = T(1.0) / amp;
        for (int u = 0; u < nx; ++u) {
            const AxisT<T> ax = axisT(i2d<T>(x0 + u) * sx + T(p.a));
            for (int v = 0; v < nz; ++v) {
                const AxisT<T> az = axisT(i2d<T>(z0 + v) * sx + T(p.c));
                const int p47 = p.d[p.d[ax.i]] + az.i, p20 = p.d[p.d[ax.i + 1]] + az.i;
                const T q0 = lerpT(ax.fd, grad2T(p.d[p47], ax.f, az.f), grad3T(p.d[p20], ax.fm1, T(0.0), az.f));
                const T q1 = lerpT(ax.fd, grad3T(p.d[p47 + 1], ax.f, T(0.0), az.fm1),
                                   grad3T(p.d[p20 + 1], ax.fm1, T(0.0), az.fm1));
                const T val = lerpT(az.fd, q0, q1) * w;
                T& out = acc[u * nz + v];
                out = o == 0 ? T(0.0) + val : out + val;
            }
        }
        amp = amp / T(2.0);
    }
}

template <class T> T simplexGradT(int g, T x, T z) {  // gradient2, NoiseGenerator2173.java:58-72
    switch (g) {
        case 0: return T(1.0) * x + T(1.0) * z;
        case 1: return T(-1.0) * x + T(1.0) * z;
        case 2: return T(1.0) * x + T(-1.0) * z;
        case 3: return T(-1.0) * x + T(-1.0) * z;
        case 4:
        case 6: return T(1.0) * x + T(0.0) * z;
        case 5:
        case 7: return T(-1.0) * x + T(0.0) * z;
        case 8:
        case 10: return T(0.0) * x + T(1.0) * z;
        default: return T(0.0) * x + T(-1.0) * z;
    }
}
template <class T> int simplexFloorT(T v) { return v > T(0.0) ? d2i(v) : d2i(v) - 1; }
template <class T> T simplexSampleT(const Simplex& s, T d5, T d6) {  // :117-172
    const T f = T(kSimplexF), g = T(kSimplexG);
    const T d7 = (d5 + d6) * f;
    const int j1 = simplexFloorT(d5 + d7), k1 = simplexFloorT(d6 + d7);
    const T d8 = i2d<T>(j1 + k1) * g;
    const T d9 = i2d<T>(j1) - d8, d10 = i2d<T>(k1) - d8;
    const T d11 = d5 - d9, d12 = d6 - d10;
    const int b0 = d11 > d12 ? 1 : 0, b1 = 1 - b0;
    const T d13 = d11 - i2d<T>(b0) + g, d14 = d12 - i2d<T>(b1) + g;
    const T d15 = d11 - T(1.0) + T(2.0) * g, d16 = d12 - T(1.0) + T(2.0) * g;
    const int l1 = j1 & 255, i2 = k1 & 255;
    const int j2 = s.e[l1 + s.e[i2]] % 12, k2 = s.e[l1 + b0 + s.e[i2 + b1]] % 12, l2 = s.e[l1 + 1 + s.e[i2 + 1]] % 12;
    T d17 = T(0.5) - d11 * d11 - d12 * d12, d18, d19 = T(0.5) - d13 * d13 - d14 * d14, d20;
    T d21 = T(0.5) - d15 * d15 - d16 * d16, d22;
    if (d17 < T(0.0)) d18 = T(0.0);
    else { d17 = d17 * d17; d18 = d17 * d17 * simplexGradT(j2, d11, d12); }
    if (d19 < T(0.0)) d20 = T(0.0);
    else { d19 = d19 * d19; d20 = d19 * d19 * simplexGradT(k2, d13, d14); }
    if (d21 < T(0.0)) d22 = T(0.0);
    else { d21 = d21 * d21; d22 = d21 * d21 * simplexGradT(l2, d15, d16); }
    return T(70.0) * (d18 + d20 + d22);
}
template <class T>
T stage2ClimateGroup(const SimplexOctaves& so, double start, double octave_scale, int bx, int bz) {
    const T d2 = T(start) / T(1.5);
    T d6 = T(1.0), d7 = T(1.0), acc = T(0.0);
    for (size_t o = 0; o < so.g.size(); ++o) {
        const Simplex& s = so.g[o];
        const T sc = d2 * d7, w = T(0.55) / d6;
        const T v = simplexSampleT(s, i2d<T>(bx) * sc + T(s.a), i2d<T>(bz) * sc + T(s.b)) * w;
        acc = o == 0 ? T(0.0) + v : acc + v;
        d7 = d7 * T(octave_scale);
        d6 = d6 * T(0.5);
    }
    return acc;
}

// The survivor's Stage-2 work over window p; out[(chunk*25 + i*5 + l)*8 + k] = density of level
// 8 + k of density column (i, l) of chunk (chunk / chunks_z, chunk % chunks_z).
template <class T>
void stage2Window(const Terrain& t, const bsfo_island_params& p, T* out) {
    const int nx = 4 * p.chunks_x + 1, nz = 4 * p.chunks_z + 1, x0 = 4 * p.chunk_x0, z0 = 4 * p.chunk_z0;
    const int nc = nx * nz;
    std::vector<std::array<T, kS2Levels>> a2(nc), a3(nc), a1(nc);
    std::vector<T> a4(nc), a5(nc);
    stage2Octaves3d<T>(t.n2, 684.412, 684.412, x0, z0, nx, nz, a2);
    stage2Octaves3d<T>(t.n3, 684.412, 684.412, x0, z0, nx, nz, a3);
    stage2Octaves3d<T>(t.n1, 684.412 / 80.0, 684.412 / 160.0, x0, z0, nx, nz, a1);
    stage2Octaves2d<T>(t.n4, 1.121, x0, z0, nx, nz, a4);
    stage2Octaves2d<T>(t.n5, 200.0, x0, z0, nx, nz, a5);
    for (int ch = 0; ch < p.chunks_x * p.chunks_z; ++ch) {
        const int cx = p.chunk_x0 + ch / p.chunks_z, cz = p.chunk_z0 + ch % p.chunks_z;
        for (int i = 0; i < 5; ++i)
            for (int l = 0; l < 5; ++l) {
                const int bx = cx * 16 + 3 * i + 1, bz = cz * 16 + 3 * l + 1;
                const T tr = stage2ClimateGroup<T>(t.climate.e, 0.02500000037252903, 0.25, bx, bz);
                const T rr = stage2ClimateGroup<T>(t.climate.f, 0.05000000074505806, 0.3333333333333333, bx, bz);
                const T cr = stage2ClimateGroup<T>(t.climate.g, 0.25, 0.5882352941176471, bx, bz);
                // WorldChunkManager173.java:100-127
                const T c0 = cr * T(1.1) + T(0.5);
                T d3 = (tr * T(0.15) + T(0.7)) * (T(1.0) - T(0.01)) + c0 * T(0.01);
                T d4 = (rr * T(0.15) + T(0.5)) * (T(1.0) - T(0.0020)) + c0 * T(0.0020);
                d3 = T(1.0) - (T(1.0) - d3) * (T(1.0) - d3);
                if (d3 < T(0.0)) d3 = T(0.0);
                if (d4 < T(0.0)) d4 = T(0.0);
                if (d3 > T(1.0)) d3 = T(1.0);
                if (d4 > T(1.0)) d4 = T(1.0);
                // column shape, BetaTerrain173.java:751-790
                const int col = (ch / p.chunks_z * 4 + i) * nz + (ch % p.chunks_z * 4 + l);
                const T rain = d4 * d3;
                T e4 = T(1.0) - rain;
                e4 = e4 * e4;
                e4 = e4 * e4;
                e4 = T(1.0) - e4;
                T d5 = (a4[col] + T(256.0)) / T(512.0);
                d5 = d5 * e4;
                if (d5 > T(1.0)) d5 = T(1.0);
                T d6 = a5[col] / T(8000.0);
                if (d6 < T(0.0)) d6 = -d6 * T(0.3);
                d6 = d6 * T(3.0) - T(2.0);
                if (d6 < T(0.0)) {
                    d6 = d6 / T(2.0);
                    if (d6 < T(-1.0)) d6 = T(-1.0);
                    d6 = d6 / T(1.4);
                    d6 = d6 / T(2.0);
                    d5 = T(0.0);
                } else {
                    if (d6 > T(1.0)) d6 = T(1.0);
                    d6 = d6 / T(8.0);
                }
                if (d5 < T(0.0)) d5 = T(0.0);
                d5 = d5 + T(0.5);
                d6 = d6 * T(17.0) / T(16.0);
                const T d7 = T(17.0) / T(2.0) + d6 * T(4.0);
                for (int k = 0; k < kS2Levels; ++k) {
                    const int y = kS2Y0 + k;
                    T d9 = (i2d<T>(y) - d7) * T(12.0) / d5;
                    if (d9 < T(0.0)) d9 = d9 * T(4.0);
                    const T d10 = a2[col][k] / T(512.0), d11 = a3[col][k] / T(512.0);
                    const T d12 = (a1[col][k] / T(10.0) + T(1.0)) / T(2.0);
                    T d8;
                    if (d12 < T(0.0)) d8 = d10;
                    else if (d12 > T(1.0)) d8 = d11;
                    else d8 = d10 + (d11 - d10) * d12;
                    d8 = d8 - d9;
                    if (y > 13) {
                        const T d13 = T(static_cast<double>(static_cast<float>(y - 13) / 3.0F));
                        d8 = d8 * (T(1.0) - d13) + T(-10.0) * d13;
                    }
                    out[(ch * 25 + i * 5 + l) * kS2Levels + k] = d8;
                }
            }
    }
}

// ---------------------------------------------------------------------------
// FP64 operation count of island Stage 1 for one seed: Stage 1a at the window's sample columns
// (lattice columns (4 cx + i, 4 cz + l), i, l in {1, 3}: blocks 4 / 12, IslandSearchEngine.java:
// 4005-4007) -- noise1, noise4, noise5 and the climate in full, the column shape, and per level the
// height bias and the noise1 blend; then noise2 / noise3 over octaves 11..15, blended per level
// with the clamped noise1 blend factor, and the density bounds that allow for the left-out octaves
// (each octave o of a 3D Perlin sum is at most 2 * 2^o in magnitude, so octaves 0..10 add at most
// 4094); only if some column of the seed is still open, octave 10 of both and the bounds again
// (octaves 0..9: 2046) -- then Stage 1b at each column the bounds leave open: noise2 / noise3 over
// all 16 octaves, the exact densities and generateBareTerrain's y-chain (BetaTerrain173.java:
// 599-664) looking for stone at y >= 68 over non-stone.  A column is open when some level k has
// lo[k] <= 0 < hi[k + 1].  The x axis of an octave is formed once per distinct lattice x, the y
// axis once per octave (Stage 1a: shared by the seed's columns; Stage 1b: per column, each open
// column being evaluated alone).  Stage 1a's noise2 / noise3 octaves only feed the bounds and take
// one FMA per level from their Y cell's corners (octaves3dAt's bound form), as the kernel does.  As
// in the kernel, noise2 / noise3 over octaves 11..15 come before noise1: when the hull of the two
// sums (any blend factor in [0, 1]) closes every column, the seed ends without noise1.
// ---------------------------------------------------------------------------
constexpr int kSplitOctave1 = 10;
// |3D Perlin octave value| <= 1.54 per unit weight: a blend of the corner gradients (+/-u +/-v,
// two different corner offsets) with nonnegative fade weights, the y offsets at most 1 (the Y-cell
// quirk's stale fraction included), so |value| <= 1 + max over (fx, fz) of
// sum_{x,z corners} w_x w_z max(|dx|, |dz|) = 1 + 0.5353 (tests/test_oracle_pinning.py checks the
// maximum and samples of the octave values).
constexpr double kOctaveBound1 = 1.54;

template <class T> T fmaT(T a, T b, T c);
template <> inline double fmaT<double>(double a, double b, double c) { return std::fma(a, b, c); }
template <> inline CD fmaT<CD>(CD a, CD b, CD c) {
    ++t_tally->mul;
    ++t_tally->add;
    return CD(std::fma(a.v, b.v, c.v));
}

// 3D octaves [o0, o1) of one group at lattice columns cols, levels 8..15, accumulated into acc
// (the first evaluated octave starts each sum at 0.0 + value).  With bound (values that only feed
// Stage 1a's bounds), each level is A + fd B from its Y cell's weighted A = lerp(fz, q00, q10) and
// B = (q01 - q00) + fz ((q11 - q10) - (q01 - q00)), one FMA (counted as a multiply and an add),
// as the Stage-1a kernel forms it; otherwise Java's three lerps.
template <class T>
void octaves3dAt(const Octaves& oct, double sxz, double sy0, int o0, int o1, const std::vector<std::pair<int, int>>& cols,
                 std::vector<std::array<T, kS2Levels>>& acc, bool bound = false) {
    T amp = T(1.0);
    for (int o = 0; o < o0; ++o) amp = amp / T(2.0);
    for (int o = o0; o < o1; ++o) {
        const Perlin& p = oct.g[static_cast<size_t>(o)];
        const T sx = T(sxz) * amp, sy = T(sy0) * amp, w = T(1.0) / amp;
        int yi[kS2Levels];
        bool fresh[kS2Levels];
        T cell_f[kS2Levels], fd[kS2Levels];
        for (int k = 0; k < kS2Levels; ++k) {
            const T v = i2d<T>(kS2Y0 + k) * sy + T(p.b);
            const int fl = floorT(v);
            yi[k] = fl & 255;
            const T f = v - i2d<T>(fl);
            fd[k] = fadeT(f);
            fresh[k] = k == 0 || yi[k] != yi[k - 1];
            cell_f[k] = fresh[k] ? f : cell_f[k - 1];
        }
        for (int y = kS2Y0 - 1; y >= 0; --y) {
            const T v = i2d<T>(y) * sy + T(p.b);
            const int fl = floorT(v);
            if ((fl & 255) != yi[0]) break;
            cell_f[0] = v - i2d<T>(fl);
        }
        T cell_fm1[kS2Levels];
        for (int k = 0; k < kS2Levels; ++k)
            if (fresh[k]) cell_fm1[k] = cell_f[k] - T(1.0);
            else cell_fm1[k] = cell_fm1[k - 1];
        std::map<int, AxisT<T>> xaxes;
        for (size_t c = 0; c < cols.size(); ++c) {
            const int X = cols[c].first, Z = cols[c].second;
            auto xa = xaxes.find(X);
            if (xa == xaxes.end()) xa = xaxes.emplace(X, axisT(i2d<T>(X) * sx + T(p.a))).first;
            const AxisT<T>& ax = xa->second;
            const AxisT<T> az = axisT(i2d<T>(Z) * sx + T(p.c));
            auto& out = acc[c];
            T q00 = T(0.0), q01 = T(0.0), q10 = T(0.0), q11 = T(0.0), A = T(0.0), B = T(0.0);
            for (int k = 0; k < kS2Levels; ++k) {
                if (fresh[k]) {
                    const int p60 = p.d[ax.i] + yi[k], p61 = p.d[p60] + az.i, p62 = p.d[p60 + 1] + az.i;
                    const int p63 = p.d[ax.i + 1] + yi[k], p33 = p.d[p63] + az.i, p64 = p.d[p63 + 1] + az.i;
                    const T yf = cell_f[k], yfm1 = cell_fm1[k];
                    q00 = lerpT(ax.fd, grad3T(p.d[p61], ax.f, yf, az.f), grad3T(p.d[p33], ax.fm1, yf, az.f));
                    q01 = lerpT(ax.fd, grad3T(p.d[p62], ax.f, yfm1, az.f), grad3T(p.d[p64], ax.fm1, yfm1, az.f));
                    q10 = lerpT(ax.fd, grad3T(p.d[p61 + 1], ax.f, yf, az.fm1), grad3T(p.d[p33 + 1], ax.fm1, yf, az.fm1));
                    q11 = lerpT(ax.fd, grad3T(p.d[p62 + 1], ax.f, yfm1, az.fm1),
                                grad3T(p.d[p64 + 1], ax.fm1, yfm1, az.fm1));
                    if (bound) {
                        const T d0 = q01 - q00, d1 = q11 - q10;
                        A = lerpT(az.fd, q00, q10) * w;
                        B = (d0 + az.fd * (d1 - d0)) * w;
                    }
                }
                T val;
                if (bound) {
                    val = fmaT(fd[k], B, A);
                } else {
                    const T r0 = lerpT(fd[k], q00, q01), r1 = lerpT(fd[k], q10, q11);
                    val = lerpT(az.fd, r0, r1) * w;
                }
                out[k] = o == o0 ? T(0.0) + val : out[k] + val;
            }
        }
        amp = amp / T(2.0);
    }
}

// The 2D octaves of one group at lattice columns cols.
template <class T>
void octaves2dAt(const Octaves& oct, double sc, const std::vector<std::pair<int, int>>& cols, std::vector<T>& acc) {
    T amp = T(1.0);
    for (size_t o = 0; o < oct.g.size(); ++o) {
        const Perlin& p = oct.g[o];
        const T sx = T(sc) * amp, w = T(1.0) / amp;
        std::map<int, AxisT<T>> xaxes;
        for (size_t c = 0; c < cols.size(); ++c) {
            auto xa = xaxes.find(cols[c].first);
            if (xa == xaxes.end()) xa = xaxes.emplace(cols[c].first, axisT(i2d<T>(cols[c].first) * sx + T(p.a))).first;
            const AxisT<T>& ax = xa->second;
            const AxisT<T> az = axisT(i2d<T>(cols[c].second) * sx + T(p.c));
            const int p47 = p.d[p.d[ax.i]] + az.i, p20 = p.d[p.d[ax.i + 1]] + az.i;
            const T q0 = lerpT(ax.fd, grad2T(p.d[p47], ax.f, az.f), grad3T(p.d[p20], ax.fm1, T(0.0), az.f));
            const T q1 = lerpT(ax.fd, grad3T(p.d[p47 + 1], ax.f, T(0.0), az.fm1),
                               grad3T(p.d[p20 + 1], ax.fm1, T(0.0), az.fm1));
            const T val = lerpT(az.fd, q0, q1) * w;
            acc[c] = o == 0 ? T(0.0) + val : acc[c] + val;
        }
        amp = amp / T(2.0);
    }
}

template <class T> T taperT(T d8, int y) {
    if (y > 13) {
        const T d13 = T(static_cast<double>(static_cast<float>(y - 13) / 3.0F));
        d8 = d8 * (T(1.0) - d13) + T(-10.0) * d13;
    }
    return d8;
}

template <class T> T densityFromT(T d9, T d12, T n2, T n3, int y) {
    const T d10 = n2 / T(512.0), d11 = n3 / T(512.0);
    T d8;
    if (d12 < T(0.0)) d8 = d10;
    else if (d12 > T(1.0)) d8 = d11;
    else d8 = d10 + (d11 - d10) * d12;
    return taperT(d8 - d9, y);
}

// Stage 1 of one seed over window p: open[c] (Stage 1a left column c open), found[c] (Stage 1b
// found a candidate start there), lattice[c * 8 + k] the column's exact densities at levels 8..15
// (plain evaluation only, for checking).  Stage 1a's operations go to *t1a, Stage 1b's to *t1b.
template <class T>
void stage1Seed(const Terrain& t, const bsfo_island_params& p, OpTally* t1a, OpTally* t1b, uint8_t* open,
                uint8_t* found, double* lattice) {
    std::vector<std::pair<int, int>> cols;
    std::vector<std::pair<int, int>> blocks;
    for (int ch = 0; ch < p.chunks_x * p.chunks_z; ++ch) {
        const int cx = p.chunk_x0 + ch / p.chunks_z, cz = p.chunk_z0 + ch % p.chunks_z;
        for (int sc = 0; sc < 4; ++sc) {
            const int i = (sc >> 1) ? 3 : 1, l = (sc & 1) ? 3 : 1;
            cols.emplace_back(cx * 4 + i, cz * 4 + l);
            blocks.emplace_back(cx * 16 + 3 * i + 1, cz * 16 + 3 * l + 1);
        }
    }
    const size_t nc = cols.size();
    // Stage 1a: noise4 / noise5 / climate (the height biases), noise2 / noise3 over octaves 11..15,
    // the hull check; noise1 (the blend factors) only if some column is still open
    t_tally = t1a;
    std::vector<std::array<T, kS2Levels>> a1(nc);
    std::vector<T> a4(nc), a5(nc);
    // noise5 first: when every column takes the low branch of the column shape (d6 * 3 - 2 < 0,
    // where d5 is set to 0 whatever noise4 and the climate are), those are not evaluated
    octaves2dAt<T>(t.n5, 200.0, cols, a5);
    auto shape_d6 = [](T n5) {
        T d6 = n5 / T(8000.0);
        if (d6 < T(0.0)) d6 = -d6 * T(0.3);
        return d6 * T(3.0) - T(2.0);
    };
    bool all_low = true;
    for (size_t c = 0; c < nc; ++c) all_low = all_low && shape_d6(a5[c]) < T(0.0);
    const T slack = T(1e-6);
    std::vector<std::array<T, kS2Levels>> d9s(nc), d12s(nc);
    if (!all_low) octaves2dAt<T>(t.n4, 1.121, cols, a4);
    for (size_t c = 0; c < nc; ++c) {
        if (all_low) {
            T d6 = shape_d6(a5[c]) / T(2.0);
            if (d6 < T(-1.0)) d6 = T(-1.0);
            d6 = d6 / T(1.4);
            d6 = d6 / T(2.0);
            d6 = d6 * T(17.0) / T(16.0);
            const T d7 = T(17.0) / T(2.0) + d6 * T(4.0);
            for (int k = 0; k < kS2Levels; ++k) {
                T d9 = (i2d<T>(kS2Y0 + k) - d7) * T(12.0) / T(0.5);
                if (d9 < T(0.0)) d9 = d9 * T(4.0);
                d9s[c][k] = d9;
            }
            continue;
        }
        const T tr = stage2ClimateGroup<T>(t.climate.e, 0.02500000037252903, 0.25, blocks[c].first, blocks[c].second);
        const T rr = stage2ClimateGroup<T>(t.climate.f, 0.05000000074505806, 0.3333333333333333, blocks[c].first,
                                           blocks[c].second);
        const T cr = stage2ClimateGroup<T>(t.climate.g, 0.25, 0.5882352941176471, blocks[c].first, blocks[c].second);
        const T c0 = cr * T(1.1) + T(0.5);
        T d3 = (tr * T(0.15) + T(0.7)) * (T(1.0) - T(0.01)) + c0 * T(0.01);
        T d4 = (rr * T(0.15) + T(0.5)) * (T(1.0) - T(0.0020)) + c0 * T(0.0020);
        d3 = T(1.0) - (T(1.0) - d3) * (T(1.0) - d3);
        if (d3 < T(0.0)) d3 = T(0.0);
        if (d4 < T(0.0)) d4 = T(0.0);
        if (d3 > T(1.0)) d3 = T(1.0);
        if (d4 > T(1.0)) d4 = T(1.0);
        const T rain = d4 * d3;
        T e4 = T(1.0) - rain;
        e4 = e4 * e4;
        e4 = e4 * e4;
        e4 = T(1.0) - e4;
        T d5 = (a4[c] + T(256.0)) / T(512.0);
        d5 = d5 * e4;
        if (d5 > T(1.0)) d5 = T(1.0);
        T d6 = a5[c] / T(8000.0);
        if (d6 < T(0.0)) d6 = -d6 * T(0.3);
        d6 = d6 * T(3.0) - T(2.0);
        if (d6 < T(0.0)) {
            d6 = d6 / T(2.0);
            if (d6 < T(-1.0)) d6 = T(-1.0);
            d6 = d6 / T(1.4);
            d6 = d6 / T(2.0);
            d5 = T(0.0);
        } else {
            if (d6 > T(1.0)) d6 = T(1.0);
            d6 = d6 / T(8.0);
        }
        if (d5 < T(0.0)) d5 = T(0.0);
        d5 = d5 + T(0.5);
        d6 = d6 * T(17.0) / T(16.0);
        const T d7 = T(17.0) / T(2.0) + d6 * T(4.0);
        for (int k = 0; k < kS2Levels; ++k) {
            const int y = kS2Y0 + k;
            T d9 = (i2d<T>(y) - d7) * T(12.0) / d5;
            if (d9 < T(0.0)) d9 = d9 * T(4.0);
            d9s[c][k] = d9;
        }
    }
    // noise2 / noise3, blended: b = (1 - cl) n2 + cl n3 with cl = d12 clamped to [0, 1] (for d12 in
    // [0, 1] the density blend n2/512 + (n3/512 - n2/512) d12 is increasing in both sums, so its
    // bounds are (b -/+ u) / 512; outside it is one sum alone).  Each group's octaves are summed
    // first, then folded in.
    std::vector<std::array<T, kS2Levels>> b(nc), g(nc), g2(nc);
    for (size_t c = 0; c < nc; ++c)
        for (int k = 0; k < kS2Levels; ++k) b[c][k] = T(0.0);
    auto fold = [&](bool n3) {
        for (size_t c = 0; c < nc; ++c)
            for (int k = 0; k < kS2Levels; ++k) {
                const T d = d12s[c][k];
                const T cl = d < T(0.0) ? T(0.0) : (d > T(1.0) ? T(1.0) : d);
                b[c][k] = b[c][k] + (n3 ? cl * g[c][k] : g[c][k] - cl * g[c][k]);
            }
    };
    auto bounds_open = [&](size_t c, T u) {
        T lo[kS2Levels + 1], hi[kS2Levels + 1];
        for (int k = 0; k < kS2Levels; ++k) {
            const int y = kS2Y0 + k;
            lo[k] = taperT((b[c][k] - u) / T(512.0) - d9s[c][k], y) - slack;
            hi[k] = taperT((b[c][k] + u) / T(512.0) - d9s[c][k], y) + slack;
        }
        lo[kS2Levels] = hi[kS2Levels] = T(-10.0);
        bool op = false;
        for (int k = 0; k < kS2Levels; ++k) {
            const bool a = lo[k] <= T(0.0);
            const bool bb = hi[k + 1] > T(0.0);
            op = op || (a && bb);
        }
        return op;
    };
    // the hull of the two sums: for any blend factor in [0, 1] the blend lies between them
    auto hull_open = [&](size_t c, T u) {
        T lo[kS2Levels + 1], hi[kS2Levels + 1];
        for (int k = 0; k < kS2Levels; ++k) {
            const int y = kS2Y0 + k;
            const T mn = g2[c][k] < g[c][k] ? g2[c][k] : g[c][k];
            const T mx = g2[c][k] < g[c][k] ? g[c][k] : g2[c][k];
            lo[k] = taperT((mn - u) / T(512.0) - d9s[c][k], y) - slack;
            hi[k] = taperT((mx + u) / T(512.0) - d9s[c][k], y) + slack;
        }
        lo[kS2Levels] = hi[kS2Levels] = T(-10.0);
        bool op = false;
        for (int k = 0; k < kS2Levels; ++k) {
            const bool a = lo[k] <= T(0.0);
            const bool bb = hi[k + 1] > T(0.0);
            op = op || (a && bb);
        }
        return op;
    };
    const T coarse = T(kOctaveBound1 * ((1 << (kSplitOctave1 + 1)) - 1));
    octaves3dAt<T>(t.n2, 684.412, 684.412, kSplitOctave1 + 1, 16, cols, g2, true);
    octaves3dAt<T>(t.n3, 684.412, 684.412, kSplitOctave1 + 1, 16, cols, g, true);
    bool any = false;
    for (size_t c = 0; c < nc; ++c) {
        open[c] = hull_open(c, coarse) ? 1 : 0;
        any = any || open[c];
    }
    auto blend_factors = [&]() {
        octaves3dAt<T>(t.n1, 684.412 / 80.0, 684.412 / 160.0, 0, 8, cols, a1);
        for (size_t c = 0; c < nc; ++c)
            for (int k = 0; k < kS2Levels; ++k) d12s[c][k] = (a1[c][k] / T(10.0) + T(1.0)) / T(2.0);
    };
    const bool with_n1 = any;
    if (any) {
        // noise1, the blend factors, and the two groups folded in (noise2 with weight 1 - c, then
        // noise3 with weight c)
        blend_factors();
        std::swap(g, g2);
        fold(false);
        std::swap(g, g2);
        fold(true);
        any = false;
        for (size_t c = 0; c < nc; ++c) {
            if (open[c]) open[c] = bounds_open(c, coarse) ? 1 : 0;
            any = any || open[c];
        }
    }
    if (any) {
        octaves3dAt<T>(t.n2, 684.412, 684.412, kSplitOctave1, kSplitOctave1 + 1, cols, g, true);
        fold(false);
        octaves3dAt<T>(t.n3, 684.412, 684.412, kSplitOctave1, kSplitOctave1 + 1, cols, g, true);
        fold(true);
        for (size_t c = 0; c < nc; ++c)
            if (open[c]) open[c] = bounds_open(c, T(kOctaveBound1 * ((1 << kSplitOctave1) - 1))) ? 1 : 0;
    }
    // Stage 1b at the open columns
    t_tally = t1b;
    for (size_t c = 0; c < nc; ++c) {
        found[c] = 0;
        if (!open[c]) continue;
        std::vector<std::pair<int, int>> one(1, cols[c]);
        std::vector<std::array<T, kS2Levels>> b2(1), b3(1);
        octaves3dAt<T>(t.n2, 684.412, 684.412, 0, 16, one, b2);
        octaves3dAt<T>(t.n3, 684.412, 684.412, 0, 16, one, b3);
        T lat[kS2Levels + 1];
        for (int k = 0; k < kS2Levels; ++k) lat[k] = densityFromT(d9s[c][k], d12s[c][k], b2[0][k], b3[0][k], kS2Y0 + k);
        lat[kS2Levels] = T(-10.0);
        bool below = false, f = false;
        for (int k1 = 0; k1 < kS2Levels; ++k1) {
            T d1 = lat[k1];
            const T d5 = (lat[k1 + 1] - d1) * T(0.125);
            for (int l1 = 0; l1 < 8; ++l1) {
                const bool stone = d1 > T(0.0);
                if (64 + k1 * 8 + l1 >= kIslandFloorY && stone && !below) f = true;
                below = stone;
                d1 = d1 + d5;
            }
        }
        found[c] = f ? 1 : 0;
    }
    t_tally = nullptr;
    // the exact densities of every column, for checking (not counted)
    if (lattice) {
        OpTally scratch;
        t_tally = &scratch;
        if (!with_n1) blend_factors();
        std::vector<std::array<T, kS2Levels>> b2(nc), b3(nc);
        octaves3dAt<T>(t.n2, 684.412, 684.412, 0, 16, cols, b2);
        octaves3dAt<T>(t.n3, 684.412, 684.412, 0, 16, cols, b3);
        for (size_t c = 0; c < nc; ++c)
            for (int k = 0; k < kS2Levels; ++k)
                lattice[c * kS2Levels + k] = raw(densityFromT(d9s[c][k], d12s[c][k], b2[c][k], b3[c][k], kS2Y0 + k));
        t_tally = nullptr;
    }
}

}  // namespace

extern "C" {

void bsfo_stage1_op_count(int64_t seed, const bsfo_island_params* p, uint64_t* counts1a, uint64_t* counts1b,
                          uint8_t* open_out, uint8_t* found_out, double* lattice_out) {
    Terrain t(seed);
    const int nc = p->chunks_x * p->chunks_z * 4;
    OpTally a, b;
    std::vector<uint8_t> open(nc), found(nc);
    stage1Seed<CD>(t, *p, &a, &b, open.data(), found.data(), nullptr);
    for (int i = 0; i < 2; ++i) {
        const OpTally& tl = i == 0 ? a : b;
        uint64_t* out = i == 0 ? counts1a : counts1b;
        out[0] = tl.add;
        out[1] = tl.mul;
        out[2] = tl.div;
        out[3] = tl.cmp;
        out[4] = tl.cvt;
    }
    if (open_out || found_out || lattice_out) {
        // the plain evaluation takes the same decisions
        OpTally pa2, pb2;
        std::vector<uint8_t> open2(nc), found2(nc);
        std::vector<double> lat(static_cast<size_t>(nc) * kS2Levels);
        stage1Seed<double>(t, *p, &pa2, &pb2, open2.data(), found2.data(), lat.data());
        for (int c = 0; c < nc; ++c) {
            if (open_out) open_out[c] = open[c] == open2[c] ? open[c] : 255;
            if (found_out) found_out[c] = found[c] == found2[c] ? found[c] : 255;
        }
        if (lattice_out) std::memcpy(lattice_out, lat.data(), lat.size() * sizeof(double));
    }
}

void bsfo_stage2_op_count(int64_t seed, const bsfo_island_params* p, uint64_t* counts5, double* lattice_out) {
    Terrain t(seed);
    const int n = p->chunks_x * p->chunks_z * 25 * kS2Levels;
    OpTally tally;
    t_tally = &tally;
    std::vector<CD> counted(n);
    stage2Window<CD>(t, *p, counted.data());
    t_tally = nullptr;
    counts5[0] = tally.add;
    counts5[1] = tally.mul;
    counts5[2] = tally.div;
    counts5[3] = tally.cmp;
    counts5[4] = tally.cvt;
    if (lattice_out) {
        std::vector<double> plain(n);
        stage2Window<double>(t, *p, plain.data());
        for (int i = 0; i < n; ++i) lattice_out[i] = plain[i];
        // the counted evaluation computes the same values
        for (int i = 0; i < n; ++i)
            if (std::memcmp(&plain[i], &counted[i].v, 8) != 0) lattice_out[i] = std::numeric_limits<double>::quiet_NaN();
    }
}

}  // extern "C"
