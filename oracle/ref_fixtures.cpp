// ref_fixtures.cpp -- writes golden fixtures from the reference's own C++ host code.
//
// Built by `make -C oracle ref` against the reference sources where they lie
// (native/src of the reference: p20_exact_math.hpp, coarse_exact_core.hpp,
// p19_monster_gate.hpp -- all host-callable inline code); the binary goes to oracle/_ref/
// and is run only in the development container.  Its outputs (small JSON files under
// tests/golden/) are data: inputs and the reference's outputs for them.  Nothing here
// is shipped or run on the GPU host.
//
//   ./_ref/ref_fixtures <out_dir>
#include "coarse_exact_core.hpp"
#include "p19_monster_gate.hpp"
#include "p20_exact_math.hpp"

#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

namespace {

uint64_t bits(double v) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    return b;
}

// FNV-1a over the little-endian bytes of a sequence of 64-bit words / bytes.
struct Fnv {
    uint64_t h = 1469598103934665603ULL;
    void byte(uint8_t b) { h = (h ^ b) * 1099511628211ULL; }
    void u64(uint64_t v) { for (int i = 0; i < 8; ++i) byte(static_cast<uint8_t>(v >> (8 * i))); }
    void f64(double v) { u64(bits(v)); }
};

FILE* open_out(const std::string& dir, const char* name) {
    std::string path = dir + "/" + name;
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::perror(path.c_str());
        std::exit(1);
    }
    return f;
}

const int64_t kSeeds[] = {0, 1, 42, -1, 1611, 6554, 123456789, -987654321, 140737488355327LL, 281474976710656LL + 5};

void random_fixture(const std::string& dir) {
    FILE* f = open_out(dir, "ref_random.json");
    std::fprintf(f, "{\"cases\": [\n");
    bool first = true;
    for (int64_t seed : kSeeds) {
        p20::JavaRandom r;
        r.setSeed(seed);
        std::fprintf(f, "%s {\"seed\": %" PRId64 ", \"doubles\": [", first ? "" : ",", seed);
        for (int i = 0; i < 8; ++i) std::fprintf(f, "%s\"%016" PRIx64 "\"", i ? ", " : "", bits(r.nextDouble()));
        std::fprintf(f, "], \"ints\": [");
        // nextInt over a bound ladder (powers of two, odd, large, and the shuffle bounds)
        const int bounds[] = {256, 255, 254, 200, 129, 128, 100, 17, 12, 3, 2, 1, 1000000, 1073741825, 2147483647};
        int k = 0;
        for (int b : bounds)
            for (int rep = 0; rep < 3; ++rep) std::fprintf(f, "%s[%d, %d]", k++ ? ", " : "", b, r.nextInt(b));
        std::fprintf(f, "], \"bits31\": [");
        for (int i = 0; i < 8; ++i) std::fprintf(f, "%s%u", i ? ", " : "", r.nextBits(31));
        std::fprintf(f, "]}\n");
        first = false;
    }
    std::fprintf(f, "]}\n");
    std::fclose(f);
}

void dump_state(FILE* f, const p20::PerlinState& p) {
    std::fprintf(f, "{\"abc\": [\"%016" PRIx64 "\", \"%016" PRIx64 "\", \"%016" PRIx64 "\"], \"perm\": [", bits(p.a),
                 bits(p.b), bits(p.c));
    for (int i = 0; i < 256; ++i) std::fprintf(f, "%s%d", i ? "," : "", p.perm[i]);
    std::fprintf(f, "]}");
}

// Terrain generator states (Random(seed), consecutive initPerlin) and climate simplex states
// (Random(seed * {9871, 39811, 543321})), plus noise primitive values on them.
void perlin_fixture(const std::string& dir) {
    FILE* f = open_out(dir, "ref_perlin.json");
    std::mt19937_64 rng(20240601);
    std::uniform_real_distribution<double> coord(-3000.0, 3000.0);
    std::uniform_real_distribution<double> yscale(0.001, 700.0);
    std::fprintf(f, "{\"cases\": [\n");
    bool first = true;
    for (int64_t seed : {int64_t(0), int64_t(42), int64_t(1611), int64_t(-5), int64_t(140737488355331LL)}) {
        p20::JavaRandom r;
        r.setSeed(seed);
        std::fprintf(f, "%s {\"seed\": %" PRId64 ", \"terrain\": [", first ? "" : ",", seed);
        for (int g = 0; g < 4; ++g) {
            p20::PerlinState p;
            p20::initPerlin(r, p);
            std::fprintf(f, "%s{\"index\": %d, \"state\": ", g ? ",\n  " : "", g);
            dump_state(f, p);
            // point values: perlin3, a 17-level column (Beta Y-cell cache), perlin2, simplex2
            std::fprintf(f, ", \"p3\": [");
            for (int k = 0; k < 6; ++k) {
                double x = coord(rng), y = coord(rng) * 0.05, z = coord(rng);
                std::fprintf(f, "%s[\"%016" PRIx64 "\", \"%016" PRIx64 "\", \"%016" PRIx64 "\", \"%016" PRIx64 "\"]",
                             k ? ", " : "", bits(x), bits(y), bits(z), bits(p20::perlin3(p, x, y, z)));
            }
            std::fprintf(f, "], \"col17\": [");
            for (int k = 0; k < 4; ++k) {
                double x = coord(rng), sy = yscale(rng), z = coord(rng);
                if (k == 0) sy = 0.0208876953125;  // the finest noise2 octave: long Y cells
                double out[17];
                coarsecore::perlin3Full17(p, x, sy, z, out);
                std::fprintf(f, "%s{\"x\": \"%016" PRIx64 "\", \"sy\": \"%016" PRIx64 "\", \"z\": \"%016" PRIx64 "\", \"v\": [",
                             k ? ", " : "", bits(x), bits(sy), bits(z));
                for (int y = 0; y < 17; ++y) std::fprintf(f, "%s\"%016" PRIx64 "\"", y ? "," : "", bits(out[y]));
                std::fprintf(f, "]}");
            }
            std::fprintf(f, "], \"p2\": [");
            for (int k = 0; k < 6; ++k) {
                double x = coord(rng), z = coord(rng);
                std::fprintf(f, "%s[\"%016" PRIx64 "\", \"%016" PRIx64 "\", \"%016" PRIx64 "\"]", k ? ", " : "", bits(x),
                             bits(z), bits(p20::perlin2(p, x, z)));
            }
            std::fprintf(f, "], \"s2\": [");
            for (int k = 0; k < 6; ++k) {
                double x = coord(rng), z = coord(rng);
                std::fprintf(f, "%s[\"%016" PRIx64 "\", \"%016" PRIx64 "\", \"%016" PRIx64 "\"]", k ? ", " : "", bits(x),
                             bits(z), bits(p20::simplex2(p, x, z)));
            }
            std::fprintf(f, "]}");
        }
        std::fprintf(f, "], \"climate\": [");
        const uint64_t mul[3] = {9871ULL, 39811ULL, 543321ULL};
        for (int t = 0; t < 3; ++t) {
            p20::JavaRandom rc;
            rc.setSeed(static_cast<int64_t>(static_cast<uint64_t>(seed) * mul[t]));
            p20::PerlinState p;
            p20::initPerlin(rc, p);
            std::fprintf(f, "%s", t ? ", " : "");
            dump_state(f, p);
        }
        std::fprintf(f, "]}\n");
        first = false;
    }
    std::fprintf(f, "]}\n");
    std::fclose(f);
}

// generateAndScoreHost (coarse_exact_core.hpp:218-386): the 61x61x17 coarse grid around the
// origin (coarse*4+2 climate) and its floating-component score, with digests of every
// intermediate array.
std::string hexlist(const std::vector<double>& v) {
    std::string out = "[";
    char buf[32];
    for (size_t i = 0; i < v.size(); ++i) {
        std::snprintf(buf, sizeof(buf), "%s\"%016" PRIx64 "\"", i ? "," : "", bits(v[i]));
        out += buf;
    }
    return out + "]";
}

std::string chunk00_lattice(int64_t seed, const coarsecore::HostScratch& s) {
    std::vector<double> n1, n2, n3, n4, n5;
    for (int x = 0; x < 5; ++x)
        for (int z = 0; z < 5; ++z) {
            const int col = (x + 28) * coarsecore::SIZE + (z + 28);
            for (int y = 0; y < 17; ++y) {
                n1.push_back(s.noise1[col * 17 + y]);
                n2.push_back(s.noise2[col * 17 + y]);
                n3.push_back(s.noise3[col * 17 + y]);
            }
            n4.push_back(s.noise4[col]);
            n5.push_back(s.noise5[col]);
        }
    char head[64];
    std::snprintf(head, sizeof(head), "{\"seed\": %" PRId64, seed);
    return std::string(head) + ", \"n1\": " + hexlist(n1) + ", \"n2\": " + hexlist(n2) + ", \"n3\": " + hexlist(n3) +
           ", \"n4\": " + hexlist(n4) + ", \"n5\": " + hexlist(n5) + "}";
}

// The Stage0 lattice at coarse offset 0 is every fourth column of the coarse grid (coarse
// -28 + 4i, stage0_exact_gpu.hpp:48-63, and columnCoordinates, coarse_exact_core.hpp:74-81, with
// the same coarse*4+2 climate), so the grid's signs give each Stage0 column's 17-level positive
// mask.  The dense per-chunk climate (getBiomeNoise over a chunk's 16x16 blocks) meets the grid's
// climate samples at the blocks whose world coordinate is 2 mod 4: chunk-local blocks
// {2, 6, 10, 14}^2 of chunk (cx, cz) are the grid's columns cx*4 + {0..3}.
const int kClimateChunks[][2] = {{0, 0}, {-7, -7}, {7, 7}, {-7, 7}, {3, -2}, {-1, 4}};

std::string stage0_climate_row(int64_t seed, const coarsecore::HostScratch& s) {
    std::string out;
    char buf[64];
    std::snprintf(buf, sizeof(buf), "{\"seed\": %" PRId64 ", \"masks\": [", seed);
    out += buf;
    for (int ix = 0; ix < 16; ++ix)
        for (int iz = 0; iz < 16; ++iz) {
            const int col = (ix * 4) * coarsecore::SIZE + iz * 4;
            unsigned mask = 0;
            for (int y = 0; y < 17; ++y)
                if (s.signs[col * 17 + y]) mask |= 1u << y;
            std::snprintf(buf, sizeof(buf), "%s%u", ix || iz ? "," : "", mask);
            out += buf;
        }
    out += "], \"climate\": [";
    bool first = true;
    for (const auto& ch : kClimateChunks) {
        std::vector<double> temp, rain;
        for (int i = 0; i < 4; ++i)
            for (int l = 0; l < 4; ++l) {
                const int col = (ch[0] * 4 + i - coarsecore::FROM_COARSE) * coarsecore::SIZE +
                                (ch[1] * 4 + l - coarsecore::FROM_COARSE);
                temp.push_back(s.temp[col]);
                rain.push_back(s.rain[col]);
            }
        std::snprintf(buf, sizeof(buf), "%s{\"chunk\": [%d, %d], \"temp\": ", first ? "" : ", ", ch[0], ch[1]);
        out += buf + hexlist(temp) + ", \"rain\": " + hexlist(rain) + "}";
        first = false;
    }
    return out + "]}";
}

void coarse_fixture(const std::string& dir, int n_seeds) {
    std::vector<std::string> lattice_rows, stage0_rows;
    FILE* f = open_out(dir, "ref_coarse.json");
    std::fprintf(f, "{\"cases\": [\n");
    coarsecore::HostScratch s;
    std::vector<int64_t> seeds = {0, 1, 42, 1611, 6554};
    for (int i = 0; static_cast<int>(seeds.size()) < n_seeds; ++i)
        seeds.push_back(static_cast<int64_t>(p20::splitMixDeterministicSeed(0xB5F0C0A5EULL, i)));
    for (size_t i = 0; i < seeds.size(); ++i) {
        const int64_t seed = seeds[i];
        const int score = coarsecore::generateAndScoreHost(seed, s);
        auto dig = [](const std::vector<double>& v) {
            Fnv h;
            for (double d : v) h.f64(d);
            return h.h;
        };
        Fnv hs;
        int positives = 0;
        for (unsigned char c : s.signs) {
            hs.byte(c);
            positives += c;
        }
        std::fprintf(f,
                     "%s {\"seed\": %" PRId64 ", \"score\": %d, \"positives\": %d, \"signs\": \"%016" PRIx64
                     "\", \"noise1\": \"%016" PRIx64 "\", \"noise2\": \"%016" PRIx64 "\", \"noise3\": \"%016" PRIx64
                     "\", \"noise4\": \"%016" PRIx64 "\", \"noise5\": \"%016" PRIx64 "\", \"temp\": \"%016" PRIx64
                     "\", \"rain\": \"%016" PRIx64 "\"}\n",
                     i ? "," : "", seed, score, positives, hs.h, dig(s.noise1), dig(s.noise2), dig(s.noise3),
                     dig(s.noise4), dig(s.noise5), dig(s.temp), dig(s.rain));
        std::fflush(f);
        if (i < 4) lattice_rows.push_back(chunk00_lattice(seed, s));
        if (i < 10) stage0_rows.push_back(stage0_climate_row(seed, s));
    }
    std::fprintf(f, "]}\n");
    std::fclose(f);
    // the raw octave sums at chunk (0,0)'s 5x17x5 lattice (coarse x, z in 0..4) for the first seeds
    FILE* g = open_out(dir, "ref_chunk00_noise.json");
    std::fprintf(g, "{\"cases\": [\n");
    for (size_t i = 0; i < lattice_rows.size(); ++i) std::fprintf(g, "%s%s\n", i ? "," : "", lattice_rows[i].c_str());
    std::fprintf(g, "]}\n");
    std::fclose(g);
    FILE* h = open_out(dir, "ref_stage0_climate.json");
    std::fprintf(h, "{\"cases\": [\n");
    for (size_t i = 0; i < stage0_rows.size(); ++i) std::fprintf(h, "%s%s\n", i ? "," : "", stage0_rows[i].c_str());
    std::fprintf(h, "]}\n");
    std::fclose(h);
}

// P19 monster gate (p19_monster_gate.hpp:48-79) on random and structured feature vectors.
void p19_fixture(const std::string& dir) {
    FILE* f = open_out(dir, "ref_p19.json");
    std::mt19937_64 rng(7);
    std::fprintf(f, "{\"cases\": [\n");
    for (int i = 0; i < 2500; ++i) {
        p19native::MonsterFeatures p{};
        auto u = [&](int lo, int hi) { return static_cast<int>(lo + rng() % static_cast<uint64_t>(hi - lo + 1)); };
        const int scale = i % 3 == 0 ? 8 : (i % 3 == 1 ? 40 : 256);
        p.stage0FullY88 = u(0, scale);
        p.stage0FullY96 = u(0, p.stage0FullY88);
        p.stage0FullY104 = u(0, p.stage0FullY96);
        p.stage0FullY112 = u(0, p.stage0FullY104);
        p.stage0Y88LargestCluster = u(0, p.stage0FullY88);
        p.stage0Y96LargestCluster = u(0, p.stage0FullY96);
        p.stage0Y88Width = u(0, 16);
        p.stage0Y88Depth = u(0, 16);
        p.stage0Y96Width = u(0, 16);
        p.stage0Y96Depth = u(0, 16);
        p.stage0Y88TouchesBorder = u(0, 1);
        p.stage0Y96TouchesBorder = u(0, 1);
        const int scout = u(0, 256);
        const double score = p19native::score(scout, p);
        std::fprintf(f,
                     "%s [%d, %d, %d, %d, %d, %d, %d, %d, %d, %d, %d, %d, %d, \"%016" PRIx64 "\", %d, %d]\n",
                     i ? "," : "", scout, p.stage0FullY88, p.stage0FullY96, p.stage0FullY104, p.stage0FullY112,
                     p.stage0Y88LargestCluster, p.stage0Y88Width, p.stage0Y88Depth, p.stage0Y88TouchesBorder ? 1 : 0,
                     p.stage0Y96LargestCluster, p.stage0Y96Width, p.stage0Y96Depth,
                     p.stage0Y96TouchesBorder ? 1 : 0, bits(score), p19native::passes(scout, p) ? 1 : 0,
                     p19native::hasExtremeTopologySignal(p) ? 1 : 0);
    }
    std::fprintf(f, "]}\n");
    std::fclose(f);
}

}  // namespace

// SplitMix64 deterministic seeds (splitMixDeterministicSeed): sequence seeds x attempts.
void splitmix_fixture(const std::string& dir) {
    FILE* f = open_out(dir, "ref_splitmix.json");
    const uint64_t seqs[] = {0ULL, 1ULL, 0xB5F00002ULL, 0x9E3779B97F4A7C15ULL, 0x8000000000000000ULL,
                             0xFFFFFFFFFFFFFFFFULL, 123456789ULL};
    const uint64_t attempts[] = {0ULL, 1ULL, 2ULL, 3ULL, 1000ULL, 1ULL << 32, (1ULL << 40) + 7, 0xFFFFFFFFFFFFFFFFULL};
    std::fprintf(f, "{\"cases\": [\n");
    bool first = true;
    for (uint64_t q : seqs)
        for (uint64_t a : attempts) {
            std::fprintf(f, "%s {\"sequence_seed\": \"%016" PRIx64 "\", \"attempt\": \"%016" PRIx64
                            "\", \"seed\": \"%016" PRIx64 "\"}",
                         first ? "" : ",\n", q, a, p20::splitMixDeterministicSeed(q, a));
            first = false;
        }
    std::fprintf(f, "\n]}\n");
    std::fclose(f);
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int n_coarse = argc > 2 ? std::atoi(argv[2]) : 24;
    random_fixture(dir);
    perlin_fixture(dir);
    p19_fixture(dir);
    coarse_fixture(dir, n_coarse);
    splitmix_fixture(dir);
    std::printf("fixtures written to %s\n", dir.c_str());
    return 0;
}
