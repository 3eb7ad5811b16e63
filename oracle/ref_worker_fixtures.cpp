// ref_worker_fixtures.cpp -- writes golden fixtures from the reference worker's own host code.
//
// The worker's gate helpers live in an unnamed namespace of its single source file
// (native/src/gpu_p20_benchmark.cpp of the reference), so this driver includes that file where it
// lies and is compiled host-only with hipcc by `make -C oracle ref` (its `main` renamed); the
// binary goes to oracle/_ref/ and runs only in the development container, with no GPU.  It calls
//
//   buildMonsterFeatures / fillMonsterShapeStats (gpu_p20_benchmark.cpp:1022-1114)
//   megaTopologyReject                           (:1116-1128)
//   parseHuntProfile / huntProfileConfig         (:7270-7300)
//
// and writes tests/golden/ref_worker.json: the four hunt profiles, and Stage0 re-entry boards
// (256 highest re-entry Y indices, 0xFF = none) with the reference's P19 features and its
// mega-topology decision for a ladder of scores.
//
//   ./_ref/ref_worker_fixtures <out_dir>
#define main bsf_reference_worker_main
#include "gpu_p20_benchmark.cpp"
#undef main

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <vector>

namespace fixtures {

uint64_t bits_of(double v) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    return b;
}

using Board = std::vector<unsigned char>;

// Random boards: each column is empty (0xFF) with probability p_empty, else a re-entry level
// drawn from 0..16 with the levels 11..14 (the feature thresholds) and their neighbours favoured.
Board random_board(std::mt19937_64& rng, double p_empty, int lo) {
    std::uniform_real_distribution<double> u(0.0, 1.0);
    Board b(256);
    for (int c = 0; c < 256; ++c) {
        if (u(rng) < p_empty) {
            b[c] = 0xFF;
            continue;
        }
        b[c] = static_cast<unsigned char>(lo + static_cast<int>(rng() % static_cast<uint64_t>(17 - lo)));
    }
    return b;
}

// Structured boards: edges, corners, rings, lines, clusters of equal size, a checkerboard.
std::vector<Board> structured_boards() {
    std::vector<Board> out;
    auto blank = [] { return Board(256, 0xFF); };
    auto set = [](Board& b, int x, int z, int v) { b[x * 16 + z] = static_cast<unsigned char>(v); };
    out.push_back(blank());
    {
        Board b(256, 16);
        out.push_back(b);
    }
    for (int v : {10, 11, 12, 13, 14, 16}) {
        Board b = blank();
        set(b, 7, 8, v);
        out.push_back(b);
    }
    for (auto xz : {std::pair<int, int>{0, 0}, {15, 15}, {0, 15}, {15, 0}, {0, 7}, {7, 15}}) {
        Board b = blank();
        set(b, xz.first, xz.second, 12);
        out.push_back(b);
    }
    {
        Board b = blank();  // a ring touching no border, with a hole
        for (int x = 3; x <= 12; ++x)
            for (int z = 3; z <= 12; ++z)
                if (x == 3 || x == 12 || z == 3 || z == 12) set(b, x, z, 13);
        out.push_back(b);
    }
    {
        Board b = blank();  // two clusters of equal size far apart: width/depth span both
        for (int x = 1; x <= 3; ++x)
            for (int z = 1; z <= 3; ++z) set(b, x, z, 11);
        for (int x = 10; x <= 12; ++x)
            for (int z = 11; z <= 13; ++z) set(b, x, z, 12);
        out.push_back(b);
    }
    {
        Board b = blank();  // checkerboard: 128 singleton clusters
        for (int x = 0; x < 16; ++x)
            for (int z = 0; z < 16; ++z)
                if ((x + z) % 2 == 0) set(b, x, z, 14);
        out.push_back(b);
    }
    {
        Board b = blank();  // a diagonal staircase, 4-connected
        for (int k = 0; k < 15; ++k) {
            set(b, k, k, 12);
            set(b, k + 1, k, 12);
        }
        out.push_back(b);
    }
    {
        Board b = blank();  // a full row and a full column of different levels
        for (int k = 0; k < 16; ++k) {
            set(b, 5, k, 11);
            set(b, k, 9, 12);
        }
        out.push_back(b);
    }
    {
        Board b = blank();  // a snake of 14 cells at Y96, large only at Y88
        for (int z = 2; z < 14; ++z) set(b, 8, z, 12);
        set(b, 7, 2, 11);
        set(b, 9, 13, 11);
        out.push_back(b);
    }
    return out;
}

const double kScores[] = {std::numeric_limits<double>::quiet_NaN(), -1.0, 0.0, 4.999999999999999, 5.0,
                          5.000000000000001, 6.7, 9.5};

void write(const std::string& dir) {
    const std::string path = dir + "/ref_worker.json";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::perror(path.c_str());
        std::exit(1);
    }
    std::fprintf(f, "{\"profiles\": [\n");
    const char* names[] = {"general", "mega", "record60", "record80"};
    for (int i = 0; i < 4; ++i) {
        const HuntProfileConfig& c = huntProfileConfig(parseHuntProfile(names[i]));
        std::fprintf(f,
                     "%s {\"name\": \"%s\", \"p20_min\": %d, \"upper_min\": %d, \"high_min\": %d, \"p19_min\": \"%016" PRIx64
                     "\", \"coarse_min\": %d, \"mega_filter\": %d}\n",
                     i ? "," : "", c.argument, c.p20MinCount, c.upperMinCount, c.highMinCount,
                     bits_of(c.p19MinScore), c.coarseMinCells, c.megaTopologyFilter ? 1 : 0);
    }
    std::fprintf(f, "],\n\"scores\": [");
    for (size_t s = 0; s < sizeof(kScores) / sizeof(kScores[0]); ++s)
        std::fprintf(f, "%s\"%016" PRIx64 "\"", s ? ", " : "", bits_of(kScores[s]));
    std::fprintf(f, "],\n\"boards\": [\n");
    std::vector<Board> boards = structured_boards();
    std::mt19937_64 rng(0xB5F0B0A4D);
    const double empties[] = {0.0, 0.3, 0.6, 0.8, 0.9, 0.95, 0.98};
    for (int k = 0; k < 280; ++k) boards.push_back(random_board(rng, empties[k % 7], k % 2 ? 9 : 0));
    for (size_t k = 0; k < boards.size(); ++k) {
        const p19native::MonsterFeatures m = buildMonsterFeatures(boards[k].data());
        // megaTopologyReject over scores x {no bypass, bypass}: bit 2*s + bypass
        uint32_t mega = 0;
        for (size_t s = 0; s < sizeof(kScores) / sizeof(kScores[0]); ++s)
            for (int bypass = 0; bypass < 2; ++bypass)
                if (megaTopologyReject(kScores[s], bypass != 0, m)) mega |= 1u << (2 * s + bypass);
        std::fprintf(f, "%s {\"highest\": \"", k ? "," : "");
        for (int c = 0; c < 256; ++c) std::fprintf(f, "%02x", boards[k][c]);
        std::fprintf(f, "\", \"features\": [%d, %d, %d, %d, %d, %d, %d, %d, %d, %d, %d, %d], \"mega\": %u}\n",
                     m.stage0FullY88, m.stage0FullY96, m.stage0FullY104, m.stage0FullY112, m.stage0Y88LargestCluster,
                     m.stage0Y88Width, m.stage0Y88Depth, m.stage0Y88TouchesBorder ? 1 : 0, m.stage0Y96LargestCluster,
                     m.stage0Y96Width, m.stage0Y96Depth, m.stage0Y96TouchesBorder ? 1 : 0, mega);
    }
    std::fprintf(f, "]}\n");
    std::fclose(f);
    std::printf("%zu boards written to %s\n", boards.size(), path.c_str());
}

}  // namespace fixtures

int main(int argc, char** argv) {
    fixtures::write(argc > 1 ? argv[1] : ".");
    return 0;
}
