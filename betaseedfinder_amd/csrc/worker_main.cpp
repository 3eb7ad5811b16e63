// worker_main.cpp -- BetaSeedFinderWorker: the drop-in search worker process for MI355X.
//
// The Java launcher (src/GpuStage0Scout.java:143-219, src/GpuP20Scout.java:45-93) spawns the
// worker with positional argv, reads an 8-byte magic from stdout, then alternates
//   request  : LE u32 count, count x LE i64 seeds          (count == 0 or EOF: orderly exit)
//   response : LE u32 count, count x centers x record bytes
// while it pumps stderr lines into its log.  This program keeps those arguments, magics,
// record layouts and stderr lines (gpu_p20_benchmark.cpp:7251-7685, 7689-7734) and runs every
// batch through libbsf (include/bsf.h).  Errors end the process with "ERROR: <what>" on stderr
// and exit status 1, as the reference's main does (:7922-7925).
//
// Modes:
//   stream [capacity=32768]
//   stage0stream [capacity=32768] [general|mega|record60|record80] [research|lean] [centers=1|8]
//   density [capacity=32768] [chunkX=0] [chunkZ=0]      magic DNSTY425: count x 425 LE f64
//   island <first_seed> <count> [chunkX0 chunkZ0 chunksX chunksZ minBlocks]   hit lines (text)
//   region [capacity=4096] [radius=7] [minBlocks minColumns minY minWidth minDepth minAvgThickness]
//          magic REGION01; request LE u32 count, count x (i64 seed, i32 centre chunk x, i32 centre
//          chunk z, i32 coarse score); response LE u32 count, count x 56-byte results
//   hunt <output_root> <regions> [general|mega|record60|record80] [research|lean] [sequence_seed=0]
//        [label] [topResultsToKeep=50] [hunterCoarseMinCells=85]
//          a headless hunt (bsf_hunt): deterministic world seeds, stage0stream, region attempts,
//          and the Java search's run folder under <output_root>/runs; summary lines on stderr
#include <cinttypes>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "bsf.h"

namespace {

struct WorkerError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(int rc, const char* what) {
    if (rc != BSF_OK) throw WorkerError(std::string(what) + ": " + bsf_last_error());
}

// tryReadLe32 (gpu_p20_benchmark.cpp:1010-1020): clean EOF before the header -> false.
bool try_read_le32(uint32_t& v) {
    unsigned char b[4];
    const size_t got = std::fread(b, 1, 4, stdin);
    if (got == 0 && std::feof(stdin)) return false;
    if (got != 4) throw WorkerError("truncated stream request header");
    v = static_cast<uint32_t>(b[0]) | (static_cast<uint32_t>(b[1]) << 8) | (static_cast<uint32_t>(b[2]) << 16) |
        (static_cast<uint32_t>(b[3]) << 24);
    return true;
}

void read_seeds(uint32_t count, std::vector<int64_t>& seeds, std::vector<unsigned char>& raw, const char* what) {
    raw.resize(static_cast<size_t>(count) * 8);
    if (count && std::fread(raw.data(), 1, raw.size(), stdin) != raw.size()) throw WorkerError(what);
    for (uint32_t i = 0; i < count; ++i) {
        uint64_t v = 0;
        for (int b = 7; b >= 0; --b) v = (v << 8) | raw[static_cast<size_t>(i) * 8 + b];
        seeds[i] = static_cast<int64_t>(v);
    }
}

bool write_all(const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, stdout) == n; }

bool write_le32(uint32_t v) {
    const unsigned char b[4] = {static_cast<unsigned char>(v), static_cast<unsigned char>(v >> 8),
                                static_cast<unsigned char>(v >> 16), static_cast<unsigned char>(v >> 24)};
    return write_all(b, 4);
}

void print_device() {
    char banner[1024];
    check(bsf_device_banner(banner, sizeof(banner)), "device banner");
    std::fputs(banner, stderr);
}

int parse_int(const char* s, const char* what) {
    char* end = nullptr;
    const long v = std::strtol(s, &end, 10);
    if (!s[0] || *end) throw WorkerError(std::string("invalid ") + what + ": " + s);
    return static_cast<int>(v);
}

// streamFilter (gpu_p20_benchmark.cpp:7614-7685): one byte per seed, the P20 count clamped.
int stream_mode(int capacity) {
    if (capacity < 1) throw WorkerError("stream capacity must be >= 1");
    print_device();
    std::fprintf(stderr, "P20 stream capacity: %d seeds\n", capacity);
    std::vector<int64_t> seeds(capacity);
    std::vector<unsigned char> raw;
    std::vector<int32_t> counts(capacity);
    std::vector<unsigned char> out(capacity);
    if (!write_all("P20STR01", 8) || std::fflush(stdout) != 0) throw WorkerError("failed writing stream magic");
    uint64_t total = 0, passed = 0;
    for (;;) {
        uint32_t count = 0;
        if (!try_read_le32(count) || count == 0) break;
        if (count > static_cast<uint32_t>(capacity)) throw WorkerError("stream batch exceeds configured capacity");
        read_seeds(count, seeds, raw, "truncated stream seed batch");
        check(bsf_p20_counts(seeds.data(), static_cast<int32_t>(count), counts.data()), "P20 batch");
        uint64_t p = 0;
        for (uint32_t i = 0; i < count; ++i) {
            const int v = counts[i];
            out[i] = static_cast<unsigned char>(v < 0 ? 0 : (v > 255 ? 255 : v));
            if (v > 0) ++p;
        }
        if (!write_le32(count) || !write_all(out.data(), count) || std::fflush(stdout) != 0)
            throw WorkerError("failed writing stream response");
        total += count;
        passed += p;
    }
    std::fprintf(stderr, "P20 stream closed | seeds=%" PRIu64 " | passed=%" PRIu64 "\n", total, passed);
    return 0;
}

// The profiles' argv names and display names (HuntProfileConfig, gpu_p20_benchmark.cpp:7270-7290);
// their thresholds come from the library (bsf_stage0_profile).
struct ProfileInfo {
    const char* arg;
    const char* display;
};
const ProfileInfo kProfiles[4] = {
    {"general", "GENERAL"},
    {"mega", "MEGA 30k+"},
    {"record60", "RECORD 60k+ (empirical)"},
    {"record80", "WORLD RECORD COARSE 700 (compact-thick safe)"},
};

// streamStage0Filter (gpu_p20_benchmark.cpp:7310-7612).
int stage0_mode(int capacity, const std::string& profile, const std::string& telemetry, int centers) {
    int pid = -1;
    for (int i = 0; i < 4; ++i)
        if (profile == kProfiles[i].arg) pid = i;
    if (pid < 0) throw WorkerError("hunt profile must be general, mega, record60, or record80");
    if (telemetry != "research" && telemetry != "lean") throw WorkerError("stage0stream telemetry must be research or lean");
    if (capacity < 1) throw WorkerError("Stage0 stream capacity must be >= 1");
    if (centers != 1 && centers != 8) throw WorkerError("Stage0 stream center count must be 1 or 8");
    const bool research = telemetry == "research";
    int p20 = 0, upper = 0, high = 0, coarse = 0, mega = 0;
    double p19 = 0.0;
    check(bsf_stage0_profile(pid, &p20, &upper, &high, &p19, &coarse, &mega), "profile");
    print_device();
    std::fprintf(stderr, "Stage0+P19+coarse stream capacity: %d worlds / %lld regions\n", capacity,
                 static_cast<long long>(capacity) * centers);
    std::fprintf(stderr,
                 "P38 production multi-center: %d exact radius-7 regions/world | 15-chunk spacing | region budget "
                 "semantics\n",
                 centers);
    std::fprintf(stderr, "P38 profile: %s | P20>=%d | Upper>=%d | High>=%d | P19>=%g/extreme | Coarse>=%d%s\n",
                 kProfiles[pid].display, p20, upper, high, p19, coarse,
                 mega ? " | Mega topology reject ON" : " | topology reject OFF");
    if (pid >= 2) std::fprintf(stderr, "WARNING: record profile is empirically validated, not mathematically lossless.\n");
    std::fprintf(stderr, "Stage0+P19+coarse pipeline: exact P20 -> full upper survivors -> lower reentry -> exact "
                         "cached Stage0.75 gate -> optional Mega topology reject -> exact radius-7 coarse score\n");
    std::fprintf(stderr, "Response telemetry: %s\n", research ? "RESEARCH 34 bytes/region" : "LEAN 16 bytes/region + exact P19 double");
    const size_t rb = research ? 34 : 16;
    std::vector<int64_t> seeds(capacity);
    std::vector<unsigned char> raw;
    std::vector<uint8_t> out(static_cast<size_t>(capacity) * centers * rb);
    if (!write_all("ST0R3816", 8) || std::fflush(stdout) != 0) throw WorkerError("failed writing stream magic");
    bsf_stage0_counters c{};
    for (;;) {
        uint32_t count = 0;
        if (!try_read_le32(count) || count == 0) break;
        if (count > static_cast<uint32_t>(capacity))
            throw WorkerError("Stage0 stream batch exceeds configured world capacity");
        read_seeds(count, seeds, raw, "truncated Stage0 stream seed batch");
        check(bsf_stage0_records(seeds.data(), static_cast<int32_t>(count), pid, research ? 1 : 0, centers, out.data(), &c),
              "Stage0 batch");
        const size_t n = static_cast<size_t>(count) * centers * rb;
        if (!write_le32(count) || !write_all(out.data(), n) || std::fflush(stdout) != 0) {
            std::fprintf(stderr, "Stage0+P19+coarse stream parent closed response pipe; exiting cleanly\n");
            break;
        }
    }
    std::fprintf(stderr,
                 "Stage0+P19+coarse stream closed | worlds=%" PRIu64 " | regions=%" PRIu64 " | P20Passed=%" PRIu64
                 " | Stage1Passed=%" PRIu64 " | Stage0.5Passed=%" PRIu64 " | Stage0.75Evaluated=%" PRIu64
                 " | Stage0.75Passed=%" PRIu64 " | Stage0.75Rejected=%" PRIu64 " | MegaTopologyRejected=%" PRIu64
                 " | CoarseEvaluated=%" PRIu64 " | CoarsePassedTarget=%" PRIu64 "\n",
                 c.worlds, c.regions, c.p20_passed, c.stage1_passed, c.stage05_passed, c.stage075_evaluated,
                 c.stage075_passed, c.stage075_rejected, c.mega_topology_rejected, c.coarse_evaluated, c.coarse_passed);
    return 0;
}

// density: the 5x17x5 generateTerrainNoise lattice of one chunk for every seed of a batch.
int density_mode(int capacity, int cx, int cz) {
    if (capacity < 1) throw WorkerError("density capacity must be >= 1");
    print_device();
    std::fprintf(stderr, "Density stream capacity: %d seeds | chunk (%d, %d) | 425 f64 per seed\n", capacity, cx, cz);
    std::vector<int64_t> seeds(capacity);
    std::vector<unsigned char> raw;
    std::vector<double> out(static_cast<size_t>(capacity) * 425);
    if (!write_all("DNSTY425", 8) || std::fflush(stdout) != 0) throw WorkerError("failed writing stream magic");
    uint64_t total = 0;
    for (;;) {
        uint32_t count = 0;
        if (!try_read_le32(count) || count == 0) break;
        if (count > static_cast<uint32_t>(capacity)) throw WorkerError("density batch exceeds configured capacity");
        read_seeds(count, seeds, raw, "truncated density seed batch");
        check(bsf_density_grid(seeds.data(), count, cx, cz, out.data()), "density batch");
        // doubles go out little-endian: this host is little-endian (x86-64)
        if (!write_le32(count) || !write_all(out.data(), static_cast<size_t>(count) * 425 * 8) || std::fflush(stdout) != 0)
            throw WorkerError("failed writing density response");
        total += count;
    }
    std::fprintf(stderr, "Density stream closed | seeds=%" PRIu64 "\n", total);
    return 0;
}

// island: the full floating-island pipeline over a seed range, one text line per hit.
int island_mode(int argc, char** argv) {
    if (argc < 4) throw WorkerError("island requires <first_seed> <count>");
    const int64_t first = std::strtoll(argv[2], nullptr, 10);
    const int64_t count = std::strtoll(argv[3], nullptr, 10);
    bsf_island_params p{-2, -2, 4, 4, 1000, 100, 60, 10, 10, 2.0};
    if (argc >= 9) {
        p.chunk_x0 = parse_int(argv[4], "chunkX0");
        p.chunk_z0 = parse_int(argv[5], "chunkZ0");
        p.chunks_x = parse_int(argv[6], "chunksX");
        p.chunks_z = parse_int(argv[7], "chunksZ");
        p.min_blocks = parse_int(argv[8], "minBlocks");
    }
    print_device();
    std::vector<bsf_island_hit> hits(1 << 20);
    int64_t n = 0;
    check(bsf_island_range(first, count, &p, hits.data(), static_cast<int64_t>(hits.size()), &n), "island range");
    std::printf("seed,chunk_x,chunk_z,min_x,max_x,min_z,max_z,min_y,max_y,blocks,columns,center_x,center_z,touches_side\n");
    for (int64_t i = 0; i < n && i < static_cast<int64_t>(hits.size()); ++i) {
        const bsf_island_hit& h = hits[i];
        std::printf("%" PRId64 ",%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d\n", h.seed, h.chunk_x, h.chunk_z, h.min_x,
                    h.max_x, h.min_z, h.max_z, h.min_y, h.max_y, h.blocks, h.columns, h.center_x, h.center_z,
                    h.flags & 1);
    }
    std::fprintf(stderr, "Island range closed | seeds=%" PRId64 " | hits=%" PRId64 "\n", count, n);
    return 0;
}

// region: region attempts of Stage-0 survivors (the exact scans that IslandSearchEngine.
// processGpuRegionAttempt runs on the CPU, :1878-1971).  Each 56-byte result: u8 found, u8
// radius used, u8 normal scan found, u8 hot scan ran, u8 hot scan won, u8 touches side, 2 zero
// bytes, then LE i32 chunk x, chunk z, min x, max x, min z, max z, min y, max y, blocks,
// columns, centre x, centre z of the match (zero when none).
int region_mode(int argc, char** argv) {
    const int capacity = argc > 2 ? parse_int(argv[2], "capacity") : 4096;
    const int radius = argc > 3 ? parse_int(argv[3], "radius") : 7;
    if (capacity <= 0) throw WorkerError("region capacity must be positive");
    if (radius < 0 || radius > 24) throw WorkerError("region radius must be 0..24");
    bsf_island_params s{0, 0, 1, 1, 1000, 100, 60, 10, 10, 2.0};  // SearchSettings defaults
    if (argc >= 10) {
        s.min_blocks = parse_int(argv[4], "minBlocks");
        s.min_columns = parse_int(argv[5], "minColumns");
        s.min_y_for_match = parse_int(argv[6], "minY");
        s.min_width = parse_int(argv[7], "minWidth");
        s.min_depth = parse_int(argv[8], "minDepth");
        s.min_avg_thickness = std::strtod(argv[9], nullptr);
    }
    print_device();
    std::fprintf(stderr, "Region stream ready | capacity=%d | radius=%d | minBlocks=%d\n", capacity, radius, s.min_blocks);
    if (!write_all("REGION01", 8) || std::fflush(stdout) != 0) throw WorkerError("failed writing region magic");
    std::vector<unsigned char> raw;
    std::vector<bsf_region_item> items(static_cast<size_t>(capacity));
    std::vector<bsf_region_result> res(static_cast<size_t>(capacity));
    std::vector<unsigned char> out(static_cast<size_t>(capacity) * 56);
    uint64_t total = 0, matches = 0, hot = 0;
    for (;;) {
        uint32_t count = 0;
        if (!try_read_le32(count) || count == 0) break;
        if (count > static_cast<uint32_t>(capacity)) throw WorkerError("region batch exceeds configured capacity");
        raw.resize(static_cast<size_t>(count) * 20);
        if (std::fread(raw.data(), 1, raw.size(), stdin) != raw.size()) throw WorkerError("truncated region batch");
        auto le = [&](size_t at, int bytes) {
            uint64_t v = 0;
            for (int b = bytes - 1; b >= 0; --b) v = (v << 8) | raw[at + b];
            return v;
        };
        for (uint32_t i = 0; i < count; ++i) {
            const size_t at = static_cast<size_t>(i) * 20;
            bsf_region_item& it = items[i];
            it.seed = static_cast<int64_t>(le(at, 8));
            it.center_chunk_x = static_cast<int32_t>(static_cast<uint32_t>(le(at + 8, 4)));
            it.center_chunk_z = static_cast<int32_t>(static_cast<uint32_t>(le(at + 12, 4)));
            it.coarse_score = static_cast<int32_t>(static_cast<uint32_t>(le(at + 16, 4)));
            it.pad = 0;
        }
        check(bsf_region_attempts(items.data(), count, &s, radius, res.data()), "region batch");
        std::memset(out.data(), 0, static_cast<size_t>(count) * 56);
        for (uint32_t i = 0; i < count; ++i) {
            unsigned char* r = out.data() + static_cast<size_t>(i) * 56;
            const bsf_region_result& x = res[i];
            r[0] = static_cast<unsigned char>(x.found);
            r[1] = static_cast<unsigned char>(x.radius_used);
            r[2] = static_cast<unsigned char>(x.normal_found);
            r[3] = static_cast<unsigned char>(x.hot_scan_ran);
            r[4] = static_cast<unsigned char>(x.hot_won);
            hot += static_cast<uint64_t>(x.hot_scan_ran);
            if (!x.found) continue;
            r[5] = static_cast<unsigned char>(x.hit.flags & 1);
            const int32_t v[12] = {x.hit.chunk_x, x.hit.chunk_z, x.hit.min_x, x.hit.max_x, x.hit.min_z, x.hit.max_z,
                                   x.hit.min_y, x.hit.max_y, x.hit.blocks, x.hit.columns, x.hit.center_x, x.hit.center_z};
            for (int k = 0; k < 12; ++k)
                for (int b = 0; b < 4; ++b)
                    r[8 + 4 * k + b] = static_cast<unsigned char>(static_cast<uint32_t>(v[k]) >> (8 * b));
            ++matches;
        }
        if (!write_le32(count) || !write_all(out.data(), static_cast<size_t>(count) * 56) || std::fflush(stdout) != 0)
            throw WorkerError("failed writing region response");
        total += count;
    }
    std::fprintf(stderr, "Region stream closed | regions=%" PRIu64 " | matches=%" PRIu64 " | hot_scans=%" PRIu64 "\n",
                 total, matches, hot);
    return 0;
}

// The Java GPU hunt without the Java: runGpuStage0Pipeline + processGpuRegionAttempt
// (IslandSearchEngine.java:1206-1480, 1878-1971) and the run output files (:4230-5361).
int hunt_mode(int argc, char** argv) {
    if (argc < 4) throw WorkerError("hunt needs <output_root> <regions>");
    bsf_run_settings s{};
    char* end = nullptr;
    s.regions = std::strtoll(argv[3], &end, 10);
    if (!argv[3][0] || *end || s.regions < 0) throw WorkerError(std::string("invalid regions: ") + argv[3]);
    const std::string profile = argc > 4 ? argv[4] : "mega";
    int pid = -1;
    for (int i = 0; i < 4; ++i)
        if (profile == kProfiles[i].arg) pid = i;
    if (pid < 0) throw WorkerError("hunt profile must be general, mega, record60, or record80");
    const std::string telemetry = argc > 5 ? argv[5] : "lean";
    if (telemetry != "research" && telemetry != "lean") throw WorkerError("hunt telemetry must be research or lean");
    s.profile = pid;
    s.research = telemetry == "research" ? 1 : 0;
    if (argc > 6) {
        s.sequence_seed = std::strtoull(argv[6], &end, 0);
        if (!argv[6][0] || *end) throw WorkerError(std::string("invalid sequence seed: ") + argv[6]);
    }
    const char* label = argc > 7 ? argv[7] : "";
    s.top_keep = argc > 8 ? parse_int(argv[8], "topResultsToKeep") : 50;
    s.coarse_min = argc > 9 ? parse_int(argv[9], "hunterCoarseMinCells") : 85;
    s.chunk_radius = 7;
    // SearchSettings defaults (SearchSettings.java:39-45)
    s.thresholds.min_blocks = 1000;
    s.thresholds.min_columns = 100;
    s.thresholds.min_y_for_match = 60;
    s.thresholds.min_width = 10;
    s.thresholds.min_depth = 10;
    s.thresholds.min_avg_thickness = 2.0;
    print_device();
    std::fprintf(stderr, "Headless hunt | profile=%s | regions=%" PRId64 " | 8 regions/world | sequence=%" PRIu64 "\n",
                 kProfiles[pid].arg, s.regions, s.sequence_seed);
    bsf_hunt_summary sum{};
    char dir[4096];
    check(bsf_hunt(argv[2], label, nullptr, &s, &sum, dir, sizeof(dir)), "hunt");
    std::fprintf(stderr,
                 "Hunt finished | worlds=%" PRId64 " | regions=%" PRId64 " | survivors=%" PRId64 " | matches=%" PRId64
                 " | duplicates=%" PRId64 " | topUpdates=%" PRId64 " | %.3f s | %.0f regions/s\n",
                 sum.worlds, sum.regions, sum.survivors, sum.matches, sum.duplicates, sum.top_updates, sum.seconds,
                 sum.seconds > 0 ? sum.regions / sum.seconds : 0.0);
    std::fprintf(stderr, "Run folder: %s\n", dir);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc < 2) {
            std::fprintf(stderr,
                         "Usage:\n"
                         "  BetaSeedFinderWorker stream [capacity=32768]\n"
                         "  BetaSeedFinderWorker stage0stream [capacity=32768] [general|mega|record60|record80] "
                         "[research|lean] [centers=1|8]\n"
                         "  BetaSeedFinderWorker density [capacity=32768] [chunkX=0] [chunkZ=0]\n"
                         "  BetaSeedFinderWorker island <first_seed> <count> [chunkX0 chunkZ0 chunksX chunksZ minBlocks]\n"
                         "  BetaSeedFinderWorker region [capacity=4096] [radius=7] [minBlocks minColumns minY minWidth "
                         "minDepth minAvgThickness]\n"
                         "  BetaSeedFinderWorker hunt <output_root> <regions> [general|mega|record60|record80] "
                         "[research|lean] [sequence_seed=0] [label] [topResultsToKeep=50] [hunterCoarseMinCells=85]\n");
            return 1;
        }
        const std::string mode = argv[1];
        const char* dev = std::getenv("BSF_DEVICE");
        check(bsf_init(dev ? std::atoi(dev) : 0), "init");
        std::setvbuf(stdout, nullptr, _IOFBF, 1 << 20);
        int rc = 1;
        if (mode == "stream") {
            rc = stream_mode(argc > 2 ? parse_int(argv[2], "capacity") : 32768);
        } else if (mode == "stage0stream") {
            rc = stage0_mode(argc > 2 ? parse_int(argv[2], "capacity") : 32768, argc > 3 ? argv[3] : "general",
                             argc > 4 ? argv[4] : "research", argc > 5 ? parse_int(argv[5], "centers") : 1);
        } else if (mode == "density") {
            rc = density_mode(argc > 2 ? parse_int(argv[2], "capacity") : 32768, argc > 3 ? parse_int(argv[3], "chunkX") : 0,
                              argc > 4 ? parse_int(argv[4], "chunkZ") : 0);
        } else if (mode == "island") {
            rc = island_mode(argc, argv);
        } else if (mode == "region") {
            rc = region_mode(argc, argv);
        } else if (mode == "hunt") {
            rc = hunt_mode(argc, argv);
        } else {
            throw WorkerError("unknown mode: " + mode);
        }
        std::fflush(stdout);
        bsf_shutdown();
        return rc;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        return 1;
    }
}
