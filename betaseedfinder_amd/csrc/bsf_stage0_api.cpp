// bsf_stage0_api.cpp -- host side of the `stream` / `stage0stream` worker modes.
//
// Per batch of worlds (streamStage0Filter, gpu_p20_benchmark.cpp:7411-7579): one generator-
// cache build per world, then for every (world, centre) item the staged gates
//   P20 (64 columns) -> full upper (192 more) -> lower re-entry -> P19 / mega (host) -> coarse,
// each stage launched only over the items its predecessor passed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/bsf.h"
#include "bsf_host.hpp"
#include "bsf_stage0.hpp"

namespace bsf {

#include "p19_trees.inc"

namespace {

// HuntProfileConfig table, gpu_p20_benchmark.cpp:7270-7290.
struct Profile {
    int p20_min, upper_min, high_min;
    double p19_min;
    int coarse_min;
    bool mega_topology;
};
const Profile kProfiles[4] = {
    {1, 5, 5, 4.02, 85, false},    // general
    {1, 8, 6, 4.02, 85, true},     // mega
    {3, 19, 12, 6.70, 95, false},  // record60
    {3, 19, 20, 6.70, 700, false}, // record80
};
// Production centres, chunk offsets (gpu_p20_benchmark.cpp:7316-7317).
const int kCenterX[8] = {0, 15, 0, 15, -15, -15, 0, 15};
const int kCenterZ[8] = {0, 0, 15, 15, 0, 15, -15, -15};

struct Features {
    int full_y88 = 0, full_y96 = 0, full_y104 = 0, full_y112 = 0;
    int y88_largest = 0, y88_width = 0, y88_depth = 0;
    bool y88_touches = false;
    int y96_largest = 0, y96_width = 0, y96_depth = 0;
    bool y96_touches = false;
};

// fillMonsterShapeStats (gpu_p20_benchmark.cpp:1022-1088): 4-neighbour clusters of a 16x16
// grid; width/depth span every positive cell (not only the largest cluster).
void shape_stats(const uint8_t* grid, int& largest, int& width, int& depth, bool& touches) {
    uint8_t visited[256] = {0};
    int queue[256];
    int big = 0, minx = 16, maxx = -1, minz = 16, maxz = -1;
    touches = false;
    for (int x = 0; x < 16; ++x)
        for (int z = 0; z < 16; ++z) {
            const int idx = x * 16 + z;
            if (!grid[idx]) continue;
            minx = std::min(minx, x);
            maxx = std::max(maxx, x);
            minz = std::min(minz, z);
            maxz = std::max(maxz, z);
            if (x == 0 || z == 0 || x == 15 || z == 15) touches = true;
            if (visited[idx]) continue;
            int head = 0, tail = 0, size = 0;
            visited[idx] = 1;
            queue[tail++] = idx;
            while (head < tail) {
                const int cur = queue[head++];
                ++size;
                const int cx = cur / 16, cz = cur % 16;
                const int nx[4] = {cx - 1, cx + 1, cx, cx};
                const int nz[4] = {cz, cz, cz - 1, cz + 1};
                for (int k = 0; k < 4; ++k) {
                    if (nx[k] < 0 || nz[k] < 0 || nx[k] > 15 || nz[k] > 15) continue;
                    const int nb = nx[k] * 16 + nz[k];
                    if (!grid[nb] || visited[nb]) continue;
                    visited[nb] = 1;
                    queue[tail++] = nb;
                }
            }
            big = std::max(big, size);
        }
    largest = big;
    width = big == 0 ? 0 : maxx - minx + 1;
    depth = big == 0 ? 0 : maxz - minz + 1;
}

// buildMonsterFeatures, gpu_p20_benchmark.cpp:1090-1114.
Features monster_features(const uint8_t* highest) {
    Features f;
    uint8_t y88[256] = {0}, y96[256] = {0};
    for (int c = 0; c < 256; ++c) {
        if (highest[c] == 0xFF) continue;
        const int h = highest[c];
        if (h >= 11) { ++f.full_y88; y88[c] = 1; }
        if (h >= 12) { ++f.full_y96; y96[c] = 1; }
        if (h >= 13) ++f.full_y104;
        if (h >= 14) ++f.full_y112;
    }
    shape_stats(y88, f.y88_largest, f.y88_width, f.y88_depth, f.y88_touches);
    shape_stats(y96, f.y96_largest, f.y96_width, f.y96_depth, f.y96_touches);
    return f;
}

// Feature expressions of the P19 trees, exactly as the reference writes them (int sums and
// products, ratios against (count + 1e-6)), all compared as doubles.
double p19_feature(int kind, int scout, const Features& p) {
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

// P19MonsterGate.score (p19_monster_gate.hpp:48-67): (tree0 + ... + tree15) / 16.
double p19_score(int scout, const Features& p) {
    double sum = 0.0;
    for (int t = 0; t < 16; ++t) {
        int node = kP19TreeRoot[t];
        while (kP19Nodes[node].feature >= 0)
            node = p19_feature(kP19Nodes[node].feature, scout, p) <= kP19Nodes[node].value ? kP19Nodes[node].left
                                                                                           : kP19Nodes[node].right;
        sum = sum + kP19Nodes[node].value;
    }
    return sum / 16;
}

bool p19_extreme(const Features& p) { return p.y88_largest >= 14 || p.y96_largest >= 14; }  // :77-79

// megaTopologyReject, gpu_p20_benchmark.cpp:1116-1128.
bool mega_reject(double score, bool extreme, const Features& f) {
    if (extreme) return false;
    if (!(score < 5.0)) return false;
    return f.full_y112 == 0 || f.y96_largest < 4;
}

template <class K>
int grid_for(K kernel, int block, int64_t items) {
    int per_cu = 0;
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0), "occupancy");
    const int64_t res = static_cast<int64_t>(std::max(1, per_cu)) * g_ctx.cu_count;
    return static_cast<int>(std::max<int64_t>(1, std::min(items, res)));
}

struct Stage0Batch {
    int count = 0, centers = 1;
    int64_t items = 0;
    std::vector<int> p20, upper, high, coarse;
    std::vector<uint8_t> p19_pass, extreme, mega;
    std::vector<double> score;
    std::vector<Features> feats;
};

void setup_args(S0Args& a, int centers, uint32_t* perms, double* offs) {
    std::memset(&a, 0, sizeof(a));
    a.perms = perms;
    a.offsets = offs;
    a.centers = centers;
    for (int c = 0; c < 8; ++c) {
        a.center_dx[c] = kCenterX[c];
        a.center_dz[c] = kCenterZ[c];
    }
}

// Generator cache + P20 gate for count worlds x centers.
void run_p20(const int64_t* h_seeds, int count, int centers, S0Args& a, std::vector<int>& p20) {
    hipStream_t s = g_ctx.stream;
    const int64_t items = static_cast<int64_t>(count) * centers;
    int64_t* d_seeds = static_cast<int64_t*>(g_ctx.seeds.ensure(static_cast<size_t>(count) * 8));
    check(hipMemcpyAsync(d_seeds, h_seeds, static_cast<size_t>(count) * 8, hipMemcpyHostToDevice, s), "H2D seeds");
    uint32_t* perms = static_cast<uint32_t*>(g_ctx.s0_a.ensure(static_cast<size_t>(count) * kGenCount * 256));
    double* offs = static_cast<double*>(g_ctx.s0_b.ensure(static_cast<size_t>(count) * kGenCount * 3 * 8));
    GenCacheArgs g{d_seeds, count, perms, offs, 0};
    {
        LaunchTimer t("stage0_cache", s);
        hipLaunchKernelGGL(gen_cache_kernel, dim3(grid_for(gen_cache_kernel, 64, (count + 63) / 64)), dim3(64), 0, s, g);
        check(hipGetLastError(), "gen_cache launch");
        t.stop(static_cast<uint64_t>(count));
    }
    setup_args(a, centers, perms, offs);
    a.n_items = items;
    a.upper_mask = static_cast<uint8_t*>(g_ctx.s0_c.ensure(static_cast<size_t>(items) * 256));
    a.d5 = static_cast<double*>(g_ctx.s0_d.ensure(static_cast<size_t>(items) * 256 * 8));
    a.d7 = static_cast<double*>(g_ctx.s0_e.ensure(static_cast<size_t>(items) * 256 * 8));
    int* counts = static_cast<int*>(g_ctx.s0_f.ensure(static_cast<size_t>(items) * 4 * 4));
    a.p20 = counts;
    a.upper = counts + items;
    a.high = counts + 2 * items;
    a.p20_gated = counts + 3 * items;
    {
        LaunchTimer t("stage0_p20", s);
        auto k = s0_upper_kernel<S0_P20>;
        hipLaunchKernelGGL(k, dim3(grid_for(k, 64, items)), dim3(64), 0, s, a);
        check(hipGetLastError(), "p20 launch");
        t.stop(static_cast<uint64_t>(items));
    }
    p20.resize(items);
    check(hipMemcpyAsync(p20.data(), a.p20, static_cast<size_t>(items) * 4, hipMemcpyDeviceToHost, s), "D2H p20");
    check(hipStreamSynchronize(s), "p20 sync");
}

void stage0_batch(const int64_t* h_seeds, int count, int profile_id, int centers, Stage0Batch& b) {
    const Profile& prof = kProfiles[profile_id];
    hipStream_t s = g_ctx.stream;
    S0Args a;
    b.count = count;
    b.centers = centers;
    b.items = static_cast<int64_t>(count) * centers;
    const int64_t items = b.items;
    run_p20(h_seeds, count, centers, a, b.p20);
    b.upper.assign(items, 0);
    b.high.assign(items, 0);
    b.coarse.assign(items, 0);
    b.p19_pass.assign(items, 0);
    b.extreme.assign(items, 0);
    b.mega.assign(items, 0);
    b.score.assign(items, std::numeric_limits<double>::quiet_NaN());
    b.feats.assign(items, Features{});

    // applyP20MinimumGateKernel + stage0UpperCompact192SharedYKernel: items with gated P20 > 0
    std::vector<int> gated(items);
    std::vector<int64_t> list1;
    for (int64_t i = 0; i < items; ++i) {
        gated[i] = b.p20[i] >= prof.p20_min ? b.p20[i] : 0;
        if (gated[i] > 0) list1.push_back(i);
    }
    int64_t* d_list = static_cast<int64_t*>(g_ctx.s0_g.ensure(static_cast<size_t>(std::max<int64_t>(1, items)) * 8));
    if (!list1.empty()) {
        check(hipMemcpyAsync(const_cast<int*>(a.p20_gated), gated.data(), static_cast<size_t>(items) * 4,
                             hipMemcpyHostToDevice, s), "H2D gated");
        check(hipMemcpyAsync(d_list, list1.data(), list1.size() * 8, hipMemcpyHostToDevice, s), "H2D list");
        a.items = d_list;
        a.n_items = static_cast<int64_t>(list1.size());
        LaunchTimer t("stage0_upper", s);
        auto k = s0_upper_kernel<S0_UPPER>;
        hipLaunchKernelGGL(k, dim3(grid_for(k, 192, a.n_items)), dim3(192), 0, s, a);
        check(hipGetLastError(), "upper launch");
        t.stop(static_cast<uint64_t>(a.n_items));
        std::vector<int> up(items);
        check(hipMemcpyAsync(up.data(), a.upper, static_cast<size_t>(items) * 4, hipMemcpyDeviceToHost, s), "D2H upper");
        check(hipStreamSynchronize(s), "upper sync");
        for (int64_t i : list1) b.upper[i] = up[i];
    }
    // stage0LowerCompactKernel: items with upper >= upperMin
    std::vector<int64_t> list2;
    for (int64_t i : list1)
        if (b.upper[i] >= prof.upper_min) list2.push_back(i);
    std::vector<uint8_t> highest;
    if (!list2.empty()) {
        check(hipMemcpyAsync(d_list, list2.data(), list2.size() * 8, hipMemcpyHostToDevice, s), "H2D list");
        a.items = d_list;
        a.n_items = static_cast<int64_t>(list2.size());
        a.highest = static_cast<uint8_t*>(g_ctx.s0_h.ensure(list2.size() * 256));
        LaunchTimer t("stage0_lower", s);
        hipLaunchKernelGGL(s0_lower_kernel, dim3(grid_for(s0_lower_kernel, 256, a.n_items)), dim3(256), 0, s, a);
        check(hipGetLastError(), "lower launch");
        t.stop(static_cast<uint64_t>(a.n_items));
        std::vector<int> hi(items);
        highest.resize(list2.size() * 256);
        check(hipMemcpyAsync(hi.data(), a.high, static_cast<size_t>(items) * 4, hipMemcpyDeviceToHost, s), "D2H high");
        check(hipMemcpyAsync(highest.data(), a.highest, highest.size(), hipMemcpyDeviceToHost, s), "D2H highest");
        check(hipStreamSynchronize(s), "lower sync");
        for (int64_t i : list2) b.high[i] = hi[i];
    }
    // Stage0.75 P19 + Mega topology on the host (gpu_p20_benchmark.cpp:7473-7489)
    std::vector<int64_t> coarse_list;
    for (size_t k = 0; k < list2.size(); ++k) {
        const int64_t i = list2[k];
        if (!(b.p20[i] >= prof.p20_min && b.upper[i] >= prof.upper_min && b.high[i] >= prof.high_min)) continue;
        const Features f = monster_features(&highest[k * 256]);
        const double sc = p19_score(b.upper[i], f);
        const bool ext = p19_extreme(f);
        const bool pass = ext || sc >= prof.p19_min;
        const bool mr = pass && prof.mega_topology && mega_reject(sc, ext, f);
        b.feats[i] = f;
        b.score[i] = sc;
        b.extreme[i] = ext ? 1 : 0;
        b.p19_pass[i] = pass ? 1 : 0;
        b.mega[i] = mr ? 1 : 0;
        if (pass && !mr) coarse_list.push_back(i);
    }
    // exact radius-7 coarse score (P35 + P36 roles)
    if (!coarse_list.empty()) {
        const int64_t chunk = 4096;
        for (size_t off = 0; off < coarse_list.size(); off += chunk) {
            const int64_t m = std::min<int64_t>(chunk, static_cast<int64_t>(coarse_list.size() - off));
            check(hipMemcpyAsync(d_list, coarse_list.data() + off, static_cast<size_t>(m) * 8, hipMemcpyHostToDevice, s),
                  "H2D coarse list");
            a.items = d_list;
            a.n_items = m;
            a.coarse_bits = static_cast<unsigned long long*>(g_ctx.lattice.ensure(static_cast<size_t>(m) * kCoarseRows * 8));
            a.coarse_score = static_cast<int*>(g_ctx.small.ensure(static_cast<size_t>(m) * 4));
            LaunchTimer t("stage0_coarse", s);
            hipLaunchKernelGGL(s0_coarse_signs_kernel, dim3(grid_for(s0_coarse_signs_kernel, 256, m)), dim3(256), 0, s, a);
            check(hipGetLastError(), "coarse signs launch");
            hipLaunchKernelGGL(s0_coarse_score_kernel, dim3(grid_for(s0_coarse_score_kernel, 256, m)), dim3(256), 0, s, a);
            check(hipGetLastError(), "coarse score launch");
            t.stop(static_cast<uint64_t>(m));
            std::vector<int> sc(m);
            check(hipMemcpyAsync(sc.data(), a.coarse_score, static_cast<size_t>(m) * 4, hipMemcpyDeviceToHost, s),
                  "D2H coarse");
            check(hipStreamSynchronize(s), "coarse sync");
            for (int64_t j = 0; j < m; ++j) b.coarse[coarse_list[off + j]] = sc[j];
        }
    }
}

inline uint8_t clamp_u8(int v) { return static_cast<uint8_t>(v < 0 ? 0 : (v > 255 ? 255 : v)); }
inline void put_u16(uint8_t* p, int v) {
    const unsigned u = static_cast<unsigned>(v < 0 ? 0 : (v > 65535 ? 65535 : v));
    p[0] = static_cast<uint8_t>(u & 0xFF);
    p[1] = static_cast<uint8_t>(u >> 8);
}

}  // namespace
}  // namespace bsf

using namespace bsf;

extern "C" {

int bsf_p20_counts(const int64_t* seeds, int32_t count, int32_t* out) {
    return guarded_call([&] {
        require_init();
        if (count < 0 || (count > 0 && (!seeds || !out))) throw ArgError("bad p20 arguments");
        if (count == 0) return;
        S0Args a;
        std::vector<int> p20;
        run_p20(seeds, count, 1, a, p20);
        std::copy(p20.begin(), p20.end(), out);
    });
}

int bsf_stage0_records(const int64_t* seeds, int32_t count, int32_t profile, int32_t research, int32_t centers,
                       uint8_t* out, bsf_stage0_counters* counters) {
    return guarded_call([&] {
        require_init();
        if (profile < 0 || profile > 3) throw ArgError("hunt profile must be general, mega, record60, or record80");
        if (centers != 1 && centers != 8) throw ArgError("Stage0 stream center count must be 1 or 8");
        if (count < 0 || (count > 0 && (!seeds || !out))) throw ArgError("bad stage0 arguments");
        if (count == 0) return;
        const Profile& prof = kProfiles[profile];
        Stage0Batch b;
        stage0_batch(seeds, count, profile, centers, b);
        const size_t rb = research ? 34 : 16;
        bsf_stage0_counters c{};
        for (int64_t i = 0; i < b.items; ++i) {
            uint8_t* r = out + static_cast<size_t>(i) * rb;  // record index world * centers + center
            const Features& f = b.feats[i];
            r[0] = clamp_u8(b.p20[i]);
            r[1] = clamp_u8(b.upper[i]);
            r[2] = clamp_u8(b.high[i]);
            r[3] = b.p19_pass[i];
            r[4] = static_cast<uint8_t>(b.coarse[i] & 0xFF);
            r[5] = static_cast<uint8_t>((b.coarse[i] >> 8) & 0xFF);
            r[6] = b.extreme[i];
            r[7] = b.mega[i];
            uint64_t sb;
            std::memcpy(&sb, &b.score[i], 8);
            for (int k = 0; k < 8; ++k) r[8 + k] = static_cast<uint8_t>(sb >> (8 * k));
            if (research) {
                put_u16(r + 16, f.full_y88);
                put_u16(r + 18, f.full_y96);
                put_u16(r + 20, f.full_y104);
                put_u16(r + 22, f.full_y112);
                put_u16(r + 24, f.y88_largest);
                r[26] = static_cast<uint8_t>(f.y88_width);
                r[27] = static_cast<uint8_t>(f.y88_depth);
                r[28] = f.y88_touches ? 1 : 0;
                put_u16(r + 29, f.y96_largest);
                r[31] = static_cast<uint8_t>(f.y96_width);
                r[32] = static_cast<uint8_t>(f.y96_depth);
                r[33] = f.y96_touches ? 1 : 0;
            }
            // closing-summary counters, gpu_p20_benchmark.cpp:7555-7571
            const int pa = b.p20[i], pb = b.upper[i], pc = b.high[i], pd = b.coarse[i];
            if (pa >= prof.p20_min) ++c.p20_passed;
            if (pa >= prof.p20_min && pb >= prof.upper_min) ++c.stage1_passed;
            if (pa >= prof.p20_min && pb >= prof.upper_min && pc >= prof.high_min) {
                ++c.stage05_passed;
                ++c.stage075_evaluated;
                if (b.p19_pass[i]) {
                    ++c.stage075_passed;
                    if (b.mega[i]) {
                        ++c.mega_topology_rejected;
                    } else {
                        ++c.coarse_evaluated;
                        if (pd >= prof.coarse_min) ++c.coarse_passed;
                    }
                } else {
                    ++c.stage075_rejected;
                }
            }
        }
        c.worlds = static_cast<uint64_t>(count);
        c.regions = static_cast<uint64_t>(b.items);
        if (counters) {
            uint64_t* dst = reinterpret_cast<uint64_t*>(counters);
            const uint64_t* src = reinterpret_cast<const uint64_t*>(&c);
            for (size_t k = 0; k < sizeof(c) / 8; ++k) dst[k] += src[k];
        }
    });
}

}  // extern "C"
