// bsf_stage0.hip -- Stage0 worker kernels (stage0stream / stream modes) for gfx950.
//
// Parallel forms:
//  * gen_cache_kernel: one world per lane builds the world's 76 used generators (66 terrain,
//    10 climate) in Random order and writes the permutations and offsets to a per-world cache
//    (the role of terrain_perlin_cache.hpp / climate_perlin_cache.hpp in the reference).
//  * column evaluation: one workgroup per (world, centre) item, one lattice column per lane.
//    Per generator the 256-byte permutation is staged once in LDS and shared by all lanes,
//    so every octave accumulator stays in registers.
//  * the staged gates (P20 -> upper -> lower re-entry + P19 / mega -> coarse): one workgroup per
//    item; each stage appends the items that pass its gate to a device list that the next
//    stage walks, so a batch needs no host round trip until its records are assembled.
//  * s0_coarse_signs_kernel: one 64-lane workgroup per item builds the 61x61x17 coarse sign
//    grid in LDS (one uint64 z-mask per (x, y) row); s0_coarse_score_kernel scores it by
//    bit-parallel component search.
#include "bsf_stage0.hpp"
#include "bsf_cache.hpp"

#include <hip/hip_runtime.h>

namespace bsf {

__constant__ DivMagic s0_magic[257];

// P19 tree table (16 trees): feature -1 marks a leaf whose value is the tree output.
struct P19DevNode {
    double value;
    int feature, left, right;
};
constexpr int kP19MaxNodes = 1024;
__constant__ P19DevNode s0_p19_nodes[kP19MaxNodes];
__constant__ int s0_p19_root[16];

namespace host_tables {
#include "p19_trees.inc"
}

hipError_t upload_p19_tables() {
    using namespace host_tables;
    constexpr int n = static_cast<int>(sizeof(kP19Nodes) / sizeof(kP19Nodes[0]));
    static_assert(n <= kP19MaxNodes, "P19 table too large");
    P19DevNode nodes[n];
    for (int i = 0; i < n; ++i) nodes[i] = P19DevNode{kP19Nodes[i].value, kP19Nodes[i].feature, kP19Nodes[i].left,
                                                      kP19Nodes[i].right};
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(s0_p19_nodes), nodes, sizeof(nodes));
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(s0_p19_root), kP19TreeRoot, sizeof(kP19TreeRoot));
}

hipError_t upload_stage0_tables() {
    DivMagic t[257];
    fill_div_magic(t);
    const hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(s0_magic), t, sizeof(t));
    if (e != hipSuccess) return e;
    return upload_p19_tables();
}

// ---------------------------------------------------------------------------
// generator cache (bsf_cache.hpp)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void gen_cache_kernel(GenCacheArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t perm_lds[64 * 256];
    const int lane = threadIdx.x;
    LanePerm P;
    P.base = perm_lds + lane * 4;
    const int64_t groups = (a.count + 63) >> 6;
    for (int64_t grp = blockIdx.x; grp < groups; grp += gridDim.x) {
        const int64_t w = (grp << 6) + lane;
        const bool active = w < a.count;
        const int64_t idx = active ? w : a.count - 1;
        const int64_t seed = a.seeds != nullptr ? a.seeds[idx] : a.first_seed + idx;
        build_world_cache(seed, active, P, a.perms + static_cast<size_t>(active ? w : 0) * kGenCount * 64,
                          a.offsets + static_cast<size_t>(active ? w : 0) * kGenCount * 3, s0_magic);
    }
}

// ---------------------------------------------------------------------------
// column evaluation from the cache (bsf_colpar.hpp), coarse climate semantics
// ---------------------------------------------------------------------------
// All five octave groups and the climate of one coarse-semantics column: climate sampled at
// block coarse*4 + 2 (BetaTerrain173.java:2524-2584).
template <int Y0, int Y1>
struct S0Column {
    ColAcc<Y0, Y1> acc;
    double temp, rain;
};

template <int Y0, int Y1, bool WITH_SHAPE>
__device__ __forceinline__ void evaluate_cached_column(GenStage& st, const uint32_t* perms, const double* offs,
                                                       double X, double Z, bool active, S0Column<Y0, Y1>& c) {
    const OneWorldBits s{&st, perms, offs};
    colpar_noise<Y0, Y1, WITH_SHAPE>(s, X, Z, active, c.acc);
    if (WITH_SHAPE) {
        const double BX = X * 4.0 + 2.0, BZ = Z * 4.0 + 2.0;
        colpar_climate_all<1>(s, &BX, &BZ, active ? 1 : 0, &c.temp, &c.rain);
    }
}

__device__ __forceinline__ bool p20_axis(int i) {
    return i == 0 || i == 2 || i == 5 || i == 7 || i == 8 || i == 10 || i == 13 || i == 15;
}
// The 192 Stage0 axis positions that are not P20 positions, compact index -> full index.
__device__ __forceinline__ int nonp20_column(int compact) {
    int rem = compact;
    for (int ix = 0; ix < 16; ++ix) {
        const int cols = p20_axis(ix) ? 8 : 16;
        if (rem < cols) {
            if (!p20_axis(ix)) return ix * 16 + rem;
            int seen = 0;
            for (int iz = 0; iz < 16; ++iz)
                if (!p20_axis(iz)) {
                    if (seen == rem) return ix * 16 + iz;
                    ++seen;
                }
        }
        rem -= cols;
    }
    return 255;
}
__device__ __forceinline__ int p20_column(int lane) {
    const int ax[8] = {0, 2, 5, 7, 8, 10, 13, 15};
    return ax[lane >> 3] * 16 + ax[lane & 7];
}

// highestSparseReentryYIndex, BetaTerrain173.java:2486-2504 / stage0_exact_gpu.hpp:552-566.
__device__ __forceinline__ int highest_reentry(uint32_t positive) {
    bool saw_pos = false, saw_gap = false;
    int highest = -1;
    for (int y = 0; y < 17; ++y) {
        const bool pos = (positive >> y) & 1u;
        if (pos) {
            if (saw_pos && saw_gap) highest = y;
            saw_pos = true;
        } else if (saw_pos) {
            saw_gap = true;
        }
    }
    return highest;
}

// ---------------------------------------------------------------------------
// P19 monster gate (p19_monster_gate.hpp:48-79, gpu_p20_benchmark.cpp:1022-1128)
// ---------------------------------------------------------------------------
// Feature expressions of the P19 trees as the reference writes them (int sums and products,
// ratios against (count + 1e-6)), all compared as doubles.
__device__ double p19_feature(int kind, int scout, const P19Features& p) {
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

// Output of one tree: go left while feature <= threshold, until a leaf.
__device__ double p19_tree(int tree, int scout, const P19Features& p) {
    int node = s0_p19_root[tree];
    while (s0_p19_nodes[node].feature >= 0) {
        const P19DevNode& n = s0_p19_nodes[node];
        node = p19_feature(n.feature, scout, p) <= n.value ? n.left : n.right;
    }
    return s0_p19_nodes[node].value;
}

// fillMonsterShapeStats (gpu_p20_benchmark.cpp:1022-1088) on a 16x16 board held as four
// 64-bit words: word w holds rows x = 4w..4w+3, row x at bits 16 (x & 3) .. + 15, bit z.
// Width and depth span every set cell; largest is the biggest 4-neighbour cluster, grown by
// bit-parallel dilation.
__device__ void board_shape(const unsigned long long* b, int& largest, int& width, int& depth, int& touches) {
    constexpr unsigned long long kZ0 = 0x0001000100010001ULL, kZ15 = 0x8000800080008000ULL;
    unsigned long long rem[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) rem[i] = b[i];
    if ((rem[0] | rem[1] | rem[2] | rem[3]) == 0) {
        largest = width = depth = touches = 0;
        return;
    }
    int minx = 16, maxx = -1;
    unsigned zm = 0;
#pragma unroll
    for (int x = 0; x < 16; ++x) {
        const unsigned row = static_cast<unsigned>(rem[x >> 2] >> (16 * (x & 3))) & 0xFFFFu;
        if (row) {
            minx = minx < x ? minx : x;
            maxx = x;
            zm |= row;
        }
    }
    const int minz = __ffs(zm) - 1, maxz = 31 - __clz(zm);
    touches = (minx == 0 || maxx == 15 || minz == 0 || maxz == 15) ? 1 : 0;
    width = maxx - minx + 1;
    depth = maxz - minz + 1;
    int big = 0;
    for (;;) {
        const int k = rem[0] ? 0 : rem[1] ? 1 : rem[2] ? 2 : rem[3] ? 3 : 4;
        if (k == 4) break;
        unsigned long long c[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) c[i] = i == k ? rem[i] & (~rem[i] + 1ULL) : 0ULL;
        for (;;) {
            unsigned long long nb[4];
            bool same = true;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                unsigned long long h = c[i] | ((c[i] << 1) & ~kZ0) | ((c[i] >> 1) & ~kZ15) | (c[i] << 16) | (c[i] >> 16);
                if (i > 0) h |= c[i - 1] >> 48;
                if (i < 3) h |= c[i + 1] << 48;
                nb[i] = h & rem[i];
                same = same && nb[i] == c[i];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) c[i] = nb[i];
            if (same) break;
        }
        const int size = __popcll(c[0]) + __popcll(c[1]) + __popcll(c[2]) + __popcll(c[3]);
        big = big > size ? big : size;
#pragma unroll
        for (int i = 0; i < 4; ++i) rem[i] &= ~c[i];
    }
    largest = big;
}

__device__ __forceinline__ void append_item(int64_t* list, int* count, int64_t item) {
    const int at = atomicAdd(count, 1);
    list[at] = item;
}

// ---------------------------------------------------------------------------
// staged Stage0 gates
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t stage_items(const S0Args& a) {
    return a.n_dev != nullptr ? static_cast<int64_t>(*a.n_dev) : a.n_items;
}

// P20 / upper: Y 11..15 of the item's Stage0 columns (Y 16 is -10 after the taper).  P20 runs
// over every item and, with a list, resets the item's later-stage outputs and appends the
// items whose count reaches p20_min; upper runs over that list, adds the P20 count (the
// gated P20 of applyP20MinimumGateKernel) and appends the items reaching upper_min.
template <int MODE>
__global__ __launch_bounds__(256, 3) void s0_upper_kernel(S0Args a) {
    __shared__ GenStage st;
    __shared__ int s_count;
    const int tid = threadIdx.x;
    const int ncols = MODE == S0_P20 ? 64 : 192;
    const int64_t n = stage_items(a);
    for (int64_t it = blockIdx.x; it < n; it += gridDim.x) {
        const int64_t item = a.items != nullptr ? a.items[it] : it;
        const int64_t world = item / a.centers;
        const int center = static_cast<int>(item % a.centers);
        const uint32_t* perms = a.perms + static_cast<size_t>(world) * kGenCount * 64;
        const double* offs = a.offsets + static_cast<size_t>(world) * kGenCount * 3;
        const bool active = tid < ncols;
        const int col = active ? (MODE == S0_P20 ? p20_column(tid) : nonp20_column(tid)) : 0;
        const double X = -28.0 + static_cast<double>((col >> 4) * 4 + a.center_dx[center] * 4);
        const double Z = -28.0 + static_cast<double>((col & 15) * 4 + a.center_dz[center] * 4);
        if (tid == 0) s_count = 0;
        S0Column<11, 16> cc;
        evaluate_cached_column<11, 16, true>(st, perms, offs, X, Z, active, cc);
        const ColAcc<11, 16>& acc = cc.acc;
        if (active) {
            const ColumnShape cs = column_shape(cc.temp, cc.rain, acc.n4, acc.n5);
            uint8_t mask = 0;
#pragma unroll
            for (int k = 0; k < 5; ++k)
                if (lattice_density(cs, 11 + k, acc.n1[k], acc.n2[k], acc.n3[k]) > 0.0) mask |= static_cast<uint8_t>(1u << k);
            const size_t ci = static_cast<size_t>(item) * 256 + col;
            a.upper_mask[ci] = mask;
            a.d5[ci] = cs.d5;
            a.d7[ci] = cs.d7;
            if (mask) atomicAdd(&s_count, 1);
        }
        __syncthreads();
        if (tid == 0) {
            if (MODE == S0_P20) {
                a.p20[item] = s_count;
                if (a.flags != nullptr) {
                    a.upper[item] = 0;
                    a.high[item] = 0;
                    a.coarse[item] = 0;
                    a.score[item] = __builtin_nan("");
                    a.flags[item] = 0;
                    a.feats[item] = P19Features{};
                }
                if (a.next_items != nullptr && s_count >= a.p20_min) append_item(a.next_items, a.next_count, item);
            } else {
                const int upper = a.p20[item] + s_count;
                a.upper[item] = upper;
                if (upper >= a.upper_min) append_item(a.next_items, a.next_count, item);
            }
        }
    }
}

// Lower re-entry (Y 0..10 of the columns whose upper mask is non-zero) and, for the items
// whose high count reaches high_min, the P19 features, score, extreme signal and mega
// topology reject (gpu_p20_benchmark.cpp:7473-7489).  Items that pass and are not rejected
// are appended for the coarse stage.  One 64-lane wave per item: it compacts the item's
// active columns (lane t reads the masks of columns 4t .. 4t + 3) and walks them 64 at a time.
__global__ __launch_bounds__(64, 2) void s0_lower_kernel(S0Args a) {
    __shared__ GenStage st;
    __shared__ uint8_t s_active[256];
    __shared__ uint8_t s_highest[256];
    __shared__ int s_high;
    __shared__ P19Features s_f;
    __shared__ double s_leaf[16];
    const int tid = threadIdx.x;
    const int64_t n = stage_items(a);
    for (int64_t it = blockIdx.x; it < n; it += gridDim.x) {
        const int64_t item = a.items[it];
        const int64_t world = item / a.centers;
        const int center = static_cast<int>(item % a.centers);
        const uint32_t* perms = a.perms + static_cast<size_t>(world) * kGenCount * 64;
        const double* offs = a.offsets + static_cast<size_t>(world) * kGenCount * 3;
        const uint32_t m4 = reinterpret_cast<const uint32_t*>(a.upper_mask + static_cast<size_t>(item) * 256)[tid];
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) cnt += ((m4 >> (8 * k)) & 0xFFu) != 0u;
        int incl = cnt;  // inclusive prefix count over the lanes
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (tid >= off) incl += v;
        }
        const int nact = __shfl(incl, 63);
        int at = incl - cnt;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (((m4 >> (8 * k)) & 0xFFu) != 0u) s_active[at++] = static_cast<uint8_t>(4 * tid + k);
        for (int c = tid; c < 256; c += 64) s_highest[c] = 0xFF;  // no re-entry / not a candidate column
        if (tid == 0) s_high = 0;
        __syncthreads();
        for (int base = 0; base < nact; base += 64) {
            const bool active = base + tid < nact;
            const int col = active ? s_active[base + tid] : 0;
            const double X = -28.0 + static_cast<double>((col >> 4) * 4 + a.center_dx[center] * 4);
            const double Z = -28.0 + static_cast<double>((col & 15) * 4 + a.center_dz[center] * 4);
            S0Column<0, 11> cc;
            evaluate_cached_column<0, 11, false>(st, perms, offs, X, Z, active, cc);
            const ColAcc<0, 11>& acc = cc.acc;
            if (active) {
                const size_t ci = static_cast<size_t>(item) * 256 + col;
                ColumnShape cs;
                cs.d5 = a.d5[ci];
                cs.d7 = a.d7[ci];
                uint32_t positive = static_cast<uint32_t>(a.upper_mask[ci]) << 11;
#pragma unroll
                for (int y = 0; y < 11; ++y)
                    if (lattice_density(cs, y, acc.n1[y], acc.n2[y], acc.n3[y]) > 0.0) positive |= 1u << y;
                const int h = highest_reentry(positive);
                s_highest[col] = h < 0 ? 0xFF : static_cast<uint8_t>(h);
                if (h >= 11) atomicAdd(&s_high, 1);
            }
        }
        __syncthreads();
        const int high = s_high;
        if (tid == 0) a.high[item] = high;
        if (high >= a.high_min) {  // buildMonsterFeatures (:1090-1114): one ballot per 64 cells
            unsigned long long b88[4], b96[4];
            int c104 = 0, c112 = 0;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const int h = s_highest[64 * w + tid];
                const bool valid = h != 0xFF;
                b88[w] = __ballot(valid && h >= 11);
                b96[w] = __ballot(valid && h >= 12);
                c104 += __popcll(__ballot(valid && h >= 13));
                c112 += __popcll(__ballot(valid && h >= 14));
            }
            if (tid == 0) {
                P19Features f;
                f.full_y88 = __popcll(b88[0]) + __popcll(b88[1]) + __popcll(b88[2]) + __popcll(b88[3]);
                f.full_y96 = __popcll(b96[0]) + __popcll(b96[1]) + __popcll(b96[2]) + __popcll(b96[3]);
                f.full_y104 = c104;
                f.full_y112 = c112;
                board_shape(b88, f.y88_largest, f.y88_width, f.y88_depth, f.y88_touches);
                board_shape(b96, f.y96_largest, f.y96_width, f.y96_depth, f.y96_touches);
                s_f = f;
            }
            __syncthreads();
            if (tid < 16) s_leaf[tid] = p19_tree(tid, a.upper[item], s_f);
            __syncthreads();
            if (tid == 0) {
                double sum = 0.0;
                for (int t = 0; t < 16; ++t) sum = sum + s_leaf[t];  // tree0 + ... + tree15
                const double sc = sum / 16;
                const P19Features& f = s_f;
                const bool extreme = f.y88_largest >= 14 || f.y96_largest >= 14;  // :77-79
                const bool pass = extreme || sc >= a.p19_min;
                // megaTopologyReject, gpu_p20_benchmark.cpp:1116-1128
                const bool mega = pass && a.mega_topology && !extreme && sc < 5.0 &&
                                  (f.full_y112 == 0 || f.y96_largest < 4);
                a.score[item] = sc;
                a.feats[item] = f;
                a.flags[item] = static_cast<uint8_t>((pass ? S0F_PASS : 0) | (extreme ? S0F_EXTREME : 0) |
                                                     (mega ? S0F_MEGA : 0));
                if (pass && !mega) append_item(a.next_items, a.next_count, item);
            }
        }
        __syncthreads();
    }
}

// The 61x61x17 coarse sign grid of one item (coarse_exact_core.hpp:218-386), built in LDS as
// one 61-bit z mask per (x, y) row and written to coarse_bits[list index].
__global__ __launch_bounds__(64, 3) void s0_coarse_signs_kernel(S0Args a) {
    __shared__ GenStage st;
    __shared__ unsigned long long rem[kCoarseRows];
    const int tid = threadIdx.x;
    const int64_t n = stage_items(a);
    for (int64_t it = blockIdx.x; it < n; it += gridDim.x) {
        const int64_t item = a.items[it];
        const int64_t world = item / a.centers;
        const int center = static_cast<int>(item % a.centers);
        const uint32_t* perms = a.perms + static_cast<size_t>(world) * kGenCount * 64;
        const double* offs = a.offsets + static_cast<size_t>(world) * kGenCount * 3;
        for (int r = tid; r < kCoarseRows; r += blockDim.x) rem[r] = 0;
        __syncthreads();
        for (int base = 0; base < 61 * 61; base += blockDim.x) {
            const int col = base + tid;
            const bool active = col < 61 * 61;
            const int cx = active ? col / 61 : 0, cz = active ? col % 61 : 0;
            const double X = static_cast<double>(-28 + cx + a.center_dx[center] * 4);
            const double Z = static_cast<double>(-28 + cz + a.center_dz[center] * 4);
            // the column shape (2D noise and climate) first, then the 3D sums of levels 0..7 and
            // of levels 8..15 in one loop body: eight levels of three octave sums at a time, with
            // nothing else live, fit two waves per SIMD
            const OneWorldBits s{&st, perms, offs};
            ColumnShape cs;
            {
                const double n4 = colpar_2d(s, kGenN4, 10, 1.121, X, Z, active);
                const double n5 = colpar_2d(s, kGenN5, 16, 200.0, X, Z, active);
                const double BX = X * 4.0 + 2.0, BZ = Z * 4.0 + 2.0;
                double temp = 0.0, rain = 0.0;
                colpar_climate_all<1>(s, &BX, &BZ, active ? 1 : 0, &temp, &rain);
                cs = column_shape(temp, rain, n4, n5);
            }
#pragma unroll 1
            for (int y0 = 0; y0 < 16; y0 += 8) {
                ColAcc<0, 8> acc;
                colpar_noise3_n<8>(s, X, Z, active, acc, y0);
                if (active) {
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                        if (lattice_density(cs, y0 + k, acc.n1[k], acc.n2[k], acc.n3[k]) > 0.0)
                            atomicOr(&rem[cx * 17 + y0 + k], 1ULL << cz);
                }
            }
        }
        __syncthreads();
        unsigned long long* bits = a.coarse_bits + static_cast<size_t>(it) * kCoarseRows;
        for (int r = tid; r < kCoarseRows; r += blockDim.x) bits[r] = rem[r];
        __syncthreads();
    }
}

// scoreSigns (coarse_exact_core.hpp:83-169) of one item's sign grid: the largest 6-connected
// positive component with maxY >= 8 that touches neither the bottom layer nor a side and has a
// positive cell below its lowest cell in some column.  Components are grown by bit-parallel
// dilation.
__global__ __launch_bounds__(256) void s0_coarse_score_kernel(S0Args a) {
    __shared__ unsigned long long rem[kCoarseRows];
    __shared__ unsigned long long sig[kCoarseRows];
    __shared__ unsigned long long comp[kCoarseRows];
    __shared__ int s_changed, s_seed, s_cells, s_maxy, s_bottom, s_side, s_reentry;
    const int tid = threadIdx.x;
    const int64_t n = stage_items(a);
    for (int64_t it = blockIdx.x; it < n; it += gridDim.x) {
        const int64_t item = a.items[it];
        const unsigned long long* bits = a.coarse_bits + static_cast<size_t>(it) * kCoarseRows;
        for (int r = tid; r < kCoarseRows; r += blockDim.x) rem[r] = bits[r];
        __syncthreads();
        for (int r = tid; r < kCoarseRows; r += blockDim.x) sig[r] = rem[r];
        __syncthreads();
        int best = 0;
        for (;;) {
            if (tid == 0) {
                s_seed = -1;
                for (int r = 0; r < kCoarseRows; ++r)
                    if (rem[r]) {
                        s_seed = r;
                        break;
                    }
            }
            __syncthreads();
            const int seed_row = s_seed;
            if (seed_row < 0) break;
            for (int r = tid; r < kCoarseRows; r += blockDim.x) comp[r] = 0;
            __syncthreads();
            if (tid == 0) comp[seed_row] = rem[seed_row] & (~rem[seed_row] + 1ULL);  // lowest set bit
            __syncthreads();
            for (;;) {
                if (tid == 0) s_changed = 0;
                __syncthreads();
                int changed = 0;
                for (int r = tid; r < kCoarseRows; r += blockDim.x) {
                    const int x = r / 17, y = r % 17;
                    const unsigned long long c = comp[r];
                    unsigned long long nb = c | (c << 1) | (c >> 1);
                    if (x > 0) nb |= comp[r - 17];
                    if (x < 60) nb |= comp[r + 17];
                    if (y > 0) nb |= comp[r - 1];
                    if (y < 16) nb |= comp[r + 1];
                    const unsigned long long nw = nb & rem[r];
                    if (nw != c) {
                        comp[r] = nw;
                        changed = 1;
                    }
                }
                if (changed) s_changed = 1;
                __syncthreads();
                const int any = s_changed;
                __syncthreads();
                if (!any) break;
            }
            if (tid == 0) {
                s_cells = 0;
                s_maxy = 0;
                s_bottom = 0;
                s_side = 0;
                s_reentry = 0;
            }
            __syncthreads();
            for (int x = tid; x < 61; x += blockDim.x) {
                int cells = 0, maxy = 0, bottom = 0, side = 0, reentry = 0;
                unsigned long long below = 0, started = 0;
                for (int y = 0; y < 17; ++y) {
                    const unsigned long long c = comp[x * 17 + y];
                    if (c) {
                        cells += __popcll(c);
                        maxy = y;
                        if (y == 0) bottom = 1;
                        if (x == 0 || x == 60 || (c & 1ULL) || (c >> 60)) side = 1;
                        const unsigned long long fresh = c & ~started;
                        if (fresh & below) reentry = 1;
                        started |= c;
                    }
                    below |= sig[x * 17 + y];
                }
                if (cells) {
                    atomicAdd(&s_cells, cells);
                    atomicMax(&s_maxy, maxy);
                    if (bottom) atomicOr(&s_bottom, 1);
                    if (side) atomicOr(&s_side, 1);
                    if (reentry) atomicOr(&s_reentry, 1);
                }
            }
            for (int r = tid; r < kCoarseRows; r += blockDim.x) rem[r] &= ~comp[r];
            __syncthreads();
            if (s_cells > best && s_maxy >= 8 && !s_bottom && !s_side && s_reentry) best = s_cells;
            __syncthreads();
        }
        if (tid == 0) a.coarse[item] = best;
        __syncthreads();
    }
}

// Records and closing-summary counters of a batch (gpu_p20_benchmark.cpp:7504-7571): one
// thread per item.  Counts clamp to u8, the coarse score is written as its low 16 bits.
__global__ __launch_bounds__(256) void s0_records_kernel(S0RecordArgs a) {
    __shared__ unsigned long long s_c[9];
    const int tid = threadIdx.x;
    if (tid < 9) s_c[tid] = 0;
    __syncthreads();
    unsigned long long c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t step = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + tid; i < a.n_items; i += step) {
        const int pa = a.p20[i], pb = a.upper[i], pc = a.high[i], pd = a.coarse[i];
        const int fl = a.flags[i];
        uint8_t* r = a.out + static_cast<size_t>(i) * a.record_bytes;
        auto u8 = [](int v) { return static_cast<uint8_t>(v < 0 ? 0 : (v > 255 ? 255 : v)); };
        auto u16 = [&](int at, int v) {
            const unsigned u = static_cast<unsigned>(v < 0 ? 0 : (v > 65535 ? 65535 : v));
            r[at] = static_cast<uint8_t>(u & 0xFFu);
            r[at + 1] = static_cast<uint8_t>(u >> 8);
        };
        r[0] = u8(pa);
        r[1] = u8(pb);
        r[2] = u8(pc);
        r[3] = (fl & S0F_PASS) ? 1 : 0;
        r[4] = static_cast<uint8_t>(pd & 0xFF);
        r[5] = static_cast<uint8_t>((pd >> 8) & 0xFF);
        r[6] = (fl & S0F_EXTREME) ? 1 : 0;
        r[7] = (fl & S0F_MEGA) ? 1 : 0;
        const unsigned long long sb = static_cast<unsigned long long>(__double_as_longlong(a.score[i]));
#pragma unroll
        for (int k = 0; k < 8; ++k) r[8 + k] = static_cast<uint8_t>(sb >> (8 * k));
        if (a.record_bytes == 34) {
            const P19Features& f = a.feats[i];
            u16(16, f.full_y88);
            u16(18, f.full_y96);
            u16(20, f.full_y104);
            u16(22, f.full_y112);
            u16(24, f.y88_largest);
            r[26] = static_cast<uint8_t>(f.y88_width);
            r[27] = static_cast<uint8_t>(f.y88_depth);
            r[28] = f.y88_touches ? 1 : 0;
            u16(29, f.y96_largest);
            r[31] = static_cast<uint8_t>(f.y96_width);
            r[32] = static_cast<uint8_t>(f.y96_depth);
            r[33] = f.y96_touches ? 1 : 0;
        }
        const bool g1 = pa >= a.p20_min, g2 = g1 && pb >= a.upper_min, g3 = g2 && pc >= a.high_min;
        c[0] += g1;
        c[1] += g2;
        if (g3) {
            ++c[2];
            ++c[3];
            if (fl & S0F_PASS) {
                ++c[4];
                if (fl & S0F_MEGA) {
                    ++c[6];
                } else {
                    ++c[7];
                    if (pd >= a.coarse_min) ++c[8];
                }
            } else {
                ++c[5];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k)
        if (c[k]) atomicAdd(&s_c[k], c[k]);
    __syncthreads();
    if (tid < 9 && s_c[tid]) atomicAdd(&a.counters[tid], s_c[tid]);
}

template __global__ void s0_upper_kernel<S0_P20>(S0Args);
template __global__ void s0_upper_kernel<S0_UPPER>(S0Args);

}  // namespace bsf
