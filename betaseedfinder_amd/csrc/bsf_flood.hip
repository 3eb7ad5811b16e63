// bsf_flood.hip -- island flood (gfx950): one 512-thread workgroup per window, block bitsets
// of the window in LDS (or, for the larger windows, in a per-workgroup global scratch that
// stays in L2).  Used by the search engine (one window per Stage-2 survivor) and by the region
// scanner (windows of up to 49 x 49 chunks).
//
// Bitsets hold one z-mask per (x, y) row for y in [68, 127]: W 64-bit words per row (z bit
// 64 w + b in word w).  The flood reproduces IslandSearchEngine.findFloatingIslandInSeed's
// exact block scan (:2861-2977): candidate starts in Java order, 6-neighbour components above
// y = 67 (floodFillFlat, :4096-4202), isMatch (:3972-4003) and compareComponents (:4079-4094).
#include <hip/hip_runtime.h>

#include "bsf_island.hpp"
#include "bsf_region.hpp"

namespace bsf {

template <int W>
struct Z {
    unsigned long long v[W];
};

template <int W>
__device__ __forceinline__ Z<W> zload(const unsigned long long* p) {
    Z<W> z;
#pragma unroll
    for (int i = 0; i < W; ++i) z.v[i] = p[i];
    return z;
}
template <int W>
__device__ __forceinline__ void zstore(unsigned long long* p, const Z<W>& z) {
#pragma unroll
    for (int i = 0; i < W; ++i) p[i] = z.v[i];
}
template <int W>
__device__ __forceinline__ Z<W> zor(const Z<W>& a, const Z<W>& b) {
    Z<W> r;
#pragma unroll
    for (int i = 0; i < W; ++i) r.v[i] = a.v[i] | b.v[i];
    return r;
}
template <int W>
__device__ __forceinline__ Z<W> zand(const Z<W>& a, const Z<W>& b) {
    Z<W> r;
#pragma unroll
    for (int i = 0; i < W; ++i) r.v[i] = a.v[i] & b.v[i];
    return r;
}
template <int W>
__device__ __forceinline__ bool zeq(const Z<W>& a, const Z<W>& b) {
    bool e = true;
#pragma unroll
    for (int i = 0; i < W; ++i) e = e && a.v[i] == b.v[i];
    return e;
}
template <int W>
__device__ __forceinline__ bool zany(const Z<W>& a) {
    unsigned long long o = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) o |= a.v[i];
    return o != 0;
}
template <int W>
__device__ __forceinline__ int zpop(const Z<W>& a) {
    int n = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) n += __popcll(a.v[i]);
    return n;
}
template <int W>
__device__ __forceinline__ Z<W> zzero() {
    Z<W> r;
#pragma unroll
    for (int i = 0; i < W; ++i) r.v[i] = 0;
    return r;
}

// Bits of m lying in a run of consecutive ones of m that contains a bit of s (s within m):
// the carry of m + s sweeps each run upward from its seed bits; the same sum on the bit-
// reversed words sweeps downward.  The W words form one (64 W)-bit value, so the carries
// cross word boundaries.
template <int W>
__device__ __forceinline__ Z<W> zfill(const Z<W>& s, const Z<W>& m) {
    Z<W> r;
    unsigned long long carry = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const unsigned long long a = m.v[i], t = a + s.v[i];
        const unsigned long long u = t + carry;
        carry = (t < a ? 1ULL : 0ULL) | (u < t ? 1ULL : 0ULL);
        r.v[i] = ((u ^ a) & a) | s.v[i];
    }
    carry = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {  // reversed word i is word W - 1 - i, bit-reversed
        const unsigned long long a = __builtin_bitreverse64(m.v[W - 1 - i]);
        const unsigned long long t = a + __builtin_bitreverse64(s.v[W - 1 - i]);
        const unsigned long long u = t + carry;
        carry = (t < a ? 1ULL : 0ULL) | (u < t ? 1ULL : 0ULL);
        r.v[W - 1 - i] |= __builtin_bitreverse64((u ^ a) & a);
    }
    return r;
}

// The stone blocks of one lattice y-cell for all four block columns along x at once: out[i2]
// holds cell_stone_bits(..., i2) (bit 4 l1 + k2).  The same sums in the same order (Java's
// d10 += d12 walk over i2 is shared by the four columns instead of redone per column).
__device__ __forceinline__ void cell_stone_bits4(double d1, double d2, double d3, double d4, double e1, double e2,
                                                 double e3, double e4, uint32_t out[4]) {
    const double mx = fmax(fmax(fmax(d1, d2), fmax(d3, d4)), fmax(fmax(e1, e2), fmax(e3, e4)));
    const double mn = fmin(fmin(fmin(d1, d2), fmin(d3, d4)), fmin(fmin(e1, e2), fmin(e3, e4)));
    if (mx <= -kSignMargin || mn > kSignMargin) {
        const uint32_t v = mx <= -kSignMargin ? 0u : 0xFFFFFFFFu;
#pragma unroll
        for (int i2 = 0; i2 < 4; ++i2) out[i2] = v;
        return;
    }
    const double d5 = (e1 - d1) * 0.125;
    const double d6 = (e2 - d2) * 0.125;
    const double d7 = (e3 - d3) * 0.125;
    const double d8 = (e4 - d4) * 0.125;
#pragma unroll
    for (int i2 = 0; i2 < 4; ++i2) out[i2] = 0;
#pragma unroll
    for (int l1 = 0; l1 < 8; ++l1) {
        const double d12 = (d3 - d1) * 0.25;
        const double d13 = (d4 - d2) * 0.25;
        double d10 = d1, d11 = d2;
#pragma unroll
        for (int i2 = 0; i2 < 4; ++i2) {
            const double d16 = (d11 - d10) * 0.25;
            double d15 = d10;
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) {
                if (d15 > 0.0) out[i2] |= 1u << (4 * l1 + k2);
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

struct CompStats {
    int blocks, columns, minx, maxx, minz, maxz, miny, maxy;
    long long sumx, sumz;
    int bottom, side;
};

__device__ __forceinline__ int sum_of_bit_positions(unsigned long long m) {
    int s = 0;
    s += 1 * __popcll(m & 0xAAAAAAAAAAAAAAAAULL);
    s += 2 * __popcll(m & 0xCCCCCCCCCCCCCCCCULL);
    s += 4 * __popcll(m & 0xF0F0F0F0F0F0F0F0ULL);
    s += 8 * __popcll(m & 0xFF00FF00FF00FF00ULL);
    s += 16 * __popcll(m & 0xFFFF0000FFFF0000ULL);
    s += 32 * __popcll(m & 0xFFFFFFFF00000000ULL);
    return s;
}

// Java Math.round of a double of small magnitude: floor(a + 1/2), ties towards +infinity.
__device__ __forceinline__ int java_round(double a) {
    const double f = floor(a);
    int r = static_cast<int>(f);
    if (a - f >= 0.5) ++r;
    return r;
}

__device__ __forceinline__ bool comp_better(const CompStats& a, const CompStats& b) {
    // compareComponents(a, b) > 0, IslandSearchEngine.java:4079-4094
    if (a.blocks != b.blocks) return a.blocks > b.blocks;
    if (a.columns != b.columns) return a.columns > b.columns;
    const int areaA = (a.maxx - a.minx + 1) * (a.maxz - a.minz + 1);
    const int areaB = (b.maxx - b.minx + 1) * (b.maxz - b.minz + 1);
    return areaA > areaB;
}

// One sweep along y of the x slices x0 .. x1 (up, then down; thread t takes slices x0 + t,
// x0 + t + blockDim.x, ...): adds the blocks of rem that touch comp within the slice or in the
// neighbouring slices' rows.  Returns whether it grew.  With s_lo, a slice outside
// [lo, hi] that gains blocks widens *s_lo / *s_hi.
template <int W>
__device__ int sweep_y(unsigned long long* comp, const unsigned long long* rem, int nx, int x0, int x1,
                       int* s_lo = nullptr, int* s_hi = nullptr, int lo = 0, int hi = 0) {
    int changed = 0;
    for (int t = x0 + threadIdx.x; t <= x1; t += blockDim.x) {
        unsigned long long* row = comp + static_cast<size_t>(t) * kRowStride * W;
        const unsigned long long* mrow = rem + static_cast<size_t>(t) * kRowStride * W;
        const unsigned long long* left = t > 0 ? row - kRowStride * W : nullptr;
        const unsigned long long* right = t < nx - 1 ? row + kRowStride * W : nullptr;
        int grew = 0;
        Z<W> prev = zzero<W>();
        for (int y = 0; y < kLevels; ++y) {
            const Z<W> m = zload<W>(mrow + y * W);
            const Z<W> c = zload<W>(row + y * W);
            Z<W> nb = zor(c, prev);
            if (left) nb = zor(nb, zload<W>(left + y * W));
            if (right) nb = zor(nb, zload<W>(right + y * W));
            const Z<W> n = zfill(zand(nb, m), m);
            if (!zeq(n, c)) {
                zstore(row + y * W, n);
                grew = 1;
            }
            prev = n;
        }
        prev = zzero<W>();
        for (int y = kLevels - 1; y >= 0; --y) {
            const Z<W> m = zload<W>(mrow + y * W);
            const Z<W> c = zload<W>(row + y * W);
            const Z<W> n = zfill(zand(zor(c, prev), m), m);
            if (!zeq(n, c)) {
                zstore(row + y * W, n);
                grew = 1;
            }
            prev = n;
        }
        if (grew && s_lo && (t < lo || t > hi)) {
            atomicMin(s_lo, t);
            atomicMax(s_hi, t);
        }
        changed |= grew;
    }
    return changed;
}

// One sweep of thread y along x over its y level (forward, then back) over x0 .. nx - 1 and
// back.  With s_lo (comp lies in slices x0 .. x1), the forward walk ends past x1 once nothing
// propagates, the backward walk likewise below x0, and the slices reached widen *s_lo / *s_hi.
template <int W>
__device__ int sweep_x(unsigned long long* comp, const unsigned long long* rem, int nx, int x0 = 0, int x1 = 0,
                       int* s_lo = nullptr, int* s_hi = nullptr) {
    const int t = threadIdx.x;
    int changed = 0;
    if (t < kLevels) {
        const bool bounded = s_lo != nullptr;
        int xmin = x0, xmax = x1;
        Z<W> prev = zzero<W>();
        for (int x = bounded ? x0 : 0; x < nx; ++x) {
            const size_t r = (static_cast<size_t>(x) * kRowStride + t) * W;
            const Z<W> m = zload<W>(rem + r);
            const Z<W> c = zload<W>(comp + r);
            Z<W> nb = zor(c, prev);
            if (t > 0) nb = zor(nb, zload<W>(comp + r - W));
            if (t < kLevels - 1) nb = zor(nb, zload<W>(comp + r + W));
            const Z<W> n = zfill(zand(nb, m), m);
            if (!zeq(n, c)) {
                zstore(comp + r, n);
                changed = 1;
                xmax = x > xmax ? x : xmax;
            }
            prev = n;
            if (bounded && x >= xmax && !zany(n)) break;
        }
        prev = zzero<W>();
        for (int x = bounded ? xmax : nx - 1; x >= 0; --x) {
            const size_t r = (static_cast<size_t>(x) * kRowStride + t) * W;
            const Z<W> m = zload<W>(rem + r);
            const Z<W> c = zload<W>(comp + r);
            const Z<W> n = zfill(zand(zor(c, prev), m), m);
            if (!zeq(n, c)) {
                zstore(comp + r, n);
                changed = 1;
                xmin = x < xmin ? x : xmin;
            }
            prev = n;
            if (bounded && x <= xmin && !zany(n)) break;
        }
        if (bounded && changed) {
            atomicMin(s_lo, xmin);
            atomicMax(s_hi, xmax);
        }
    }
    return changed;
}

// Number of blocks in rem and not in comp, over the whole window (block-wide sum).
template <int W>
__device__ int count_outside(const unsigned long long* rem, const unsigned long long* comp, int rows, int* s_total) {
    if (threadIdx.x == 0) *s_total = 0;
    __syncthreads();
    int mine = 0;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
        const Z<W> m = zload<W>(rem + static_cast<size_t>(r) * W);
        const Z<W> c = zload<W>(comp + static_cast<size_t>(r) * W);
#pragma unroll
        for (int i = 0; i < W; ++i) mine += __popcll(m.v[i] & ~c.v[i]);
    }
    atomicAdd(s_total, mine);
    __syncthreads();
    return *s_total;
}

// One directional pass of slice x along y (UP: y ascending, else descending): each row takes the
// blocks of rem that touch comp in the row itself, the row passed before it or the neighbouring
// slices' rows.  Returns whether it grew.
template <int W, bool UP>
__device__ int pass_y(unsigned long long* comp, const unsigned long long* rem, int nx, int x) {
    unsigned long long* row = comp + static_cast<size_t>(x) * kRowStride * W;
    const unsigned long long* mrow = rem + static_cast<size_t>(x) * kRowStride * W;
    const unsigned long long* left = x > 0 ? row - kRowStride * W : nullptr;
    const unsigned long long* right = x < nx - 1 ? row + kRowStride * W : nullptr;
    int grew = 0;
    Z<W> prev = zzero<W>();
    for (int k = 0; k < kLevels; ++k) {
        const int y = UP ? k : kLevels - 1 - k;
        const Z<W> m = zload<W>(mrow + y * W);
        const Z<W> c = zload<W>(row + y * W);
        Z<W> nb = zor(c, prev);
        if (left) nb = zor(nb, zload<W>(left + y * W));
        if (right) nb = zor(nb, zload<W>(right + y * W));
        const Z<W> n = zfill(zand(nb, m), m);
        if (!zeq(n, c)) {
            zstore(row + y * W, n);
            grew = 1;
        }
        prev = n;
    }
    return grew;
}

// The same along x for level y (FWD: x ascending, else descending), with the rows above and
// below as neighbours.
template <int W, bool FWD>
__device__ int pass_x(unsigned long long* comp, const unsigned long long* rem, int nx, int y) {
    int grew = 0;
    Z<W> prev = zzero<W>();
    for (int k = 0; k < nx; ++k) {
        const int x = FWD ? k : nx - 1 - k;
        const size_t r = (static_cast<size_t>(x) * kRowStride + y) * W;
        const Z<W> m = zload<W>(rem + r);
        const Z<W> c = zload<W>(comp + r);
        Z<W> nb = zor(c, prev);
        if (y > 0) nb = zor(nb, zload<W>(comp + r - W));
        if (y < kLevels - 1) nb = zor(nb, zload<W>(comp + r + W));
        const Z<W> n = zfill(zand(nb, m), m);
        if (!zeq(n, c)) {
            zstore(comp + r, n);
            grew = 1;
        }
        prev = n;
    }
    return grew;
}

// Grow comp to its fixpoint inside rem: 6-neighbour connectivity over the rows of the window
// (x, y in [68, 127]) with z along the bits (z runs fill in one step, zfill).  Returns false,
// without finishing, when after the first round fewer than min_blocks blocks of rem lie outside
// comp (comp only grows, so that bounds what the fixpoint leaves outside).  Each round runs
// the four directional passes at once, one per wave (y up and y down over the slices, waves 0 /
// 4 and 1 / 5 taking alternate groups of 64 slices; x forward and backward over the levels,
// waves 2 and 3).  Every step only adds blocks of rem adjacent to comp, so whatever a pass
// reads of the others' concurrent writes, comp stays inside the component.  A round in which
// no pass grew wrote nothing, so all four read the same comp, and each row was closed against
// the row below and its x neighbours (y up), the row above (y down) and its x and y neighbours
// (x passes): comp is then the whole component.
template <int W>
__device__ bool grow_component(unsigned long long* comp, const unsigned long long* rem, int nx, int rows,
                               int min_blocks, int* s_total) {
    static_assert(kFloodThreads == 512, "the directional passes take eight waves");
    const int w = static_cast<int>(threadIdx.x) >> 6, lane = static_cast<int>(threadIdx.x) & 63;
    for (bool first = true;; first = false) {
        int changed = 0;
        if ((w & 3) == 0) {
            for (int x = lane + 64 * (w >> 2); x < nx; x += 128) changed |= pass_y<W, true>(comp, rem, nx, x);
        } else if ((w & 3) == 1) {
            for (int x = lane + 64 * (w >> 2); x < nx; x += 128) changed |= pass_y<W, false>(comp, rem, nx, x);
        } else if (w == 2) {
            for (int y = lane; y < kLevels; y += 64) changed |= pass_x<W, true>(comp, rem, nx, y);
        } else if (w == 3) {
            for (int y = lane; y < kLevels; y += 64) changed |= pass_x<W, false>(comp, rem, nx, y);
        }
        if (!__syncthreads_or(changed)) break;
        if (first && count_outside<W>(rem, comp, rows, s_total) < min_blocks) return false;
    }
    return true;
}

// The same for a connected comp whose blocks lie in slices *s_lo .. *s_hi (a connected set's
// x projection is an interval): each round sweeps y over that interval widened by one slice
// and x from it outwards while blocks propagate, and widens the interval as comp grows.  A
// round without change has looked at every slice that holds comp or borders it, so comp is
// closed.
template <int W>
__device__ void grow_component_in(unsigned long long* comp, const unsigned long long* rem, int nx, int* s_lo,
                                  int* s_hi) {
    for (;;) {
        const int lo = *s_lo, hi = *s_hi;
        __syncthreads();  // every thread has the same slices before any widens them
        int changed = sweep_y<W>(comp, rem, nx, lo > 0 ? lo - 1 : 0, hi < nx - 1 ? hi + 1 : nx - 1, s_lo, s_hi,
                                 lo, hi);
        __syncthreads();
        changed |= sweep_x<W>(comp, rem, nx, *s_lo, *s_hi, s_lo, s_hi);
        if (!__syncthreads_or(changed)) break;
    }
}

// Shared state of one workgroup's flood.
template <int W>
struct FloodShared {
    CompStats cs;
    CompStats best;                  // best matching component so far (written by thread 0)
    int have_best;
    unsigned long long zor[W];
    int total;
    int lo, hi;                      // x slices of the growing component
    unsigned long long starts[kFloodThreads];  // candidate-start levels of a slice of sample columns
};

// The flood of one window: upsampling of its lattice (levels 8..15 of every density column,
// [chunk x][chunk z][25][8] doubles) into the bitsets at base, ground removal, candidate
// starts in Java order.  Returns (in every thread) whether a component matched, and the best
// one in *best.  Every thread of the workgroup calls it.
template <int W>
__device__ bool flood_window(unsigned long long* const base, const double* __restrict__ lat,
                             const bsf_island_params& p, FloodShared<W>& sh, CompStats* best_out) {
    const int tid = threadIdx.x;
    const int nx = p.chunks_x * 16;
    const int nz = p.chunks_z * 16;
    const int rows = nx * kRowStride;
    unsigned long long* const rem = base;
    unsigned long long* const comp = base + static_cast<size_t>(rows) * W;
    unsigned long long* const st67 = base + static_cast<size_t>(2 * rows) * W;
    const int minWZ = p.chunk_z0 * 16;

    // 1. upsample the window: blocks y in [67, 127] from lattice y-cells 8..15.  A task is one
    //    y-cell level of one x cell (four block columns) in one 16-block z slot (one chunk): its
    //    four lattice cells along z, each for all four block columns (cell_stone_bits4), written
    //    as the 16-bit pieces of the rows' z words that the slot covers (zero past the window).
    //    Level-major: the lanes of a wave share a level, so the all-air levels above the
    //    terrain take the short path together.
    {
        const int nxc = nx >> 2, nzc = 4 * W;
        uint16_t* const rem16 = reinterpret_cast<uint16_t*>(rem);
        uint16_t* const st16 = reinterpret_cast<uint16_t*>(st67);
        for (int task = tid; task < 8 * nxc * nzc; task += blockDim.x) {
            const int k1 = 8 + task / (nxc * nzc), rest = task % (nxc * nzc);
            const int xc = rest % nxc, zc = rest / nxc;
            uint32_t piece[4][8];
#pragma unroll
            for (int i2 = 0; i2 < 4; ++i2)
#pragma unroll
                for (int l1 = 0; l1 < 8; ++l1) piece[i2][l1] = 0;
            if (zc < p.chunks_z) {
                const int i1 = xc & 3;
                const double* cl = lat + static_cast<size_t>((xc >> 2) * p.chunks_z + zc) * 25 * 8;
                auto L = [&](int ii, int jj, int kk) -> double {
                    return kk >= 16 ? -10.0 : cl[(ii * 5 + jj) * 8 + (kk - 8)];
                };
#pragma unroll 1
                for (int j1 = 0; j1 < 4; ++j1) {
                    uint32_t m[4];
                    cell_stone_bits4(L(i1, j1, k1), L(i1, j1 + 1, k1), L(i1 + 1, j1, k1), L(i1 + 1, j1 + 1, k1),
                                     L(i1, j1, k1 + 1), L(i1, j1 + 1, k1 + 1), L(i1 + 1, j1, k1 + 1),
                                     L(i1 + 1, j1 + 1, k1 + 1), m);
#pragma unroll
                    for (int i2 = 0; i2 < 4; ++i2)
#pragma unroll
                        for (int l1 = 0; l1 < 8; ++l1) piece[i2][l1] |= ((m[i2] >> (4 * l1)) & 0xFu) << (4 * j1);
                }
            }
            // piece zc of word zc / 4 of a row: its 16-bit half-word 4 (row word) + zc % 4
            const int wq = (zc >> 2) * 4 + (zc & 3);
#pragma unroll
            for (int i2 = 0; i2 < 4; ++i2) {
                const int x = 4 * xc + i2;
#pragma unroll
                for (int l1 = 0; l1 < 8; ++l1) {
                    const int y = k1 * 8 + l1;
                    if (y >= kFloorY)
                        rem16[(static_cast<size_t>(x) * kRowStride + (y - kFloorY)) * W * 4 + wq] =
                                static_cast<uint16_t>(piece[i2][l1]);
                    else if (y == kFloorY - 1)
                        st16[static_cast<size_t>(x) * W * 4 + wq] = static_cast<uint16_t>(piece[i2][l1]);
                }
                if (k1 == 8) rem16[(static_cast<size_t>(x) * kRowStride + kLevels) * W * 4 + wq] = 0;  // row padding
            }
        }
    }
    __syncthreads();
    // 2. every component that reaches a y = 68 block with stone below touches the bottom
    //    and can never match (isMatch): grow them all at once from their ground contacts
    //    and drop them, as Java's flood would mark them visited.
    for (int r = tid; r < rows; r += blockDim.x) {
        const int x = r / kRowStride, y = r % kRowStride;
        const Z<W> v = y == 0 ? zand(zload<W>(rem + static_cast<size_t>(r) * W), zload<W>(st67 + static_cast<size_t>(x) * W))
                              : zzero<W>();
        zstore(comp + static_cast<size_t>(r) * W, v);
    }
    __syncthreads();
    // The first round from the contacts already reaches most grounded stone; what it leaves
    // bounds the floating stone from above, and fewer than minBlocks floating blocks cannot
    // hold a match (most windows end there).  Otherwise the ground components in full.
    if (!grow_component<W>(comp, rem, nx, rows, p.min_blocks, &sh.total)) return false;
    // 3. all remaining stone is floating
    const int floating = count_outside<W>(rem, comp, rows, &sh.total);
    if (floating < p.min_blocks) return false;
    for (int r = tid; r < rows; r += blockDim.x) {
        Z<W> left = zload<W>(rem + static_cast<size_t>(r) * W);
        const Z<W> c = zload<W>(comp + static_cast<size_t>(r) * W);
#pragma unroll
        for (int i = 0; i < W; ++i) left.v[i] &= ~c.v[i];
        zstore(rem + static_cast<size_t>(r) * W, left);
        zstore(comp + static_cast<size_t>(r) * W, zzero<W>());
    }
    __syncthreads();

    // The best component lives in LDS, updated by thread 0: kept in registers across the
    // candidate loop, hipcc (ROCm 7.2, gfx950) dropped the copy of its min y on the path where
    // a later component replaces an earlier match.
    if (tid == 0) sh.have_best = 0;
    // 4. candidate starts in Java order: chunk x, chunk z, x in {4,12}, z in {4,12}, y up.
    //    Sample columns are taken in slices of blockDim.x: each thread finds the start levels
    //    of one column (stone over non-stone: y = 68 over the y = 67 row, above that over a
    //    block that is not in rem -- a stone block below a floating block is floating too),
    //    then the slice's starts are walked in order, skipping blocks an earlier flood took.
    const int npairs = p.chunks_x * p.chunks_z * 4;
    for (int p0 = 0; p0 < npairs; p0 += blockDim.x) {
        {
            const int pi = p0 + tid;
            unsigned long long m = 0;
            if (pi < npairs) {
                const int ch = pi >> 2, sc = pi & 3;
                const int x = (ch / p.chunks_z) * 16 + ((sc >> 1) ? 12 : 4);
                const int z = (ch % p.chunks_z) * 16 + ((sc & 1) ? 12 : 4);
                const int zw = z >> 6;
                const unsigned long long bitz = 1ULL << (z & 63);
                const unsigned long long* col = rem + static_cast<size_t>(x) * kRowStride * W + zw;
                bool below = (st67[static_cast<size_t>(x) * W + zw] & bitz) != 0;
                for (int yy = 0; yy < kLevels; ++yy) {
                    const bool here = (col[static_cast<size_t>(yy) * W] & bitz) != 0;
                    if (here && !below) m |= 1ULL << yy;
                    below = here;
                }
            }
            sh.starts[tid] = m;
        }
        __syncthreads();
        const int slice = npairs - p0 < static_cast<int>(blockDim.x) ? npairs - p0 : static_cast<int>(blockDim.x);
        for (int k = 0; k < slice; ++k) {
            unsigned long long todo = sh.starts[k];
            const int ch = (p0 + k) >> 2, sc = (p0 + k) & 3;
            const int x = (ch / p.chunks_z) * 16 + ((sc >> 1) ? 12 : 4);
            const int z = (ch % p.chunks_z) * 16 + ((sc & 1) ? 12 : 4);
            const int zw = z >> 6;
            const unsigned long long bitz = 1ULL << (z & 63);
            while (todo) {
                const int yy = __ffsll(static_cast<long long>(todo)) - 1;
                todo &= todo - 1;
                // stone not yet visited (visited or grounded blocks are no longer in rem)
                if ((rem[(static_cast<size_t>(x) * kRowStride + yy) * W + zw] & bitz) == 0) continue;
                // comp is all clear here: each component's slices are cleared after its statistics
                if (tid == 0) {
                    comp[(static_cast<size_t>(x) * kRowStride + yy) * W + zw] = bitz;
                    sh.lo = x;
                    sh.hi = x;
                }
                __syncthreads();
                grow_component_in<W>(comp, rem, nx, &sh.lo, &sh.hi);
                const int xlo = sh.lo, xhi = sh.hi;
                // statistics of the component
                if (tid == 0) {
                    sh.cs.blocks = 0;
                    sh.cs.columns = 0;
                    sh.cs.minx = 0x7fffffff;
                    sh.cs.maxx = -0x7fffffff;
                    sh.cs.miny = 0x7fffffff;
                    sh.cs.maxy = -0x7fffffff;
                    sh.cs.minz = 0;
                    sh.cs.maxz = 0;
                    sh.cs.sumx = 0;
                    sh.cs.sumz = 0;
                    sh.cs.bottom = 0;
                    sh.cs.side = 0;
#pragma unroll
                    for (int i = 0; i < W; ++i) sh.zor[i] = 0;
                }
                __syncthreads();
                for (int rx = xlo + tid; rx <= xhi; rx += blockDim.x) {
                    Z<W> col = zzero<W>();
                    int blk = 0, ymin = 0x7fffffff, ymax = -0x7fffffff;
                    for (int ry = 0; ry < kLevels; ++ry) {
                        const Z<W> c = zload<W>(comp + (static_cast<size_t>(rx) * kRowStride + ry) * W);
                        if (zany(c)) {
                            col = zor(col, c);
                            blk += zpop(c);
                            ymin = min(ymin, ry + kFloorY);
                            ymax = max(ymax, ry + kFloorY);
                        }
                    }
                    if (zany(col)) {
                        const int ncol = zpop(col);
                        long long zsum = static_cast<long long>(ncol) * minWZ;
#pragma unroll
                        for (int i = 0; i < W; ++i)
                            zsum += sum_of_bit_positions(col.v[i]) + 64LL * i * __popcll(col.v[i]);
                        atomicAdd(&sh.cs.blocks, blk);
                        atomicAdd(&sh.cs.columns, ncol);
                        atomicMin(&sh.cs.minx, rx);
                        atomicMax(&sh.cs.maxx, rx);
                        atomicMin(&sh.cs.miny, ymin);
                        atomicMax(&sh.cs.maxy, ymax);
                        atomicAdd(reinterpret_cast<unsigned long long*>(&sh.cs.sumx),
                                  static_cast<unsigned long long>(static_cast<long long>(ncol) * (p.chunk_x0 * 16 + rx)));
                        atomicAdd(reinterpret_cast<unsigned long long*>(&sh.cs.sumz), static_cast<unsigned long long>(zsum));
#pragma unroll
                        for (int i = 0; i < W; ++i)
                            if (col.v[i]) atomicOr(&sh.zor[i], col.v[i]);
                        const Z<W> c0 = zload<W>(comp + static_cast<size_t>(rx) * kRowStride * W);
                        if (zany(zand(c0, zload<W>(st67 + static_cast<size_t>(rx) * W)))) atomicOr(&sh.cs.bottom, 1);
                        if (rx == 0 || rx == nx - 1) atomicOr(&sh.cs.side, 1);
                    }
                }
                __syncthreads();  // the statistics have read comp
                for (size_t r = static_cast<size_t>(xlo) * kRowStride * W + tid;
                     r < static_cast<size_t>(xhi + 1) * kRowStride * W; r += blockDim.x) {
                    rem[r] &= ~comp[r];
                    comp[r] = 0;
                }
                __syncthreads();
                CompStats cs = sh.cs;
                int lo = -1, hi = -1;  // lowest and highest z of the component
#pragma unroll
                for (int i = 0; i < W; ++i) {
                    const unsigned long long zo = sh.zor[i];
                    if (zo != 0) {
                        if (lo < 0) lo = 64 * i + __ffsll(static_cast<long long>(zo)) - 1;
                        hi = 64 * i + 63 - __clzll(static_cast<long long>(zo));
                    }
                }
                cs.minz = lo;
                cs.maxz = hi;
                if (cs.minz == 0 || cs.maxz == nz - 1) cs.side = 1;
                // isMatch, IslandSearchEngine.java:3972-4003
                const int width = cs.maxx - cs.minx + 1;
                const int depth = cs.maxz - cs.minz + 1;
                const bool match = cs.blocks >= p.min_blocks && cs.columns >= p.min_columns &&
                                   cs.maxy >= p.min_y_for_match && !cs.bottom && width >= p.min_width &&
                                   depth >= p.min_depth &&
                                   cs.blocks / static_cast<double>(cs.columns > 1 ? cs.columns : 1) >=
                                           p.min_avg_thickness;
                if (tid == 0 && match && (!sh.have_best || comp_better(cs, sh.best))) {
                    sh.best = cs;
                    sh.have_best = 1;
                }
                __syncthreads();
            }
        }
        __syncthreads();  // the slice's start masks are read before the next slice rewrites them
    }
    *best_out = sh.best;
    return sh.have_best != 0;
}

// The hit record of a window's best component.
__device__ bsf_island_hit make_hit(int64_t seed, const bsf_island_params& p, const CompStats& best, int32_t tag) {
    const int minWX = p.chunk_x0 * 16, minWZ = p.chunk_z0 * 16;
    bsf_island_hit h;
    h.seed = seed;
    h.min_x = minWX + best.minx;
    h.max_x = minWX + best.maxx;
    h.min_z = minWZ + best.minz;
    h.max_z = minWZ + best.maxz;
    h.min_y = best.miny;
    h.max_y = best.maxy;
    h.blocks = best.blocks;
    h.columns = best.columns;
    h.center_x = java_round(static_cast<double>(best.sumx) / best.columns);
    h.center_z = java_round(static_cast<double>(best.sumz) / best.columns);
    h.chunk_x = h.center_x >> 4;
    h.chunk_z = h.center_z >> 4;
    h.flags = best.side ? 1 : 0;
    h.window_index = tag;
    return h;
}

template <int W, bool GLOBAL>
__global__ __launch_bounds__(kFloodThreads, 4) void pass_flood_kernel(EngineArgs a, int cls) {
    extern __shared__ __attribute__((aligned(16))) unsigned long long flood_lds[];
    __shared__ FloodShared<W> sh;
    __shared__ unsigned long long s_item;
    const int tid = threadIdx.x;
    if (a.audit != kAuditNone) return;
    unsigned long long* const base =
            GLOBAL ? a.flood_scratch + static_cast<size_t>(blockIdx.x) * a.flood_scratch_words : flood_lds;
    const unsigned long long n_items = a.ctr[kCtrFloodCount + cls];
    const int32_t* list = a.flood_list + static_cast<size_t>(cls) * a.pass_slots;

    for (;;) {
        if (tid == 0) s_item = atomicAdd(&a.ctr[kCtrFloodNext + cls], 1ULL);
        __syncthreads();
        const unsigned long long item = s_item;
        __syncthreads();
        if (item >= n_items) break;
        const int64_t sv = list[item];
        const int32_t jix = a.surv_job[sv];
        const bsf_island_params p = a.jobs[jix].p;
        CompStats best;
        const bool found = flood_window<W>(base, a.lattice + static_cast<size_t>(sv) * a.lat_stride, p, sh, &best);
        if (tid == 0 && found) {
            const unsigned long long slot = atomicAdd(&a.ctr[kCtrHits], 1ULL);
            atomicAdd(&a.jobs[jix].hits, 1ULL);
            if (static_cast<int64_t>(slot) < a.pass_slots) a.hits[slot] = make_hit(a.surv_seed[sv], p, best, a.jobs[jix].tag);
        }
        __syncthreads();
    }
}

// Region scanner: one workgroup per window of one flood bucket, bitsets in the workgroup's
// slice of the global scratch.
template <int W>
__global__ __launch_bounds__(kFloodThreads) void region_flood_kernel(RegionArgs a, int bucket) {
    __shared__ FloodShared<W> sh;
    __shared__ unsigned long long s_item;
    const int tid = threadIdx.x;
    unsigned long long* const base = a.flood_scratch + static_cast<size_t>(blockIdx.x) * a.flood_scratch_words;
    const int32_t* list = a.flood_list + static_cast<size_t>(bucket) * a.ntasks;
    const unsigned long long n_items = static_cast<unsigned long long>(a.flood_count[bucket]);
    for (;;) {
        if (tid == 0) s_item = atomicAdd(&a.ctr[kRegionCtrFlood + bucket], 1ULL);
        __syncthreads();
        const unsigned long long item = s_item;
        __syncthreads();
        if (item >= n_items) break;
        const int k = list[item];
        const RegionTask& t = a.tasks[k];
        const bsf_island_params p = t.p;
        CompStats best;
        const bool found = flood_window<W>(base, a.lattice + t.lat_off, p, sh, &best);
        if (tid == 0) {
            a.found[k] = found ? 1 : 0;
            if (found) a.hits[k] = make_hit(t.seed, p, best, 0);
        }
        __syncthreads();
    }
}

template __global__ void region_flood_kernel<1>(RegionArgs, int);
template __global__ void region_flood_kernel<2>(RegionArgs, int);
template __global__ void region_flood_kernel<4>(RegionArgs, int);
template __global__ void region_flood_kernel<7>(RegionArgs, int);
template __global__ void region_flood_kernel<9>(RegionArgs, int);
template __global__ void region_flood_kernel<13>(RegionArgs, int);

template __global__ void pass_flood_kernel<1, false>(EngineArgs, int);
template __global__ void pass_flood_kernel<2, false>(EngineArgs, int);
template __global__ void pass_flood_kernel<2, true>(EngineArgs, int);

}  // namespace bsf
