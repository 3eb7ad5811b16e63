// bsf_host.hpp -- host-side context and helpers of libbsf.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "bsf_kernels.hpp"

namespace bsf {

struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct CapacityError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// Two exact forms of the same computation gave different results: a defect, never expected.
struct ParityError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char* what);
void set_error(const std::string& m);

// Grow-only device allocation.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t size = 0;
    void* ensure(size_t bytes);
    void release();
};

struct HostColumns {
    std::vector<int2> noise;
    std::vector<int4> dens;
    ColumnSet dev{};
    void upload(DeviceBuffer& nbuf, DeviceBuffer& dbuf, hipStream_t s);
};

struct Context {
    bool initialised = false;
    int device = 0;
    int cu_count = 0;
    int clock_khz = 0;  // peak shader clock (hipDeviceProp_t::clockRate)
    int density_resident = 0;  // resident 64-thread workgroups of the density kernel
    std::string name, arch;
    size_t global_mem = 0;
    hipStream_t stream = nullptr;
    DeviceBuffer seeds, out, scratch, scratch2, survivors, counters, lattice, hits, small;
    DeviceBuffer cols_noise, cols_dens, cols_noise2, cols_dens2;
    DeviceBuffer s0_a, s0_b, s0_c, s0_d, s0_e, s0_f, s0_g, s0_h;
    DeviceBuffer rg_seeds, rg_perms, rg_offs, rg_tasks, rg_lattice, rg_lists, rg_ctr, rg_scratch, rg_hits, rg_found;
    void release_all() {
        for (DeviceBuffer* b : {&seeds, &out, &scratch, &scratch2, &survivors, &counters, &lattice, &hits, &small,
                                &cols_noise, &cols_dens, &cols_noise2, &cols_dens2, &s0_a, &s0_b, &s0_c, &s0_d, &s0_e,
                                &s0_f, &s0_g, &s0_h, &rg_seeds, &rg_perms, &rg_offs, &rg_tasks, &rg_lattice, &rg_lists,
                                &rg_ctr, &rg_scratch, &rg_hits, &rg_found})
            b->release();
    }
};

extern Context g_ctx;

// HIP-event timing of named kernel launches (enabled by bsf_timing_enable).  Each launch is
// bracketed by two events on the stream it is launched on; durations resolve lazily.
struct LaunchTimer {
    const char* name;
    hipEvent_t start = nullptr;
    hipStream_t stream = nullptr;
    bool active = false;
    LaunchTimer(const char* n, hipStream_t s, bool allow = true);
    void stop(uint64_t units);
};
void resolve_timers();
// Adds units to a named timer whose launches were timed before the units were known (a stage
// that reads its item count from device memory).
void credit_timer_units(const char* name, uint64_t units);
bool timing_enabled();

void require_init();

// Chunk coordinates the noise keeps inside int range: the octave-0 coordinate of noise2 /
// noise3 is (4 chunk + i) * 684.412 + offset < 2^31 for |chunk| <= 2^19.  Beyond that Java's
// (int) casts saturate (the Far Lands); windows and chunks there are refused.
constexpr long long kMaxChunkCoord = 1LL << 19;
inline bool chunk_span_ok(long long first, long long count) {
    return first >= -kMaxChunkCoord && first + count - 1 <= kMaxChunkCoord;
}

extern std::mutex g_mutex;

// Runs f under the library lock and maps exceptions to the C ABI's status codes.
template <class F>
int guarded_call(F&& f) {
    std::lock_guard<std::mutex> lock(g_mutex);
    try {
        f();
        return BSF_OK;
    } catch (const ArgError& e) {
        set_error(e.what());
        return BSF_ERR_ARG;
    } catch (const CapacityError& e) {
        set_error(e.what());
        return BSF_ERR_CAPACITY;
    } catch (const ParityError& e) {
        set_error(e.what());
        return BSF_ERR_PARITY;
    } catch (const HipError& e) {
        set_error(e.what());
        return BSF_ERR_HIP;
    } catch (const std::exception& e) {
        set_error(e.what());
        return BSF_ERR_STATE;
    }
}

template <class K>
int resident_blocks(K kernel, int block, size_t dyn_lds);

}  // namespace bsf
