// runtime.hpp - host runtime: device handle, resident scenes, kernel launches.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/embloks_hip.h"
#include "common.hpp"
#include "setup.hpp"

namespace emb {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define EMB_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return ::emb::fail(EMB_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Body of a C ABI function: C++ exceptions become an error code.
#define EMB_GUARD(...)                                                         \
    try {                                                                      \
        __VA_ARGS__                                                            \
    } catch (const std::exception &e) {                                        \
        return ::emb::fail(EMB_E_INVALID, std::string("exception: ") + e.what()); \
    }

struct DeviceTables {
    float *g = nullptr;
    float *square = nullptr;
    int square_n = 0;
    double half = 0.0;
    int reach = 0;
    float *ion_sum = nullptr;
    int ion_a0 = 0;
    float2 *g_lerp = nullptr;
    float2 *square_f2 = nullptr;
};

}  // namespace emb

namespace emb {
// The DST-I matrix of one length on the device, whole and as its halves
// (sine_halves): the tick's products use the halves.
struct SineSet {
    float *full = nullptr, *e = nullptr, *e_t = nullptr, *o = nullptr;
    size_t bytes = 0;       // device bytes of the four matrices
    int refs = 0;           // resident scenes and Poisson solvers using this length
    uint64_t last_use = 0;  // EmbDevice::sine_clock when last taken or released
};
// Bytes of the sine set of length n (sine_matrix plus sine_halves).
size_t sine_set_bytes(int n);
// Unreferenced sine sets are kept for reuse up to this many bytes per device
// (least recently used ones freed first).
constexpr size_t SINE_IDLE_BYTES = (size_t)512 << 20;
}  // namespace emb

struct EmbDevice {
    int ordinal = 0;
    int n_cu = 0;
    double *scratch = nullptr;   // this device's copy of runtime.hip's `scratch` array (hipGetSymbolAddress)
    std::mutex lock;
    std::map<int, emb::SineSet> sine;   // DST-I matrices per length, shared by every scene
    size_t sine_bytes = 0;              // device bytes of all sets in `sine`
    uint64_t sine_clock = 0;
    std::map<std::tuple<double, int, double>, emb::DeviceTables> tables;
};

namespace emb {

// One scene resident on the device: its arena, descriptor and host-side state.
struct Resident {
    EmbDevice *dev = nullptr;
    EmbSettings s{};
    HostScene hs;                 // FP64 preparation (electron arrays dropped after upload)
    void *arena = nullptr;
    size_t bytes = 0;
    DevScene ds{};                // host copy of the descriptor
    DevScene *d_scene = nullptr;  // the descriptor in device memory
    int32_t *d_prefix = nullptr;  // single-scene block prefix
    Batch batch{};                // single-scene batch
    float *pin_tab = nullptr;     // pinned staging of the per-tick tables
    hipEvent_t tab_event = nullptr;
    size_t tab_floats = 0;
    double time = 0.0;
    uint64_t tick = 0;            // ticks enqueued
    int last_substeps = 0;
    long long *dep_dump = nullptr;
    int sine_nx = 0, sine_ny = 0; // sine lengths this scene holds references to (0: none)
};

// The close-range tables of (softening, passes, cutoff), built on `dev` in FP64
// and cached for the process (tables.hip); a no-op when cutoff <= 0 or built.
int ensure_tables(EmbDevice *dev, double softening, int passes, double cutoff);

// The FP32 DST-I matrix of length n on the device (built once per device, then
// shared), with a reference taken: release it with release_sine.
int device_sine(EmbDevice *dev, int n, const SineSet **out);
void release_sine(EmbDevice *dev, int n);
// Device bytes that admitting a scene of nx x ny nodes adds in sine sets not yet
// held, and the bytes held now (referenced or idle).
size_t sine_missing_bytes(EmbDevice *dev, int nx, int ny);
size_t sine_held_bytes(EmbDevice *dev);
// Free idle (unreferenced) sine sets, least recently used first, until at most
// `keep` bytes of sine sets remain or none is idle.
void trim_sine(EmbDevice *dev, size_t keep);

// Bytes a scene of this shape needs (arena only; shared tables excluded).
size_t scene_bytes(const EmbSettings &s, int cols, int rows, int64_t n, int n_stacks, int n_stack_cells, int n_bats, int n_ac);

// Allocate the arena, fill the descriptor and upload the prepared scene on `stream`.
int make_resident(EmbDevice *dev, const EmbSettings &s, HostScene &&hs, hipStream_t stream, Resident &out);
void free_resident(Resident &r);

// The message of a scene's device-side numeric flags (SceneCounters::rest_error,
// charge_overflow), empty when neither is set.
inline std::string numeric_fault(int32_t rest_error, int32_t charge_overflow) {
    if (rest_error) return "electron substeps after the first were dropped: k_rest wait limit";
    if (charge_overflow) return "a node's charge passed the fixed-point range of the smoothing (too many electrons in one place)";
    return std::string();
}

// Electron positions (n x 2, FP64) of a resident scene; synchronous.
int read_positions(const DevScene &d, double *pos);

// Blocks per kernel class for one scene.
void scene_blocks(const DevScene &d, int out[KC_COUNT]);

// Kernel sequences over a batch (kernels.hip).
hipError_t sort_kernels(const Batch &b, hipStream_t stream);
hipError_t setup_kernels(const Batch &b, hipStream_t stream);   // sort, deposit, Gauss (no FDTD)
constexpr int TICK_PHASES = 11;
extern const char *const TICK_PHASE_NAMES[];
hipError_t tick_kernels(const Batch &b, int fdtd_substeps, hipStream_t stream, hipEvent_t *marks = nullptr);
hipError_t deposit_kernel(const Batch &b, hipStream_t stream);
hipError_t close_force_kernel(const Batch &b, float2 *out, hipStream_t stream);
hipError_t finish_sync_kernel(const Batch &b, hipStream_t stream);
hipError_t diagnostics_kernel(const DevScene *scene, double c2, double *out, int blocks, hipStream_t stream);
hipError_t probe_kernel(const DevScene *scene, int cell, int stride, float c2, float inv_mu0, float *out, unsigned *hist,
                        hipStream_t stream);

// Write the per-tick AC phases and EMF targets of ticks [r.tick, r.tick + count)
// into the pinned staging and enqueue the copy. Advances r.time and r.tick.
int enqueue_tick_tables(Resident &r, int count, hipStream_t stream);

}  // namespace emb
