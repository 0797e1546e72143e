#include "runtime.h"

#include <cstdio>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>
#include <sstream>

namespace ctp {

namespace {
struct Pending {
    std::string name;
    hipEvent_t a, b;
};
struct Stat {
    double ms = 0;
    long n = 0;
};
std::mutex g_t_mutex;
bool g_t_on = false;
std::vector<Pending> g_pending;
std::map<std::string, Stat> g_stats;
thread_local std::string g_last_error;
std::mutex g_tune_mutex;
std::map<std::string, int> g_tune = {
    {"cone_fp_tile_u", 0},   // detector columns per wavefront tile of the Siddon FP (8, 16, 32, 64)
    {"cone_plan_bp_launch", 0},  // cone plan backward: 1 one persistent launch for every window shape (default), 2 one launch per shape
    {"cone_bp_zrun", 0},     // voxels along z per thread of the Siddon BP (1, 2, 4, 8; default 8 for one volume with a 2x2 window, else 4)
    {"cone_fp_order", 0},    // Siddon FP block order: 1 plain view-major, 2 XCD bands (default per form)
    {"cone_fp_views_per_block", 0},   // column-driven Siddon FP: views per block (1, 2, 4; default 2)
    {"cone_fp_mode", 0},     // Siddon FP: 0 column-driven form (rays of a detector column together), 1 ray per lane
    {"cone_bp_mode", 0},     // Siddon BP: 0 column mode where the geometry allows it, 1 general window, 2 ray-driven scatter
    {"cone_slab_slices", 0}, // Siddon FP/BP: at most this many z slices per slab (0: slabs only for 2^29+ voxels)
    {"cone_plan_persistent", 0},   // cone plans: 1 persistent blocks with work counters, 2 one block per tile
    {"dd_batch_persistent", 0},    // DD batch plans: 1 persistent blocks with work counters, 2 one block per tile (default)
    {"dd_batch_streams", 0},       // DD batch plans: streams the kernel classes of a direction run on (1 to 4; default 4)
    {"dd_fp_mode", 0},       // DD FP: 1 direct (default), 2 LDS-staged rows where the footprint fits
    {"dd_fp_extra_channel", 0},  // DD FP direct kernel: 1 keeps the next chunk's channel with one chunk too
    {"dd_fp_pixel_stride", 0},   // DD FP direct kernel, one channel chunk: floats per pixel of the channel-last copy
    {"dd_fp_views_per_block", 0},  // DD FP direct kernel, one row segment: views per block (1, 2, 4; default 4)
    {"dd_fp_lanes_per_column", 0}, // DD FP: lanes per detector column (1 = whole channel vector per lane; 2, 4, 8; default CH / 8 from 16 channels on, else 1)
    {"dd_bp_lanes_per_pixel", 0},  // DD BP (pixel driven): lanes per pixel (1 = whole channel vector per lane; 2, 4; default 4 at 32 channels, else 2)
    {"fbp_bp_lanes_per_pixel", 0}, // fan FBP / parallel pixel BP: lanes per pixel (1 = whole channel vector per lane; 2, 4; default 4, 2 at 8 channels)
    {"fbp_bp_stage", 0},     // fan FBP / parallel pixel BP, split form: cells staged in LDS per wavefront: 1 always, 2 never (default: at 32 channels with 4 lanes per pixel)
    {"dd_fp_segments", 0},   // DD FP direct kernel: row segments per column (1, 2; default by wave count)
    {"tv_mode", 0},          // TV prior: 1 z-marching tile kernels (default), 2 one lane per voxel
    {"ramp_mode", 0},        // ramp filter: 0 FFT convolution where 2nu-1 <= 4096 (default), 1 direct, 2 generic FFT kernel
    {"bp_view_groups", 0},   // view groups of the pixel-driven DD / FBP backprojections (default: enough for 1024 blocks)
};
// Arenas are keyed by (device, stream, host thread): two threads calling on the
// same stream (the null stream is every thread's default) never carve the
// same scratch; within one arena calls are ordered by their thread.
using WsKey = std::tuple<int, hipStream_t, std::thread::id>;
std::mutex g_ws_mutex;
std::map<WsKey, std::unique_ptr<Workspace>> g_ws;
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error((int)e, std::string(what) + ": " + hipGetErrorString(e));
}

void fail(const std::string& msg) { throw Error((int)hipErrorInvalidValue, msg); }

Workspace::Workspace(hipStream_t s) : stream_(s) {}

void Workspace::reset() { used_ = 0; }

void Workspace::reserve(size_t bytes) {
    used_ = 0;
    if (bytes <= cap_) return;
    if (base_) {
        // the previous block may still be read by queued kernels on this stream
        CTP_CHECK(hipStreamSynchronize(stream_));
        CTP_CHECK(hipFree(base_));
        base_ = nullptr;
        cap_ = 0;
    }
    size_t want = bytes + bytes / 8;   // some head-room for slightly larger calls
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        e = hipMalloc(&p, bytes);
        want = bytes;
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error((int)e, "workspace allocation of " + std::to_string(bytes) + " bytes failed: " +
                                hipGetErrorString(e));
    }
    base_ = static_cast<char*>(p);
    cap_ = want;
}

void* Workspace::take(size_t bytes) {
    size_t n = ((bytes + 255) / 256) * 256;
    if (used_ + n > cap_) throw Error((int)hipErrorOutOfMemory, "workspace overrun (internal planning error)");
    void* p = base_ + used_;
    used_ += n;
    return p;
}

Workspace& workspace_for(hipStream_t stream) {
    int dev = 0;
    CTP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    auto key = WsKey(dev, stream, std::this_thread::get_id());
    auto it = g_ws.find(key);
    if (it == g_ws.end()) it = g_ws.emplace(key, std::make_unique<Workspace>(stream)).first;
    return *it->second;
}

void release_all_workspaces() {
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto& kv : g_ws) {
        Workspace& w = *kv.second;
        if (w.base_) {
            (void)hipSetDevice(std::get<0>(kv.first));
            (void)hipStreamSynchronize(w.stream_);
            (void)hipFree(w.base_);
        }
    }
    g_ws.clear();
    (void)hipSetDevice(cur);
}

DeviceBuffer::DeviceBuffer(size_t bytes) : n(bytes) {
    if (bytes == 0) return;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        throw Error((int)e, "device allocation of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
}

DeviceBuffer::~DeviceBuffer() {
    if (p) (void)hipFree(p);
}

int tuning_get(const char* key, int dflt) {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    auto it = g_tune.find(key);
    return (it == g_tune.end() || it->second == 0) ? dflt : it->second;
}
std::atomic<long> g_tune_gen{0};

bool tuning_set(const char* key, int value) {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    it->second = value;
    g_tune_gen.fetch_add(1);
    return true;
}

long tuning_generation() { return g_tune_gen.load(); }

void timing_enable(bool on) {
    std::lock_guard<std::mutex> lk(g_t_mutex);
    g_t_on = on;
}
bool timing_enabled() {
    std::lock_guard<std::mutex> lk(g_t_mutex);
    return g_t_on;
}
void timing_reset() {
    std::lock_guard<std::mutex> lk(g_t_mutex);
    g_stats.clear();
}
void timing_collect() {
    std::lock_guard<std::mutex> lk(g_t_mutex);
    for (auto& p : g_pending) {
        float ms = 0;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            Stat& st = g_stats[p.name];
            st.ms += ms;
            st.n += 1;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    g_pending.clear();
    (void)hipGetLastError();
}
bool timing_get(const char* name, double* total_ms, long* count) {
    std::lock_guard<std::mutex> lk(g_t_mutex);
    auto it = g_stats.find(name);
    if (it == g_stats.end()) return false;
    if (total_ms) *total_ms = it->second.ms;
    if (count) *count = it->second.n;
    return true;
}
std::string timing_report() {
    std::lock_guard<std::mutex> lk(g_t_mutex);
    std::ostringstream o;
    for (auto& kv : g_stats) o << kv.first << " " << kv.second.ms << " " << kv.second.n << "\n";
    return o.str();
}

namespace {
thread_local bool t_timing_suspended = false;
}
TimingSuspended::TimingSuspended() : prev_(t_timing_suspended) { t_timing_suspended = true; }
TimingSuspended::~TimingSuspended() { t_timing_suspended = prev_; }

TimedLaunch::TimedLaunch(const char* name, hipStream_t s) : name_(name), s_(s) {
    if (t_timing_suspended || !timing_enabled()) return;
    // no event records into a stream that is being captured into a graph
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        return;
    }
    if (hipEventCreate(&a_) != hipSuccess || hipEventCreate(&b_) != hipSuccess) {
        (void)hipGetLastError();
        a_ = b_ = nullptr;
        return;
    }
    (void)hipEventRecord(a_, s_);
}
TimedLaunch::~TimedLaunch() {
    if (!a_) return;
    (void)hipEventRecord(b_, s_);
    std::lock_guard<std::mutex> lk(g_t_mutex);
    g_pending.push_back(Pending{name_, a_, b_});
}

}  // namespace ctp
