// Host runtime of the projector library: error reporting, a per-(device,
// stream) grow-only workspace arena, and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>

namespace ctp {

// Error raised inside the library; converted to a non-zero return code at the
// C ABI, never propagated as an exception or an abort.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check(hipError_t e, const char* what);
#define CTP_CHECK(expr) ::ctp::check((expr), #expr)
#define CTP_LAUNCH_CHECK() ::ctp::check(hipGetLastError(), "kernel launch")

// Validation failure: invalid argument.
[[noreturn]] void fail(const std::string& msg);

// Thread-local message of the last failure (ctp_last_error()).
void set_last_error(const std::string& msg);
const char* last_error();

// Device scratch memory owned by the library.  One arena per (device, stream,
// host thread), so that calls on different streams never share scratch and
// threads sharing a stream never share an arena's host-side state; calls into
// one arena come from one thread and are ordered by the stream.  Grows on demand (the old block is
// released after the stream drains) and is kept until ctp_release_workspace().
class Workspace {
public:
    explicit Workspace(hipStream_t s);
    // Returns a pointer to at least 'bytes' bytes, 256-byte aligned, valid until
    // the next call on the same stream.  Sub-allocations are carved with take().
    void reset();
    void* take(size_t bytes);
    template <typename T> T* take_n(size_t n) { return static_cast<T*>(take(n * sizeof(T))); }
    void reserve(size_t bytes);
    size_t capacity() const { return cap_; }

private:
    hipStream_t stream_;
    char* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    friend Workspace& workspace_for(hipStream_t);
    friend void release_all_workspaces();
};

Workspace& workspace_for(hipStream_t stream);
void release_all_workspaces();

// Size planner: callers first add up what they need, then reserve.
struct Plan {
    size_t bytes = 0;
    template <typename T> void add(size_t n) { bytes += ((n * sizeof(T) + 255) / 256) * 256; }
};

// Host staging for the reference's host-pointer (numpy) entry points.
struct DeviceBuffer {
    void* p = nullptr;
    size_t n = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    template <typename T> T* as() { return static_cast<T*>(p); }
};

// Launch-shape knobs (tile shapes, run lengths) read by the kernels' host
// code; set by the autotuner or by the caller through ctpSetTuning().  Unknown
// keys are rejected.  Value 0 means "use the built-in default".
int tuning_get(const char* key, int dflt);
bool tuning_set(const char* key, int value);
// Incremented by every tuning_set: tables built from tuning values record it.
long tuning_generation();

// Per-kernel timing with HIP events on the launching stream (off by default).
// Launch sites wrap their kernel with TimedLaunch; the C ABI collects the
// elapsed times after each call has synchronised.
void timing_enable(bool on);
bool timing_enabled();
void timing_reset();
void timing_collect();                       // after the stream(s) drained
bool timing_get(const char* name, double* total_ms, long* count);
std::string timing_report();

// While an instance lives, TimedLaunch records nothing on the calling thread
// (the allocation-free entry point creates no HIP events).
class TimingSuspended {
public:
    TimingSuspended();
    ~TimingSuspended();
private:
    bool prev_;
};

class TimedLaunch {
public:
    TimedLaunch(const char* name, hipStream_t s);
    ~TimedLaunch();
private:
    const char* name_;
    hipStream_t s_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

}  // namespace ctp
