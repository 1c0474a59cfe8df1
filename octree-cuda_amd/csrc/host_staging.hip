#include "host_staging.h"

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_copy.h"
#include "slot_builder.h"

namespace oc {

namespace {

constexpr uint64_t STAGE_BYTES    = 32ull << 20;
constexpr uint64_t STAGE_MIN_COPY = 8ull << 20;  // smaller transfers go straight to their destination

void check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// the two buffers, allocated on first use and kept until the process ends; one transfer at a time
struct HostStaging
{
    std::mutex mutex;
    uint8_t* buffer[2] = {nullptr, nullptr};
    void allocate()
    {
        for (auto& b : buffer)
            if (!b)
                b = static_cast<uint8_t*>(pinnedAllocate(STAGE_BYTES));
    }
};

HostStaging& hostStaging()
{
    static HostStaging s;
    return s;
}

// true when p is page-locked host memory the copy engines can reach directly
bool pageLocked(const void* p)
{
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, p) != hipSuccess)
    {
        (void)hipGetLastError();  // pageable memory: the query fails; clear it
        return false;
    }
    return attr.type == hipMemoryTypeHost;
}

// two events for the two buffers, destroyed with the transfer
struct EventPair
{
    hipEvent_t e[2] = {};
    EventPair()
    {
        for (auto& x : e)
            check(hipEventCreateWithFlags(&x, hipEventDisableTiming), "hipEventCreate");
    }
    ~EventPair()
    {
        for (auto& x : e)
            (void)hipEventDestroy(x);
    }
};

}  // namespace

void uploadSegments(const HostSegment* segments, size_t count, hipStream_t stream)
{
    std::vector<HostSegment> staged;
    uint64_t stagedBytes = 0;
    for (size_t i = 0; i < count; ++i)
    {
        if (segments[i].bytes == 0)
            continue;
        if (pageLocked(segments[i].host))
            check(hipMemcpyAsync(segments[i].device, segments[i].host, segments[i].bytes, hipMemcpyHostToDevice,
                                 stream),
                  "upload");
        else
        {
            staged.push_back(segments[i]);
            stagedBytes += segments[i].bytes;
        }
    }
    if (stagedBytes < STAGE_MIN_COPY)
    {
        for (const HostSegment& s : staged)
            check(hipMemcpyAsync(s.device, s.host, s.bytes, hipMemcpyHostToDevice, stream), "upload");
        return;
    }
    HostStaging& st = hostStaging();
    std::lock_guard<std::mutex> lock(st.mutex);
    st.allocate();
    EventPair drained;
    hipError_t failure = hipSuccess;
    size_t seg         = 0;
    uint64_t segOff    = 0;
    std::vector<CopyPiece> pieces;
    for (uint64_t k = 0; seg < staged.size() && failure == hipSuccess; ++k)
    {
        // the buffer of chunk k - 2 must have been copied out before chunk k fills it
        if (k >= 2)
            failure = hipEventSynchronize(drained.e[k & 1]);
        if (failure != hipSuccess)
            break;
        uint8_t* buf  = st.buffer[k & 1];
        uint64_t fill = 0;
        pieces.clear();
        const size_t firstSeg = seg;
        const uint64_t firstOff = segOff;
        while (seg < staged.size() && fill < STAGE_BYTES)
        {
            const uint64_t len = std::min(STAGE_BYTES - fill, staged[seg].bytes - segOff);
            pieces.push_back({buf + fill, static_cast<const uint8_t*>(staged[seg].host) + segOff, len});
            fill += len;
            segOff += len;
            if (segOff == staged[seg].bytes)
            {
                ++seg;
                segOff = 0;
            }
        }
        parallelMemcpy(pieces.data(), pieces.size());
        // the same pieces, from the buffer to their devices
        size_t s = firstSeg;
        uint64_t o = firstOff, at = 0;
        for (const CopyPiece& p : pieces)
        {
            if (failure == hipSuccess)
                failure = hipMemcpyAsync(static_cast<uint8_t*>(staged[s].device) + o, buf + at, p.bytes,
                                         hipMemcpyHostToDevice, stream);
            at += p.bytes;
            o += p.bytes;
            if (o == staged[s].bytes)
            {
                ++s;
                o = 0;
            }
        }
        if (failure == hipSuccess)
            failure = hipEventRecord(drained.e[k & 1], stream);
    }
    // the buffers are shared: no copy may read them once the lock is released
    const hipError_t drain = hipStreamSynchronize(stream);
    check(failure, "staged upload");
    check(drain, "staged upload");
}

void downloadStaged(void* host, const void* device, uint64_t bytes, hipStream_t stream)
{
    if (bytes < STAGE_MIN_COPY || pageLocked(host))
    {
        check(hipMemcpyAsync(host, device, bytes, hipMemcpyDeviceToHost, stream), "download");
        check(hipStreamSynchronize(stream), "download");
        return;
    }
    HostStaging& st = hostStaging();
    std::lock_guard<std::mutex> lock(st.mutex);
    st.allocate();
    EventPair landed;
    const auto* from   = static_cast<const uint8_t*>(device);
    auto* to           = static_cast<uint8_t*>(host);
    const uint64_t n   = (bytes + STAGE_BYTES - 1) / STAGE_BYTES;
    hipError_t failure = hipSuccess;
    auto issue         = [&](uint64_t k) {
        const uint64_t off = k * STAGE_BYTES, len = std::min(STAGE_BYTES, bytes - off);
        hipError_t e       = hipMemcpyAsync(st.buffer[k & 1], from + off, len, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess)
            e = hipEventRecord(landed.e[k & 1], stream);
        return e;
    };
    failure = issue(0);
    for (uint64_t k = 0; k < n && failure == hipSuccess; ++k)
    {
        // chunk k + 1 goes to the buffer chunk k - 1 was emptied from
        if (k + 1 < n)
            failure = issue(k + 1);
        if (failure == hipSuccess)
            failure = hipEventSynchronize(landed.e[k & 1]);
        if (failure == hipSuccess)
        {
            const uint64_t off = k * STAGE_BYTES;
            parallelMemcpy(to + off, st.buffer[k & 1], std::min(STAGE_BYTES, bytes - off));
        }
    }
    if (failure != hipSuccess)
        (void)hipStreamSynchronize(stream);  // no copy may still write a staging buffer
    check(failure, "staged download");
}

}  // namespace oc
