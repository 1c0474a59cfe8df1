// The C ABI (include/octreeApi.h). Every entry point runs inside execute(), which maps any
// exception to UNEXPECTED_ERROR (-1) like the reference's Execute wrapper (octreeApi.cpp:9-21)
// and keeps the message for ocpi_get_last_error; HIP failures never terminate the process.
#include "octreeApi.h"

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <cstring>
#include <exception>
#include <string>

#include "batch_builder.h"
#include "bounds.h"
#include "log.h"
#include "reports.h"
#include "potree_writer.h"
#include "service.h"
#include "session.h"

namespace {

thread_local std::string tLastError;

template <typename F>
int execute(F&& f)
{
    try
    {
        f();
        tLastError.clear();
        return SUCCESS;
    }
    catch (const std::exception& e)
    {
        tLastError = e.what();
    }
    catch (...)
    {
        tLastError = "unknown error";
    }
    oc::log(oc::LOG_ERROR, tLastError);
    return UNEXPECTED_ERROR;
}

oc::Session* S(void* h)
{
    return oc::Session::fromHandle(h);
}

}  // namespace

int ocpi_set_logging_level(int level)
{
    return execute([&] {
        switch (level)  // octreeApi.cpp:26-40
        {
        case 1: oc::setLogLevel(oc::LOG_INFO); break;
        case 2: oc::setLogLevel(oc::LOG_WARN); break;
        case 3: oc::setLogLevel(oc::LOG_ERROR); break;
        default: oc::setLogLevel(oc::LOG_DEBUG); break;
        }
    });
}

int ocpi_create_session(void** session, int device)
{
    return execute([&] {
        if (!session)
            throw std::runtime_error("null session pointer");
        *session = nullptr;
        *session = new oc::Session(device);
    });
}

int ocpi_destroy_session(void* session)
{
    return execute([&] { delete S(session); });
}

int ocpi_set_point_cloud_host(void* session, uint8_t* pointCloud)
{
    return execute([&] { S(session)->setPointCloudHost(pointCloud); });
}

int ocpi_set_point_cloud_device(void* session, uint8_t* devicePointCloud)
{
    return execute([&] { S(session)->setPointCloudDevice(devicePointCloud); });
}

int ocpi_init_octree(void* session)
{
    return execute([&] { S(session)->initOctree(); });
}

int ocpi_generate_octree(void* session)
{
    return execute([&] { S(session)->generateOctree(); });
}

int ocpi_export_potree(void* session, const char* directory)
{
    return execute([&] {
        if (!directory)
            throw std::runtime_error("null directory");
        S(session)->exportPotree(directory);
    });
}

int ocpi_configure_chunking(void* session, uint32_t chunkingGrid, uint32_t mergingThreshold, float outputFactor)
{
    return execute([&] { S(session)->configureChunking(chunkingGrid, mergingThreshold, outputFactor); });
}

int ocpi_configure_subsampling(void* session, uint32_t subsamplingGrid, bool intraCellAveraging,
                               bool interCellAveraging, bool replacementScheme, bool useRandomSubsampling,
                               bool useWeightingFunction)
{
    return execute([&] {
        S(session)->configureSubsampling(subsamplingGrid, intraCellAveraging, interCellAveraging, replacementScheme,
                                         useRandomSubsampling, useWeightingFunction);
    });
}

int ocpi_export_memory_report(void* session, const char* filename)
{
    return execute([&] {
        if (!filename)
            throw std::runtime_error("null filename");
        S(session)->exportMemoryReport(filename);
    });
}

int ocpi_export_json_report(void* session, const char* filename)
{
    return execute([&] {
        if (!filename)
            throw std::runtime_error("null filename");
        S(session)->exportJsonReport(filename);
    });
}

int ocpi_export_distribution_histogram(void* session, const char* filename, uint32_t binWidth)
{
    return execute([&] {
        if (!filename)
            throw std::runtime_error("null filename");
        S(session)->exportDistributionHistogram(filename, binWidth);
    });
}

int ocpi_set_cloud_type(void* session, uint8_t cloudType)
{
    return execute([&] { S(session)->setCloudType(cloudType); });
}

int ocpi_set_cloud_point_amount(void* session, uint32_t pointAmount)
{
    return execute([&] { S(session)->setCloudPointAmount(pointAmount); });
}

int ocpi_set_cloud_data_stride(void* session, uint32_t dataStride)
{
    return execute([&] { S(session)->setCloudDataStride(dataStride); });
}

int ocpi_set_cloud_scale(void* session, double x, double y, double z)
{
    return execute([&] { S(session)->setCloudScale(x, y, z); });
}

int ocpi_set_cloud_offset(void* session, double x, double y, double z)
{
    return execute([&] { S(session)->setCloudOffset(x, y, z); });
}

int ocpi_set_cloud_bb(void* session, double minX, double minY, double minZ, double maxX, double maxY, double maxZ)
{
    return execute([&] { S(session)->setCloudBoundingBox(minX, minY, minZ, maxX, maxY, maxZ); });
}

int ocpi_set_random_seed(void* session, uint64_t seed)
{
    return execute([&] { S(session)->setRandomSeed(seed); });
}

int ocpi_get_last_error(char* buffer, uint32_t bufferSize)
{
    const std::string& e = tLastError;
    if (buffer && bufferSize > 0)
    {
        size_t n = std::min<size_t>(e.size(), bufferSize - 1);
        std::memcpy(buffer, e.data(), n);
        buffer[n] = '\0';
    }
    return static_cast<int>(e.size());
}

int ocpi_get_build_summary(void* session, uint64_t* summary, uint32_t n)
{
    return execute([&] {
        oc::Session* s          = S(session);
        const oc::BuildResult& b = s->result();
        const oc::HostOctree& t = s->hostOctree();
        oc::OctreeInfo info     = oc::nodeStatistics(b, t);
        uint64_t exported       = 0;
        uint64_t nodes          = oc::hierarchyEntries(t, &exported).size();
        uint64_t v[9] = {b.sparseNodes(), b.cloud().pointAmount, b.capacity(), b.used(),
                         info.internalNodeAmount, info.leafNodeAmount, b.depth(), b.denseNodes(), nodes};
        for (uint32_t i = 0; i < n && i < 9; ++i)
            summary[i] = v[i];
    });
}

int ocpi_copy_nodes(void* session, uint32_t* denseKey, uint32_t* denseCount, uint32_t* pointCount, uint64_t* dataIdx,
                    int32_t* parent, int32_t* children, uint8_t* flags)
{
    return execute([&] {
        const oc::HostOctree& t = S(session)->hostOctree();
        const size_t m          = t.flags.size();
        if (denseKey)
            std::copy(t.denseKey.begin(), t.denseKey.end(), denseKey);
        if (denseCount)
            std::copy(t.count.begin(), t.count.end(), denseCount);
        if (pointCount)
            std::copy(t.pointCount.begin(), t.pointCount.end(), pointCount);
        if (dataIdx)
            std::copy(t.dataIdx.begin(), t.dataIdx.end(), dataIdx);
        if (parent)
            std::copy(t.parent.begin(), t.parent.end(), parent);
        if (children)
            std::copy(t.children.begin(), t.children.begin() + 8 * m, children);
        if (flags)
            std::copy(t.flags.begin(), t.flags.end(), flags);
    });
}

int ocpi_copy_point_lut(void* session, uint32_t* destination, uint64_t count)
{
    return execute([&] { S(session)->result().copyLut(destination, count); });
}

int ocpi_copy_output(void* session, uint8_t* destination, uint64_t count)
{
    return execute([&] { S(session)->result().copyOutput(destination, count); });
}

int ocpi_get_stage_timings(void* session, float* milliseconds, uint32_t n)
{
    return execute([&] {
        const auto& t = S(session)->result().stageTimings();
        for (uint32_t i = 0; i < n; ++i)
            milliseconds[i] = i < t.size() ? t[i] : 0.f;
    });
}

// ---- build service (continuous batching) ------------------------------------------------------

namespace {

oc::ServiceBase* V(void* h)
{
    return oc::ServiceBase::fromHandle(h);
}

}  // namespace

int ocpi_create_service(void** service, int device, uint32_t slots, uint64_t memoryBudget)
{
    return execute([&] {
        if (!service)
            throw std::runtime_error("service handle pointer is null");
        *service = static_cast<oc::ServiceBase*>(new oc::BuildService(device, slots, memoryBudget));
    });
}

int ocpi_create_batch_service(void** service, int device, uint32_t workers, uint32_t maxBatch, uint64_t memoryBudget)
{
    return execute([&] {
        if (!service)
            throw std::runtime_error("service handle pointer is null");
        *service = static_cast<oc::ServiceBase*>(new oc::BatchService(device, workers, maxBatch, memoryBudget));
    });
}

int ocpi_destroy_service(void* service)
{
    return execute([&] { delete V(service); });
}

int ocpi_service_submit(void* service, void* session)
{
    return execute([&] { V(service)->submit(S(session)); });
}

int ocpi_service_wait(void* service, void* session, uint32_t timeoutMs)
{
    return execute([&] {
        if (!V(service)->wait(S(session), timeoutMs))
            throw std::runtime_error("timed out waiting for the build");
    });
}

int ocpi_service_set_active_slots(void* service, uint32_t slots)
{
    return execute([&] { V(service)->setActiveSlots(slots); });
}

int ocpi_service_get_stats(void* service, double* stats, uint32_t n)
{
    return execute([&] {
        const std::vector<double> v = V(service)->stats();
        for (uint32_t i = 0; i < n; ++i)
            stats[i] = i < v.size() ? v[i] : 0.0;
    });
}

int ocpi_service_reset_stats(void* service)
{
    return execute([&] { V(service)->resetStats(); });
}

int ocpi_get_request_timing(void* session, double* milliseconds, uint32_t n)
{
    return execute([&] {
        auto& r = S(session)->request();
        if (r.state != oc::Session::Request::DONE && r.state != oc::Session::Request::FAILED)
            throw std::runtime_error("the session has no finished service request");
        using ms     = std::chrono::duration<double, std::milli>;
        double v[3] = {ms(r.started - r.submitted).count(), ms(r.finished - r.started).count(),
                       ms(r.finished - r.submitted).count()};
        for (uint32_t i = 0; i < n; ++i)
            milliseconds[i] = i < 3 ? v[i] : 0.0;
    });
}

int ocpi_estimate_footprint(void* session, uint64_t* bytes)
{
    return execute([&] {
        oc::Session* s = S(session);
        s->validate();
        *bytes = oc::OctreeBuilder::estimateFootprint(s->cloudInfo(), s->processingInfo(), s->hostCloud() != nullptr);
    });
}

int ocpi_compute_bounding_box(const uint8_t* devicePointCloud, uint32_t pointAmount, uint32_t dataStride,
                              uint8_t cloudType, int device, double* realBB, double* cubicBB)
{
    return execute([&] {
        if (!realBB || !cubicBB)
            throw std::runtime_error("bounding box output pointers are null");
        oc::realBoundingBox(devicePointCloud, pointAmount, dataStride, cloudType, device, realBB);
        oc::cubicBoundingBox(realBB, cubicBB);
    });
}

int ocpi_compute_bounding_box_noalloc(const uint8_t* devicePointCloud, uint32_t pointAmount, uint32_t dataStride,
                                      uint8_t cloudType, int device, double* realBB, double* cubicBB)
{
    return execute([&] {
        if (!realBB || !cubicBB)
            throw std::runtime_error("bounding box output pointers are null");
        oc::realBoundingBoxNoAlloc(devicePointCloud, pointAmount, dataStride, cloudType, device, realBB);
        oc::cubicBoundingBox(realBB, cubicBB);
    });
}

// ---- batched builds -------------------------------------------------------------------------

namespace {

oc::BatchBuilder* BB(void* h)
{
    if (!h)
        throw std::runtime_error("No BatchBuilder is currently initialized!");
    return static_cast<oc::BatchBuilder*>(h);
}

}  // namespace

int ocpi_create_batch_builder(void** builder, int device)
{
    return execute([&] {
        if (!builder)
            throw std::runtime_error("batch builder handle pointer is null");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
            throw std::runtime_error("device " + std::to_string(device) + " does not exist");
        *builder = new oc::BatchBuilder(device);
    });
}

int ocpi_destroy_batch_builder(void* builder)
{
    return execute([&] { delete BB(builder); });
}

int ocpi_batch_build(void* builder, void** sessions, uint32_t count, int32_t* status)
{
    return execute([&] {
        oc::BatchBuilder* b = BB(builder);
        if (count > 0 && !sessions)
            throw std::runtime_error("session array is null");
        std::vector<oc::Session*> list(count);
        std::vector<oc::BatchRequest> requests(count);
        for (uint32_t i = 0; i < count; ++i)
        {
            list[i] = S(sessions[i]);
            list[i]->validate();
            requests[i].cloud       = list[i]->cloudInfo();
            requests[i].proc        = list[i]->processingInfo();
            requests[i].hostCloud   = list[i]->hostCloud();
            requests[i].deviceCloud = list[i]->deviceCloud();
        }
        std::vector<oc::BatchOutcome> outcomes = b->build(requests);
        std::string firstError;
        for (uint32_t i = 0; i < count; ++i)
        {
            const bool ok = static_cast<bool>(outcomes[i].result);
            if (ok)
                list[i]->adoptResult(std::move(outcomes[i].result));
            else if (firstError.empty())
                firstError = "session " + std::to_string(i) + ": " + outcomes[i].error;
            if (status)
                status[i] = ok ? SUCCESS : UNEXPECTED_ERROR;
        }
        if (!firstError.empty())
            throw std::runtime_error(firstError);
    });
}

int ocpi_batch_set_launch(void* builder, uint32_t threadGrid, uint32_t tileGrid, uint32_t sortGrid)
{
    return execute([&] {
        if (threadGrid == 0 || tileGrid == 0 || sortGrid == 0 || threadGrid > 65536 || tileGrid > 65536 ||
            sortGrid > 65536)
            throw std::runtime_error("launch sizes must be in [1, 65536] workgroups");
        BB(builder)->setLaunch(oc::BatchLaunch{threadGrid, tileGrid, sortGrid});
    });
}

int ocpi_batch_use_graphs(void* builder, uint32_t on)
{
    return execute([&] { BB(builder)->useGraphs(on != 0); });
}

int ocpi_batch_get_timings(void* builder, float* milliseconds, uint32_t n)
{
    return execute([&] {
        const std::vector<float>& t = BB(builder)->timings();
        for (uint32_t i = 0; i < n && i < t.size(); ++i)
            milliseconds[i] = t[i];
    });
}

int ocpi_estimate_batch_footprint(void** sessions, uint32_t count, uint64_t* bytes)
{
    return execute([&] {
        if (!bytes || (count > 0 && !sessions))
            throw std::runtime_error("null argument");
        std::vector<oc::BatchRequest> requests(count);
        for (uint32_t i = 0; i < count; ++i)
        {
            oc::Session* s = S(sessions[i]);
            s->validate();
            requests[i].cloud       = s->cloudInfo();
            requests[i].proc        = s->processingInfo();
            requests[i].hostCloud   = s->hostCloud();
            requests[i].deviceCloud = s->deviceCloud();
        }
        *bytes = oc::BatchBuilder::estimateFootprint(requests);
    });
}
