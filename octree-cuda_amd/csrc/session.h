// Session: the object behind the void* handle of the C ABI (OctreeApi/src/session.h:13-58).
// Holds the cloud description and processing configuration set through the ocpi_* setters,
// validates them at init, owns the OctreeBuilder, and runs the exports.
#pragma once

#include <chrono>
#include <cstdint>
#include <memory>
#include <string>

#include "build_request.h"
#include "builder.h"

namespace oc {

class Session final : public BuildRequest
{
public:
    explicit Session(int device);
    ~Session();
    Session(const Session&) = delete;
    Session& operator=(const Session&) = delete;

    static Session* fromHandle(void* handle);

    void setPointCloudHost(const uint8_t* cloud);
    void setPointCloudDevice(const uint8_t* cloud);
    void setCloudType(uint8_t type);
    void setCloudPointAmount(uint32_t n);
    void setCloudDataStride(uint32_t stride);
    void setCloudScale(double x, double y, double z);
    void setCloudOffset(double x, double y, double z);
    void setCloudBoundingBox(double minX, double minY, double minZ, double maxX, double maxY, double maxZ);
    void configureChunking(uint32_t chunkingGrid, uint32_t mergingThreshold, float outputFactor);
    void configureSubsampling(uint32_t subsamplingGrid, bool intra, bool inter, bool replacement, bool random,
                              bool weighted);
    void setRandomSeed(uint64_t seed);

    void initOctree();
    void generateOctree();

    void exportPotree(const std::string& directory);
    void exportJsonReport(const std::string& filename);
    void exportMemoryReport(const std::string& filename);
    void exportDistributionHistogram(const std::string& filename, uint32_t binWidth);

    // The last finished build: the session's own builder, or the snapshot of a build the
    // BuildService ran for it.
    const BuildResult& result() const;
    const HostOctree& hostOctree();  // lazily copied after each build

    // Build service hooks (BuildRequest). A submitted session is owned by the service until its
    // wait returns.
    void validate() const override;
    const CloudInfo& cloudInfo() const override { return itsCloud; }
    const ProcessingInfo& processingInfo() const override { return itsProc; }
    const uint8_t* hostCloud() const override { return itsHostCloud; }
    const uint8_t* deviceCloud() const override { return itsDeviceCloud; }
    void adoptResult(std::unique_ptr<HostBuildResult> result) override;

private:

    int itsDevice;
    const uint8_t* itsHostCloud   = nullptr;
    const uint8_t* itsDeviceCloud = nullptr;
    CloudInfo itsCloud;
    ProcessingInfo itsProc;
    std::unique_ptr<OctreeBuilder> itsBuilder;
    std::unique_ptr<HostBuildResult> itsSnapshot;
    const BuildResult* itsResult = nullptr;
    std::unique_ptr<HostOctree> itsHost;
    uint64_t itsBuildSerial     = 0;
    uint64_t itsHostSerial      = ~0ull;
    std::string itsMemoryReport;
};

}  // namespace oc
