// Session implementation (OctreeApi/src/session.cpp:17-154), with configuration validation at
// init and no process exits.
#include "session.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>

#include "log.h"
#include "potree_writer.h"
#include "reports.h"
#include "service.h"

namespace oc {

Session* Session::fromHandle(void* handle)  // session.cpp:17-25
{
    auto* s = static_cast<Session*>(handle);
    if (!s)
        throw std::runtime_error("No Session is currently initialized!");
    return s;
}

Session::Session(int device) : itsDevice(device)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        throw std::runtime_error("no HIP device is available");
    if (device < 0 || device >= count)
        throw std::runtime_error("device " + std::to_string(device) + " does not exist");
    e = hipSetDevice(device);
    if (e != hipSuccess)
        throw std::runtime_error(std::string("hipSetDevice failed: ") + hipGetErrorString(e));
    hipDeviceProp_t props{};
    if (hipGetDeviceProperties(&props, device) == hipSuccess)
        log(LOG_INFO, std::string("Using GPU device: ") + props.name);
}

Session::~Session()
{
    ServiceBase* owner = itsRequest.owner.load();
    const Request state = itsRequest.state.load();
    if (owner && (state == Request::QUEUED || state == Request::RUNNING))
        owner->forget(this);
    if (!itsMemoryReport.empty())
    {
        try
        {
            writeMemoryReport(itsMemoryReport);
        }
        catch (...)
        {
        }
    }
    itsBuilder.reset();
}

void Session::setPointCloudHost(const uint8_t* cloud)
{
    itsHostCloud   = cloud;
    itsDeviceCloud = nullptr;
}

void Session::setPointCloudDevice(const uint8_t* cloud)
{
    itsDeviceCloud = cloud;
    itsHostCloud   = nullptr;
}

void Session::setCloudType(uint8_t type)
{
    itsCloud.cloudType = type;
}

void Session::setCloudPointAmount(uint32_t n)
{
    itsCloud.pointAmount = n;
}

void Session::setCloudDataStride(uint32_t stride)
{
    itsCloud.dataStride = stride;
}

void Session::setCloudScale(double x, double y, double z)
{
    itsCloud.scale[0] = x;
    itsCloud.scale[1] = y;
    itsCloud.scale[2] = z;
}

void Session::setCloudOffset(double x, double y, double z)
{
    itsCloud.offset[0] = x;
    itsCloud.offset[1] = y;
    itsCloud.offset[2] = z;
}

void Session::setCloudBoundingBox(double minX, double minY, double minZ, double maxX, double maxY, double maxZ)
{
    itsCloud.bbMin[0] = minX;
    itsCloud.bbMin[1] = minY;
    itsCloud.bbMin[2] = minZ;
    itsCloud.bbMax[0] = maxX;
    itsCloud.bbMax[1] = maxY;
    itsCloud.bbMax[2] = maxZ;
}

void Session::configureChunking(uint32_t chunkingGrid, uint32_t mergingThreshold, float outputFactor)
{
    itsProc.chunkingGrid     = chunkingGrid;
    itsProc.mergingThreshold = mergingThreshold;
    itsProc.outputFactor     = outputFactor;
}

void Session::configureSubsampling(uint32_t subsamplingGrid, bool intra, bool inter, bool replacement, bool random,
                                   bool weighted)
{
    itsProc.subsamplingGrid      = subsamplingGrid;
    itsProc.useIntraCellAvg      = intra;
    itsProc.useInterCellAvg      = inter;
    itsProc.useReplacementScheme = replacement;
    itsProc.useRandomSubsampling = random;
    itsProc.useWeightingFunction = weighted;
}

void Session::setRandomSeed(uint64_t seed)
{
    itsProc.seed = seed;
}

void Session::validate() const
{
    const auto& c = itsCloud;
    const auto& p = itsProc;
    if (!itsHostCloud && !itsDeviceCloud)
        throw std::runtime_error("no point cloud set (ocpi_set_point_cloud_host)");
    if (c.pointAmount == 0)
        throw std::runtime_error("the point amount is zero");
    if (c.cloudType > 1)
        throw std::runtime_error("cloud type must be 0 (float) or 1 (double)");
    uint32_t coordBytes = c.cloudType == 0 ? 4 : 8;
    if (c.dataStride < 3 * coordBytes + 3)
        throw std::runtime_error("data stride is smaller than three coordinates plus RGB");
    if (static_cast<uint64_t>(c.pointAmount) * c.dataStride > (1ull << 40))
        throw std::runtime_error("cloud too large");
    bool pow2 = false;
    for (uint32_t d = 1; d <= 9; ++d)
        pow2 |= p.chunkingGrid == (1u << d);
    if (!pow2)  // tools.cu:63-88 (any other grid gives depth 0)
        throw std::runtime_error("chunking grid must be a power of two in [2, 512]");
    if (p.subsamplingGrid == 0 || p.subsamplingGrid > 1024)
        throw std::runtime_error("subsampling grid must be in [1, 1024]");
    if (p.useIntraCellAvg && p.useWeightingFunction)  // octree_processor_impl.cu:59-71 vs rp_subsampling.cu:65-77
        throw std::runtime_error("the weighting function needs inter-cell filtering without intra-cell filtering");
    double side = c.bbMax[0] - c.bbMin[0];
    if (!(side > 0.0) || !std::isfinite(side))
        throw std::runtime_error("the cubic bounding box has no extent");
    if (!(c.scale[0] != 0.0 && c.scale[1] != 0.0 && c.scale[2] != 0.0))
        throw std::runtime_error("scale must be non-zero");
    uint64_t e = static_cast<uint64_t>(static_cast<float>(c.pointAmount) * p.outputFactor);
    if (!(p.outputFactor > 0.f) || e < c.pointAmount)
        throw std::runtime_error("output factor gives fewer output records than points");
    if (e >= 0xFFFFFFFFull)
        throw std::runtime_error("output capacity exceeds 2^32 records");
}

void Session::initOctree()
{
    validate();
    if (!itsBuilder)
        itsBuilder = std::make_unique<OctreeBuilder>(itsDevice);
    itsResult = nullptr;
    itsBuilder->setup(itsCloud, itsProc, itsHostCloud, itsDeviceCloud);
}

void Session::generateOctree()
{
    if (!itsBuilder)
        throw std::runtime_error("ocpi_init_octree has not been called");
    itsResult = nullptr;
    itsBuildSerial++;
    itsBuilder->generate();
    itsResult = itsBuilder.get();
}

void Session::adoptResult(std::unique_ptr<HostBuildResult> result)
{
    itsSnapshot = std::move(result);
    itsBuildSerial++;
    itsResult = itsSnapshot.get();
}

const BuildResult& Session::result() const
{
    if (!itsResult)
        throw std::runtime_error("the octree has not been generated");
    return *itsResult;
}

const HostOctree& Session::hostOctree()
{
    const BuildResult& b = result();
    if (!itsHost || itsHostSerial != itsBuildSerial)
    {
        itsHost = std::make_unique<HostOctree>();
        b.copyNodes(*itsHost);
        itsHostSerial = itsBuildSerial;
    }
    return *itsHost;
}

void Session::exportPotree(const std::string& directory)
{
    writePotree(directory, result(), hostOctree());
}

void Session::exportJsonReport(const std::string& filename)
{
    writeJsonReport(filename, result(), hostOctree());
}

void Session::exportMemoryReport(const std::string& filename)
{
    itsMemoryReport = filename;
}

void Session::exportDistributionHistogram(const std::string& filename, uint32_t binWidth)
{
    writeHistogram(filename, binWidth, result(), hostOctree());
}

}  // namespace oc
