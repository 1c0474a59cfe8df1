// octree-amd: the command-line caller of libOctreeApi.so, with the reference CLI's options and
// call sequence (src/main.cpp:24-109, src/argparser.h:24-160). The cloud file holds raw records
// (N x stride bytes, as the reference reads it); it is uploaded to HBM once, its bounding box is
// reduced on the GPU (ocpi_compute_bounding_box) and the build reads the device copy.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "octreeApi.h"

namespace {

struct Input  // src/argparser.h:7-23
{
    std::string inputFile, outputPath;
    uint32_t pointAmount = 0, pointDataStride = 0;
    float scale                   = 0.f;
    bool isIntraCellAveraging     = false;
    bool isInterCellAveraging     = false;
    bool useReplacementScheme     = true;
    bool performRandomSubsampling = false;
    bool useWeightingFunction     = false;
    uint32_t chunkingGrid = 0, subsamplingGrid = 0, mergingThreshold = 10000;
    uint8_t cloudType  = 0;
    float outputFactor = 2.2f;
    int device         = 0;
    uint64_t seed      = 0;
};

const char* kUsage =
        "Generates Potree compatible LOD structures on an MI355X\n"
        "usage: octree-amd -f FILE -o DIR -p POINTS -t float|double -d STRIDE,SCALE -g CHUNKING,SUBSAMPLING\n"
        "                  [-m MERGE_THRESHOLD=10000] [-e ESTIMATED_OUTPUT=2.2] [-a] [-i] [-w] [-r]\n"
        "                  [--device N] [--seed S] [-h]\n"
        "  -f, --file              point cloud file: raw records of 3 coordinates + RGB\n"
        "  -o, --output            output directory for the Potree data\n"
        "  -p, --points            point amount of the cloud\n"
        "  -t, --type              coordinate type: float or double\n"
        "  -d, --data              record stride and scale, e.g. 15,0.001\n"
        "  -g, --grids             chunking and subsampling grid, e.g. 512,128\n"
        "  -m, --merge_threshold   merging threshold\n"
        "  -e, --estimated_output  estimated output point amount factor\n"
        "  -a, --averaging-intra   intra-cell colour averaging\n"
        "  -i, --averaging-inter   inter-cell colour averaging\n"
        "  -w, --weight            distance weighting for inter-cell averaging\n"
        "  -r, --random            explicit random subsampling (default: implicit)\n"
        "  --device                HIP device (default 0)\n"
        "  --seed                  seed of the explicit random subsampling (default 0)\n";

[[noreturn]] void fail(const std::string& message)
{
    std::fprintf(stderr, "[error] %s\n", message.c_str());
    std::exit(-1);
}

std::vector<std::string> split(const std::string& s)
{
    std::vector<std::string> out;
    size_t a = 0;
    while (a <= s.size())
    {
        size_t b = s.find(',', a);
        if (b == std::string::npos)
            b = s.size();
        out.push_back(s.substr(a, b - a));
        a = b + 1;
    }
    return out;
}

Input parse(int argc, char** argv)
{
    Input in;
    bool haveFile = false, havePoints = false, haveType = false, haveData = false, haveGrids = false,
         haveOutput = false;
    for (int i = 1; i < argc; ++i)
    {
        const std::string a = argv[i];
        auto value          = [&]() -> std::string {
            if (i + 1 >= argc)
                fail("option " + a + " needs a value");
            return argv[++i];
        };
        if (a == "-h" || a == "--help")
        {
            std::printf("%s", kUsage);
            std::exit(0);
        }
        else if (a == "-f" || a == "--file")
            in.inputFile = value(), haveFile = true;
        else if (a == "-o" || a == "--output")
            in.outputPath = value(), haveOutput = true;
        else if (a == "-p" || a == "--points")
            in.pointAmount = static_cast<uint32_t>(std::stoul(value())), havePoints = true;
        else if (a == "-t" || a == "--type")
        {
            const std::string t = value();
            if (t != "float" && t != "double")
                fail(R"(Data type must be "float" or "double")");
            in.cloudType = t == "float" ? 0 : 1;
            haveType     = true;
        }
        else if (a == "-d" || a == "--data")
        {
            auto v = split(value());
            if (v.size() < 2)
                fail("Data info has to contain stride and scale");
            in.pointDataStride = static_cast<uint32_t>(std::stof(v[0]));
            in.scale           = std::stof(v[1]);
            haveData           = true;
        }
        else if (a == "-g" || a == "--grids")
        {
            auto v = split(value());
            if (v.size() < 2)
                fail("Grid sizes have to contain [chunking_grid, subsampling_grid]");
            in.chunkingGrid    = static_cast<uint32_t>(std::stoul(v[0]));
            in.subsamplingGrid = static_cast<uint32_t>(std::stoul(v[1]));
            haveGrids          = true;
        }
        else if (a == "-m" || a == "--merge_threshold")
            in.mergingThreshold = static_cast<uint32_t>(std::stoul(value()));
        else if (a == "-e" || a == "--estimated_output")
            in.outputFactor = std::stof(value());
        else if (a == "-a" || a == "--averaging-intra")
            in.isIntraCellAveraging = true;
        else if (a == "-i" || a == "--averaging-inter")
            in.isInterCellAveraging = true;
        else if (a == "-w" || a == "--weight")
            in.useWeightingFunction = true;
        else if (a == "-r" || a == "--random")
            in.performRandomSubsampling = true;
        else if (a == "--device")
            in.device = std::stoi(value());
        else if (a == "--seed")
            in.seed = std::stoull(value());
        else
            fail("unknown option " + a);
    }
    // src/argparser.h:42-112
    if (!haveFile)
        fail("No input file (-f) specified");
    if (!havePoints)
        fail("No point amount (-p) specified");
    if (!haveType)
        fail("No coordinate datatype (-t) specified");
    if (!haveData)
        fail("No data info (-d) specified");
    if (!haveGrids)
        fail("Grid sizes (-g) have to be specified: [chunking_grid, subsampling_grid]");
    if (!haveOutput)
        fail("No output path (-o) specified");
    if (in.isIntraCellAveraging && in.isInterCellAveraging)
        fail("Either intra-cell (-a) or inter-cell (-i) can be applied");
    if (!in.isInterCellAveraging && in.useWeightingFunction)
        fail("Weighting function can only be applied when performing inter-cell (-i) filtering");
    return in;
}

void printInput(const Input& in)  // src/argparser.h:114-131
{
    std::printf("[info] Input file: %s\n[info] Output path: %s\n[info] pointAmount: %u\n[info] cloudType: %u\n"
                "[info] dataStride: %u\n[info] scale: %g\n[info] chunkingGrid: %u\n[info] subsamplingGrid: %u\n"
                "[info] mergingThreshold: %u\n[info] outputFactor: %g\n[info] perform intra-cell averaging: %d\n"
                "[info] perform inter-cell averaging: %d\n[info] Use weighting function: %d\n"
                "[info] Subsampling method: %s\n",
                in.inputFile.c_str(), in.outputPath.c_str(), in.pointAmount, in.cloudType, in.pointDataStride,
                in.scale, in.chunkingGrid, in.subsamplingGrid, in.mergingThreshold, in.outputFactor,
                in.isIntraCellAveraging, in.isInterCellAveraging, in.useWeightingFunction,
                in.performRandomSubsampling ? "explicit random subsampling" : "implicit random subsampling");
}

void call(int rc, const char* what)
{
    if (rc != SUCCESS)
    {
        char message[1024];
        ocpi_get_last_error(message, sizeof(message));
        fail(std::string(what) + ": " + message);
    }
}

double seconds(std::chrono::steady_clock::time_point a)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
}

}  // namespace

int main(int argc, char** argv)
{
    Input in = parse(argc, argv);
    ocpi_set_logging_level(1);
    printInput(in);

    namespace fs = std::filesystem;
    if (!fs::exists(in.outputPath))
        fs::create_directories(in.outputPath);

    const uint64_t bytes = static_cast<uint64_t>(in.pointAmount) * in.pointDataStride;
    auto start           = std::chrono::steady_clock::now();
    std::ifstream file(in.inputFile, std::ios::binary | std::ios::ate);
    if (!file)
        fail("cannot open " + in.inputFile);
    if (static_cast<uint64_t>(file.tellg()) < bytes)
        fail("the file holds fewer than points x stride bytes");
    file.seekg(0);
    std::vector<uint8_t> host(bytes);
    file.read(reinterpret_cast<char*>(host.data()), static_cast<std::streamsize>(bytes));
    if (!file)
        fail("cannot read " + in.inputFile);

    if (hipSetDevice(in.device) != hipSuccess)
        fail("HIP device " + std::to_string(in.device) + " is not available");
    uint8_t* cloud = nullptr;
    if (hipMalloc(&cloud, bytes) != hipSuccess || hipMemcpy(cloud, host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        fail("cannot place the cloud in device memory");
    double realBB[6], cubicBB[6];
    call(ocpi_compute_bounding_box(cloud, in.pointAmount, in.pointDataStride, in.cloudType, in.device, realBB, cubicBB),
         "ocpi_compute_bounding_box");
    std::printf("[info] Reading cloud and calc bounding box took: %g [s]\n", seconds(start));

    void* session = nullptr;
    call(ocpi_create_session(&session, in.device), "ocpi_create_session");
    call(ocpi_set_cloud_type(session, in.cloudType), "ocpi_set_cloud_type");
    call(ocpi_set_cloud_point_amount(session, in.pointAmount), "ocpi_set_cloud_point_amount");
    call(ocpi_set_cloud_data_stride(session, in.pointDataStride), "ocpi_set_cloud_data_stride");
    call(ocpi_set_cloud_scale(session, in.scale, in.scale, in.scale), "ocpi_set_cloud_scale");
    call(ocpi_set_cloud_offset(session, cubicBB[0], cubicBB[1], cubicBB[2]), "ocpi_set_cloud_offset");
    call(ocpi_set_cloud_bb(session, cubicBB[0], cubicBB[1], cubicBB[2], cubicBB[3], cubicBB[4], cubicBB[5]),
         "ocpi_set_cloud_bb");
    call(ocpi_set_point_cloud_device(session, cloud), "ocpi_set_point_cloud_device");
    call(ocpi_configure_chunking(session, in.chunkingGrid, in.mergingThreshold, in.outputFactor),
         "ocpi_configure_chunking");
    call(ocpi_configure_subsampling(session, in.subsamplingGrid, in.isIntraCellAveraging, in.isInterCellAveraging,
                                    in.useReplacementScheme, in.performRandomSubsampling, in.useWeightingFunction),
         "ocpi_configure_subsampling");
    call(ocpi_set_random_seed(session, in.seed), "ocpi_set_random_seed");

    start = std::chrono::steady_clock::now();
    call(ocpi_init_octree(session), "ocpi_init_octree");
    call(ocpi_generate_octree(session), "ocpi_generate_octree");
    std::printf("[info] Generating the octree took: %g [s]\n", seconds(start));

    start = std::chrono::steady_clock::now();
    call(ocpi_export_potree(session, in.outputPath.c_str()), "ocpi_export_potree");
    call(ocpi_export_distribution_histogram(session, (in.outputPath + "/point_distribution.html").c_str(), 0),
         "ocpi_export_distribution_histogram");
    call(ocpi_export_json_report(session, (in.outputPath + "/statistics.json").c_str()), "ocpi_export_json_report");
    call(ocpi_export_memory_report(session, (in.outputPath + "/memory_report.html").c_str()),
         "ocpi_export_memory_report");
    std::printf("[info] Exporting took: %g [s]\n", seconds(start));
    call(ocpi_destroy_session(session), "ocpi_destroy_session");
    (void)hipFree(cloud);
    return 0;
}
