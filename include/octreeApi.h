/*
 * octreeApi.h - C ABI of the MI355X LOD octree builder (libOctreeApi.so).
 *
 * Drop-in for the reference's session API (OctreeApi/include/octreeApi.h:42-72): the same 18
 * functions with the same signatures, argument meaning and return codes (SUCCESS 0 /
 * UNEXPECTED_ERROR -1, octreeApi.h:37-38). The entry points after the first block are additive;
 * each says what it is for.
 *
 * Behavioural differences from the reference, all on paths where the reference crashes, exits
 * or is undefined:
 *   - no call ever terminates the process: HIP failures, invalid configurations and call-order
 *     violations return -1; ocpi_get_last_error() gives the reason;
 *   - the configuration is validated in ocpi_init_octree (chunking grid a power of two in
 *     [2, 512], intra-cell + weighted rejected, non-empty cloud, E = floor(N * factor) >= N);
 *   - ocpi_generate_octree returns -1 when the subsampled points do not fit the estimated output
 *     (N + I > E), where the reference writes past its buffers;
 *   - ocpi_generate_octree may be called again on the same session (a rebuild of the same cloud);
 *   - every session owns its device and stream, so separate sessions may be used from separate
 *     threads.
 */
#ifndef OCTREE_API_H
#define OCTREE_API_H

#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
#define EXPORTED extern "C" __attribute__((visibility("default")))
#else
#define EXPORTED __attribute__((visibility("default")))
#endif

#define SUCCESS          (0)
#define UNEXPECTED_ERROR (-1)

/* ---- the reference's API (OctreeApi/include/octreeApi.h:42-72) -------------------------------- */

/* octreeApi.h:42 - creates a session bound to HIP device `device` (session.cpp:27-46). */
EXPORTED int ocpi_create_session (void** session, int device);
/* octreeApi.h:43 - frees the session and every device buffer it owns. */
EXPORTED int ocpi_destroy_session (void* session);
/* octreeApi.h:44 - 1 info, 2 warn, 3 error, anything else debug (octreeApi.cpp:23-42). */
EXPORTED int ocpi_set_logging_level (int level);

/* octreeApi.h:46 - 0 = 3 x float32 + 3 x uint8, 1 = 3 x float64 + 3 x uint8 (metadata.cuh:25-29). */
EXPORTED int ocpi_set_cloud_type (void* session, uint8_t cloudType);
/* octreeApi.h:47 */
EXPORTED int ocpi_set_cloud_point_amount (void* session, uint32_t pointAmount);
/* octreeApi.h:48 - record stride in bytes; the colour follows the three coordinates. */
EXPORTED int ocpi_set_cloud_data_stride (void* session, uint32_t dataStride);
/* octreeApi.h:49 - quantisation scale of the exported int32 positions. */
EXPORTED int ocpi_set_cloud_scale (void* session, double x, double y, double z);
/* octreeApi.h:50 - stored; like the reference it does not change the output. */
EXPORTED int ocpi_set_cloud_offset (void* session, double x, double y, double z);
/* octreeApi.h:51-52 - the cubic bounding box (equal sides, max = min + side). */
EXPORTED int ocpi_set_cloud_bb (
        void* session, double minX, double minY, double minZ, double maxX, double maxY, double maxZ);

/* octreeApi.h:54 - borrowed host buffer of pointAmount * stride bytes; copied to HBM by
 * ocpi_init_octree, so it must stay valid until that call returns. */
EXPORTED int ocpi_set_point_cloud_host (void* session, uint8_t* pointCloud);

/* octreeApi.h:56-57 */
EXPORTED int ocpi_configure_chunking (
        void* session, uint32_t chunkingGrid, uint32_t mergingThreshold, float outputFactor);
/* octreeApi.h:58-65 - useRandomSubsampling selects explicit random subsampling, otherwise the
 * implicit (first point) method; replacementScheme only sets a metadata flag. */
EXPORTED int ocpi_configure_subsampling (
        void* session,
        uint32_t subsamplingGrid,
        bool intraCellAveraging,
        bool interCellAveraging,
        bool replacementScheme,
        bool useRandomSubsampling,
        bool useWeightingFunction);

/* octreeApi.h:67 - records the file name; the report is written when the session is destroyed. */
EXPORTED int ocpi_export_memory_report (void* session, const char* filename);
/* octreeApi.h:68 - leaf point-count histogram as HTML; binWidth 0 picks Scott's rule. */
EXPORTED int ocpi_export_distribution_histogram (void* session, const char* filename, uint32_t binWidth);
/* octreeApi.h:69 - statistics.json. */
EXPORTED int ocpi_export_json_report (void* session, const char* filename);
/* octreeApi.h:70 - writes hierarchy.bin, octree.bin and metadata.json into an existing directory. */
EXPORTED int ocpi_export_potree (void* session, const char* directory);
/* octreeApi.h:71 - validates the configuration, allocates HBM and uploads the cloud. */
EXPORTED int ocpi_init_octree (void* session);
/* octreeApi.h:72 - counting, merging, distribution and subsampling on the GPU. */
EXPORTED int ocpi_generate_octree (void* session);

/* ---- additive entry points ------------------------------------------------------------------- */

/* Borrowed device buffer (HBM) of pointAmount * stride bytes, used in place by init and generate.
 * The reference has this cloud memory type (CloudMemory ClOUD_DEVICE, metadata.cuh:34-38,
 * point_cloud.cu:24-44) but no setter for it. */
EXPORTED int ocpi_set_point_cloud_device (void* session, uint8_t* devicePointCloud);

/* Seed of the explicit random subsampling. The reference seeds with std::time
 * (octree_processor_impl.cu:82); here the selection is a reproducible function of the seed. */
EXPORTED int ocpi_set_random_seed (void* session, uint64_t seed);

/* Message of the last failing call on this thread ("" if none). Returns the message length. */
EXPORTED int ocpi_get_last_error (char* buffer, uint32_t bufferSize);

/* Sizes of the last build. summary receives up to n values:
 * [0] sparse nodes M, [1] points N, [2] output capacity E, [3] used output records N + I,
 * [4] internal nodes, [5] leaf nodes, [6] depth, [7] dense nodes D, [8] exported (hierarchy) nodes */
EXPORTED int ocpi_get_build_summary (void* session, uint64_t* summary, uint32_t n);

/* Copies the M sparse nodes of the last build (ids in canonical order: level ascending from the
 * finest, then linear index). flags: bit0 finished, bit1 internal. children: 8 ints per node,
 * slot = 4 * dx + 2 * dy + dz. Any pointer may be NULL. */
EXPORTED int ocpi_copy_nodes (
        void* session,
        uint32_t* denseKey,
        uint32_t* denseCount,
        uint32_t* pointCount,
        uint64_t* dataIdx,
        int32_t* parent,
        int32_t* children,
        uint8_t* flags);

/* Copies the first `count` entries of the point LUT (source point index per output record). */
EXPORTED int ocpi_copy_point_lut (void* session, uint32_t* destination, uint64_t count);

/* Copies the first `count` 18-byte output records (int32 x, y, z; uint16 r, g, b). */
EXPORTED int ocpi_copy_output (void* session, uint8_t* destination, uint64_t count);

/* Device-side durations of the last ocpi_generate_octree, in milliseconds, measured with HIP
 * events on the session stream: [0] counting kernel, [1] merge (pyramid, nodes, leaf blocks),
 * [2] distribution kernel, [3] subsampling (all levels and the final placement), [4] whole
 * build, [5] summed subsampling insert kernels, [6] summed subsampling emit kernels. */
EXPORTED int ocpi_get_stage_timings (void* session, float* milliseconds, uint32_t n);

/* ---- build service: continuous batching of many builds on one GPU ---------------------------
 * None of these has a counterpart in the reference, which builds one cloud per process
 * (src/main.cpp:55-109). A service owns `slots` workers, each with its own HIP stream and a
 * device arena that stays resident between requests. A configured session (setters and
 * configure_* as for ocpi_init_octree) is a request: ocpi_service_submit queues it, a free slot
 * takes it when the projected device memory of all slots stays within memoryBudget bytes
 * (footprint estimated from N and the output factor), builds overlap on the device, and the
 * finished octree is copied to host memory and attached to the session, whose exports and
 * copies then work as after ocpi_generate_octree. The session and its cloud buffer must stay
 * untouched until ocpi_service_wait has returned for it. */
EXPORTED int ocpi_create_service (void** service, int device, uint32_t slots, uint64_t memoryBudget);
/* A service that builds in batches (continuous batching over the batched pipeline): `workers`
 * workers, each with a batch builder; a free worker takes every queued request that fits the
 * memory budget next to the other workers' arenas, up to maxBatch, and builds them together (one
 * launch per stage for all of them); requests arriving meanwhile wait for the next batch, which
 * another worker may build while the first batch is still on the GPU. Submit, wait, stats and
 * destroy work as for ocpi_create_service; set_active_slots sets the active workers. */
EXPORTED int ocpi_create_batch_service (
        void** service, int device, uint32_t workers, uint32_t maxBatch, uint64_t memoryBudget);
/* Fails requests still queued and waits for the running ones. */
EXPORTED int ocpi_destroy_service (void* service);
/* Validates the session's configuration and queues it (-1 if invalid or above the budget). */
EXPORTED int ocpi_service_submit (void* service, void* session);
/* Waits up to timeoutMs for the session's build: 0 done, -1 failed or timed out (see last error). */
EXPORTED int ocpi_service_wait (void* service, void* session, uint32_t timeoutMs);
/* Number of slots that take requests (1..slots). */
EXPORTED int ocpi_service_set_active_slots (void* service, uint32_t slots);
/* stats: [0] submitted [1] completed [2] failed [3] queued now [4] running now [5] max queue depth
 * [6] p50 latency ms [7] p99 latency ms [8] mean latency ms [9] reserved device bytes
 * [10] memory budget [11] active slots [12] slots [13] mean queue wait ms [14] max latency ms
 * [15] peak reserved device bytes [16] [17] [18] mean time per completed request in setup (arena
 * sizing and cloud upload), generate and result capture, ms (batch service: [17] is the time of
 * the request's batch) [19] batches built [20] mean requests per batch. Latency = submit to result
 * attached to the session. */
EXPORTED int ocpi_service_get_stats (void* service, double* stats, uint32_t n);
EXPORTED int ocpi_service_reset_stats (void* service);
/* Of the session's last service request: [0] queue wait ms [1] build + copy ms [2] total ms. */
EXPORTED int ocpi_get_request_timing (void* session, double* milliseconds, uint32_t n);

/* Device bytes a build of the session's configured cloud may need (the estimate the service admits
 * requests by; E = N * outputFactor drives most of it). Validates the configuration. */
EXPORTED int ocpi_estimate_footprint (void* session, uint64_t* bytes);

/* Bounding box of a cloud already in HBM, reduced on the GPU: realBB = {min x, y, z, max x, y, z}
 * of the coordinates (widened to double), cubicBB = the cube the reference CLI builds from it
 * (src/boundingbox.h:10-52, replacing its host loop). No counterpart in the library API; it
 * serves callers such as the CLI that would otherwise scan the cloud on the host. */
EXPORTED int ocpi_compute_bounding_box (
        const uint8_t* devicePointCloud,
        uint32_t pointAmount,
        uint32_t dataStride,
        uint8_t cloudType,
        int device,
        double* realBB,
        double* cubicBB);

/* ---- batched builds: many clouds, one launch per stage ---------------------------------------
 * No counterpart in the reference, which builds one cloud per process and launches every kernel
 * per phase and per node (kernel_executor.cuh:8-28). A batch builder takes configured sessions
 * (setters and configure_* as for ocpi_init_octree; host or device clouds) and builds them
 * together: the sessions that share record type, chunking grid, colour filter and subsampling
 * method go through every stage in one launch per stage, with per-cloud offsets into a shared
 * device arena. Each session's result is attached as after ocpi_generate_octree (node arrays,
 * data offsets and record sets equal a build of that session alone). The builder's arena stays
 * resident between batches. */
EXPORTED int ocpi_create_batch_builder (void** builder, int device);
EXPORTED int ocpi_destroy_batch_builder (void* builder);
/* Builds sessions[0 .. count). status (may be NULL) receives 0 or -1 per session; the call returns
 * -1 when any session failed (the first failure is the last error), and the others keep their
 * results. The sessions' host clouds must stay valid until the call returns. */
EXPORTED int ocpi_batch_build (void* builder, void** sessions, uint32_t count, int32_t* status);
/* Persistent workgroups per launch of the batched kernels: those with one item per thread, one
 * subsampling tile per workgroup and one sort chunk per workgroup (each 1..65536). */
EXPORTED int ocpi_batch_set_launch (void* builder, uint32_t threadGrid, uint32_t tileGrid, uint32_t sortGrid);
/* 1 (the default): each batch's two launch segments (before and after the host reads the node
 * counts) run as HIP graphs, captured once per launch shape and replayed; 0: plain launches. */
EXPORTED int ocpi_batch_use_graphs (void* builder, uint32_t on);
/* Of the last ocpi_batch_build, summed over its groups: device ms of [0] counting [1] merge
 * [2] distribution [3] subsampling and placement [4] the whole pipeline; host ms of [5] the whole
 * call [6] the result copies. With graphs, [0] covers segment 1 (counting and the count pyramid),
 * [3] segment 2 (node links to placement) and [1], [2] are 0. */
EXPORTED int ocpi_batch_get_timings (void* builder, float* milliseconds, uint32_t n);
/* Device bytes a batch of these sessions needs (validates them). */
EXPORTED int ocpi_estimate_batch_footprint (void** sessions, uint32_t count, uint64_t* bytes);

/* ocpi_compute_bounding_box without any device allocation. No counterpart in the reference; it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same arguments, results and errors as ocpi_compute_bounding_box; the
 * partial minima and maxima live in a static 49,200-byte device array (one per GPU, held until the
 * process ends), which bounds the reduction to 1024 workgroups and so fits any point amount. It
 * runs on the null stream and returns after the result is copied. Calls on one GPU share that
 * array: within a process they are serialised by a lock; calls from separate processes on one GPU
 * each have their own copy of it. */
EXPORTED int ocpi_compute_bounding_box_noalloc (
        const uint8_t* devicePointCloud,
        uint32_t pointAmount,
        uint32_t dataStride,
        uint8_t cloudType,
        int device,
        double* realBB,
        double* cubicBB);

#endif
