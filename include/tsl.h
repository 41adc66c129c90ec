/*
 * tsl.h -- C ABI of the MI355X (gfx950) MCL sensor update.
 *
 * This is the drop-in boundary for the hot path of uos/tsdf_localization:
 * TSDF map upload, per-particle scan evaluation, normalisation + pose
 * estimate, and systematic resampling.  Plain C types only: pointers, sizes,
 * POD structs whose layouts equal the reference's (static-asserted in the
 * implementation).  Every function returns TSL_OK or an error code and
 * leaves a thread-local message in tsl_last_error().
 *
 * Threading: a handle may be used by one thread at a time.  Different
 * handles (several maps, evaluators, filters) may be used concurrently; the
 * library keeps no global device state, unlike the reference's file-scope
 * __constant__/device globals (cuda_data.h:14-34).
 *
 * Streams: entry points taking `void* stream` enqueue asynchronously on that
 * hipStream_t (NULL = the null stream) and do not synchronise.
 */
#ifndef TSL_H
#define TSL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TSL_OK 0
#define TSL_ERR_INVALID_ARGUMENT 1
#define TSL_ERR_HIP 2               /* a HIP runtime call failed */
#define TSL_ERR_NO_VALID_PARTICLE 3 /* weight sum is zero: "No particle is valid!" */
#define TSL_ERR_OUT_OF_MEMORY 4
#define TSL_ERR_CAPACITY 5          /* an output buffer is too small */
#define TSL_ERR_UNSUPPORTED 6       /* input outside what this build handles */
#define TSL_ERR_ADMISSION 7         /* serving: request refused by the HBM budget */

/* ---- reference-compatible POD types ------------------------------------ */

/* Particle (include/tsdf_localization/particle.h:26-55): pose
 * (x, y, z, roll, pitch, yaw) and weight.  28 bytes. */
typedef struct tsl_particle {
    float first[6];
    float second;
} tsl_particle;

/* CudaPoint (include/tsdf_localization/cuda/cuda_evaluator.h:22-39).  12 bytes. */
typedef struct tsl_point {
    float x, y, z;
} tsl_point;

/* CudaSubVoxelMap<float,float>::MapCoef
 * (include/tsdf_localization/cuda/cuda_sub_voxel_map.h:22-50), same field
 * order and types, 120 bytes; a reference map's coef() can be memcpy'd here. */
typedef struct tsl_map_coef {
    uint64_t dim_x, dim_y, dim_z;
    float min_x, min_y, min_z;
    float max_x, max_y, max_z;
    float resolution;
    float init_value;
    uint64_t up_dim_x, up_dim_y, up_dim_z, up_dim_2;
    uint64_t sub_dim, sub_dim_2;
    uint64_t grid_occ_size, data_size;
} tsl_map_coef;

/* geometry_msgs::PoseWithCovariance as returned by Evaluator::evaluate
 * (cuda_evaluator.h:102): position, orientation quaternion (x, y, z, w),
 * row-major 6x6 covariance (the reference writes zeros on the diagonal). */
typedef struct tsl_pose {
    double position[3];
    double orientation[4];
    double covariance[36];
} tsl_pose;

typedef struct tsl_map tsl_map;
typedef struct tsl_evaluator tsl_evaluator;
typedef struct tsl_filter tsl_filter;

/* ---- library ----------------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char* tsl_last_error(void);

/* Library ABI version (major*10000 + minor*100 + patch). */
int tsl_version(void);

/* Number of visible HIP devices.  No counterpart in the reference. */
int tsl_device_count(int* count);

/* Select the device that handles created afterwards on this thread use. */
int tsl_set_device(int device);

/* ---- map: CudaSubVoxelMap + the upload in CudaEvaluator's constructor --- */

/* Constructor arithmetic of CudaSubVoxelMap
 * (include/tsdf_localization/cuda/cuda_sub_voxel_map.tcc:5-45): dims,
 * block dims and sub dims from the bounding box; data_size = 0. */
int tsl_map_coef_init(float min_x, float min_y, float min_z, float max_x, float max_y,
                      float max_z, float resolution, float init_value, tsl_map_coef* out);

/* Upload a map given in the reference's representation: block table
 * (grid_occ_size int32 offsets, -1 = empty block) and one float value per
 * voxel (data_size floats), i.e. rawGridOcc()/rawData() of a
 * CudaSubVoxelMap (cuda_sub_voxel_map.h:141,166).  Replaces the map upload of
 * CudaEvaluator::CudaEvaluator (src/cuda/cuda_evaluator.cu:21-59).  Voxel
 * values are stored on the device as 16-bit codes into a value palette
 * when the map holds at most 4096 distinct values, else as the float values
 * themselves (4 bytes per voxel); results are the same either way. */
int tsl_map_create(const tsl_map_coef* coef, const int32_t* grid_occ, const float* data,
                   tsl_map** out);

/* Upload a map given as TSDF distances in millimetres, one int16 per voxel
 * (INT16_MIN = voxel never set, reads init_value), plus the likelihood
 * parameters of createTSDFMap (include/tsdf_localization/map/map_util.h:17-154):
 * a voxel with |tsdf| < truncation_mm gets (N(d; sigma))^3 computed with the
 * reference's float/double arithmetic (map_util.h:68-73,124-126); others read
 * init_value.  coef->init_value is used as given. */
int tsl_map_create_tsdf(const tsl_map_coef* coef, const int32_t* grid_occ,
                        const int16_t* tsdf_mm, int truncation_mm, float sigma, tsl_map** out);

/* The 64-bit extension of tsl_map_create_tsdf, for maps of more voxels than
 * the reference's int block offsets can address (OCC_T = int,
 * cuda_sub_voxel_map.h:13; setData's int running offset, .tcc:202-211).
 * grid_occ: grid_occ_size int64 block offsets (-1 = empty block), each a
 * whole multiple of sub_dim^3 with its block inside the data_size voxels;
 * tsdf_mm: data_size int16 millimetres in the reference's linear layout
 * (offset + x + y*sub_dim + z*sub_dim^2).  sub_dim must be a multiple of 4 in
 * [4, 256].  Lookups report 64-bit reference indices and read the same
 * values as tsl_map_create_tsdf would.  The millimetres are converted on the
 * device a few hundred MB at a time. */
int tsl_map_create_tsdf_wide(const tsl_map_coef* coef, const int64_t* grid_occ, const int16_t* tsdf_mm,
                             int truncation_mm, float sigma, tsl_map** out);

/* Build the block table and voxel array from a list of cells exactly as
 * CudaSubVoxelMap::setData does (cuda_sub_voxel_map.tcc:171-231: cells are
 * placed through the float lookup of their corner coordinates).  coef is
 * updated (data_size).  grid_occ receives grid_occ_size entries; data
 * receives data_size floats and must have room for data_capacity floats
 * (TSL_ERR_CAPACITY, with coef->data_size set, if it does not).  Host only. */
int tsl_map_set_data(tsl_map_coef* coef, const float* cell_xyz, const float* cell_value,
                     size_t num_cells, int32_t* grid_occ, float* data, size_t data_capacity);

/* The cell list createTSDFMap builds from a map file's chunks
 * (include/tsdf_localization/map/map_util.h:17-154), for tsl_map_set_data.
 * chunk_pos holds (x, y, z) chunk coordinates per chunk (the /map dataset
 * name "x_y_z"), chunk_raw CHUNK_SIZE^3 = 262144 raw TSDFValue words per chunk
 * (x-major; low 16 bits the TSDF value in mm, high 16 bits the weight), in
 * the order the reference lists them (the group's name order).  coef receives
 * the map geometry (bounding box of the chunks and the origin, resolution
 * 0.064 m, init value from sigma).  A voxel with weight != 0 and |tsdf| < 600
 * mm becomes a cell (cell_xyz, cell_value = its likelihood); weight != 0 and
 * |tsdf| >= 600 mm a free-space point (free_xyz, the reference's free_map).
 * Sizes are reported through num_cells / num_free; TSL_ERR_CAPACITY if a
 * buffer is too small, so a first call with zero capacities sizes them.
 * The HDF5 read itself is the caller's.  Host only. */
int tsl_map_chunks_to_cells(const int32_t* chunk_pos, const uint32_t* chunk_raw, size_t num_chunks, float sigma,
                            tsl_map_coef* coef, float* cell_xyz, float* cell_value, size_t cell_capacity,
                            size_t* num_cells, float* free_xyz, size_t free_capacity, size_t* num_free);

/* Likelihood value of a valid TSDF cell and the map's init value
 * (map_util.h:68-73,124-126).  Host only. */
float tsl_likelihood_value(int16_t tsdf_mm, float sigma);
float tsl_likelihood_init(float sigma);

/* Device bytes held by the map. */
int tsl_map_device_bytes(const tsl_map* map, uint64_t* bytes);

int tsl_map_destroy(tsl_map* map);

/* ---- evaluator: CudaEvaluator / TSDFEvaluator -------------------------- */

/* CudaEvaluator(map, per_point, a_hit, a_range, a_max, max_range)
 * (include/tsdf_localization/cuda/cuda_evaluator.h:112) together with the
 * TSDFEvaluator constants (tsdf_evaluator.h:72-76).  per_point selects a
 * point-split tile shape; the scoring is the reference CPU evaluator's in
 * both cases (the reference's per-pair CUDA kernel tests the particle
 * position instead of the point range, cuda_eval_particles.h:499). */
int tsl_evaluator_create(tsl_map* map, int per_point, float a_hit, float a_range, float a_max,
                         float max_range, tsl_evaluator** out);

int tsl_evaluator_destroy(tsl_evaluator* ev);

/* sin/cos of the particle angles in the pose matrices
 * (tsdf_evaluator.cpp:102-145), the pose estimate's circular means
 * (:206-213) and the motion model's particle rotation
 * (particle_cloud.cpp:549-581).  The reference's CPU evaluator calls
 * unqualified sin(float)/cos(float), i.e. the C library's sinf/cosf; glibc
 * (2.28+) computes them by one algorithm built two ways on x86-64, with FMA
 * contraction (selected on CPUs with FMA3) and without.  The library
 * restates both, bit-exact with glibc on every float, on host and device:
 *   TSL_TRIG_GLIBC_FMA   glibc sinf/cosf as on x86-64 with FMA3 (default)
 *   TSL_TRIG_GLIBC_NOFMA glibc sinf/cosf as on x86-64 without FMA3
 *   TSL_TRIG_DOUBLE      (float)sin((double)x), correctly rounded but for rare ties
 * The reference's CUDA path uses CUDA's device sinf/cosf
 * (cuda_eval_particles.h:290-295), which has no counterpart here.  Applies
 * to every filter, batch and server of the evaluator created afterwards and
 * to its host-buffer entry points. */
#define TSL_TRIG_GLIBC_FMA 0
#define TSL_TRIG_GLIBC_NOFMA 1
#define TSL_TRIG_DOUBLE 2
int tsl_evaluator_set_trig(tsl_evaluator* ev, int mode);
int tsl_evaluator_get_trig(const tsl_evaluator* ev, int* mode);

/* sin and cos of n angles in a trig mode, on the device (on_device != 0) or
 * by the host build of the same code; for parity checking.  Synchronous. */
int tsl_trig(int mode, const float* x, size_t n, float* s, float* c, int on_device);

/* Host self-check: the host build of a trig mode against this process's C
 * library sinf/cosf over every non-NaN float whose bit pattern is a
 * multiple of stride.  No counterpart in the reference. */
int tsl_check_trig_libm(int mode, uint32_t stride, uint64_t* mismatches, uint64_t* checked);

/* Host self-check of the inverse trigonometry of the motion model's
 * getAngleFromMat (src/util/util.cpp:86-111): the host build of its asinf
 * and atanf against this process's C library over every non-NaN float
 * whose bit pattern is a multiple of stride, and of its atan2f on
 * atan2_samples seeded pseudo-random argument pairs.  No counterpart in
 * the reference. */
int tsl_check_invtrig_libm(uint32_t stride, uint64_t atan2_samples, uint64_t* mismatches, uint64_t* checked);

/* Evaluator::evaluate(particles, points, tf_matrix)
 * (cuda_evaluator.h:102; src/cuda/cuda_evaluator.cu:118-428;
 * CPU semantics src/evaluation/tsdf_evaluator.cpp:78-245).
 * particles (host, in/out): on return particles[i].second holds the
 * normalised weight.  points and tf_matrix (row-major 4x4, entries 0..11
 * used) are read only.  No points: *out_pose is zeroed and particles are
 * left untouched (cuda_evaluator.cu:122-125).  Zero weight sum:
 * TSL_ERR_NO_VALID_PARTICLE.  Synchronous. */
int tsl_evaluate(tsl_evaluator* ev, tsl_particle* particles, size_t num_particles,
                 const tsl_point* points, size_t num_points, const float tf_matrix[16],
                 tsl_pose* out_pose);

/* Evaluator::evaluate(particles, PointCloud2, tf)
 * (cuda_evaluator.h:97-100; src/cuda/cuda_evaluator.cu:78-116): the cloud's
 * x, y, z and ring fields as arrays, reduced as tsl_filter_set_scan_cloud
 * describes, then evaluated as tsl_evaluate.  Synchronous. */
int tsl_evaluate_cloud(tsl_evaluator* ev, tsl_particle* particles, size_t num_particles, const tsl_point* raw,
                       const int32_t* ring, size_t num_raw, const float tf_matrix[16], tsl_pose* out_pose);

/* Raw (unnormalised) per-particle weights of one evaluation, written to
 * weights_out[num_particles]; particles are not modified.  Synchronous. */
int tsl_evaluate_raw(tsl_evaluator* ev, const tsl_particle* particles, size_t num_particles,
                     const tsl_point* points, size_t num_points, const float tf_matrix[16],
                     float* weights_out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tsl_evaluate_raw on device-resident inputs.
 * particles (num_particles tsl_particle) and points (num_points tsl_point)
 * are device pointers read in place; weights_out (num_particles floats,
 * device) receives the raw weights, bit-identical to tsl_evaluate_raw's.
 * Every intermediate lives in a fixed 64 MiB device array per GPU, which
 * the library holds until the process ends; TSL_ERR_CAPACITY if the call
 * needs more (about 52 B per particle, 60 B per point, plus 4 B per particle
 * per scan chunk).  Enqueued on `stream` with no allocation, no host copy
 * and no synchronisation.  Calls that overlap on one GPU share the array
 * and race: serialise them (one stream, or events). */
int tsl_evaluate_raw_noalloc(tsl_evaluator* ev, const tsl_particle* particles, size_t num_particles,
                             const tsl_point* points, size_t num_points, const float tf_matrix[16],
                             float* weights_out, void* stream);

/* Voxel index that CudaSubVoxelMap::getIndex (cuda_sub_voxel_map.tcc:51-137)
 * returns for every (particle, point) pair, row-major [particle][point]
 * (coef.data_size = rejected).  Runs the production lookup on the device; for
 * parity checking.  out has num_particles*num_points entries.  Synchronous. */
int tsl_trace_indices(tsl_evaluator* ev, const tsl_particle* particles, size_t num_particles,
                      const tsl_point* points, size_t num_points, const float tf_matrix[16],
                      uint64_t* out);

/* Voxel value CudaSubVoxelMap::getEntry (cuda_sub_voxel_map.tcc:140-158)
 * returns for every (particle, point) pair, read through the device map's
 * storage by the production kernel; for parity checking.  Synchronous. */
int tsl_trace_values(tsl_evaluator* ev, const tsl_particle* particles, size_t num_particles,
                     const tsl_point* points, size_t num_points, const float tf_matrix[16],
                     float* out);

/* Self-checks of the host-built lookup tables, host only (no GPU): the
 * sub-voxel threshold table of a resolution against IEEE division over every
 * float in [0, 1) whose bit pattern is a multiple of stride, and the
 * constant-divisor multipliers of the brick layout over every x < 2^31 that is
 * a multiple of stride.  No counterpart in the reference. */
int tsl_check_sub_table(float resolution, uint32_t stride, uint64_t* mismatches, uint32_t* entries);
int tsl_check_magic_divisor(uint32_t divisor, uint32_t stride, uint64_t* mismatches);

/* Pose matrices L[12] (tsdf_evaluator.cpp:102-145) computed on the device
 * with the sin/cos of trig mode `trig` (TSL_TRIG_*). */
int tsl_pose_matrices(const tsl_particle* particles, size_t num_particles,
                      const float tf_matrix[16], int trig, float* out);

/* ---- resampling: SystematicResampler ----------------------------------- */

/* SystematicResampler::resample
 * (include/tsdf_localization/resampling/novel_resampling.h:38-73) on the
 * device: in[M] are particles with normalised weights in .second, u0 the
 * uniform draw from [0, 1/M).  Writes the emitted particles to out (room
 * for out_capacity) and their source indices to out_indices (may be NULL),
 * and the number emitted (which, as in the reference, can differ from M) to
 * *out_count.  Indices are bit-exact with the reference's double running
 * sum and float U recurrence.  Synchronous. */
int tsl_resample_systematic(const tsl_particle* in, size_t M, float u0, tsl_particle* out,
                            size_t out_capacity, int32_t* out_indices, size_t* out_count);

/* Upper bound on the number of particles tsl_resample_systematic can emit
 * for M particles and draw u0 (for sizing out). */
int tsl_resample_capacity(size_t M, float u0, size_t* capacity);

/* ---- device-resident filter (particle set and scan live in HBM) -------- */

/* A filter owns a device-resident particle set (capacity particles) bound
 * to an evaluator; the update path below never copies particles to the
 * host.  No direct counterpart in the reference, whose particles live in
 * ParticleCloud on the host (particle_cloud.h:46-259). */
int tsl_filter_create(tsl_evaluator* ev, size_t capacity, tsl_filter** out);
int tsl_filter_destroy(tsl_filter* f);

/* Copy particles in / out (synchronous). */
int tsl_filter_set_particles(tsl_filter* f, const tsl_particle* particles, size_t n);
int tsl_filter_get_particles(tsl_filter* f, tsl_particle* particles, size_t capacity, size_t* n);

/* Upload a scan and its scanner->base transform (synchronous). */
int tsl_filter_set_scan(tsl_filter* f, const tsl_point* points, size_t n,
                        const float tf_matrix[16]);

/* The scan reduction of TSDFEvaluator::evaluateParticles
 * (src/evaluation/tsdf_evaluator.cpp:304-376) on the device, then as
 * tsl_filter_set_scan: raw points with their ring numbers (the PointCloud2
 * fields x, y, z and ring); points closer than 1 m are dropped, the rest are
 * deduplicated by (ring, centre of their `cell`-sized voxel) keeping the
 * first, and ordered by ring then index.  The reference's ring iterator
 * advances only on kept points (:319-331); that is reproduced: the k-th
 * kept point takes ring[k].  TSL_ERR_INVALID_ARGUMENT if a kept point's ring
 * is outside [0, 64). */
int tsl_filter_set_scan_raw(tsl_filter* f, const tsl_point* raw, const int16_t* ring, size_t n, float cell,
                            const float tf_matrix[16]);

/* The point reduction of CudaEvaluator::evaluate(particles, PointCloud2,
 * tf) (src/cuda/cuda_evaluator.cu:78-116) on the device, then as
 * tsl_filter_set_scan: every raw point becomes the centre of its 0.064 m
 * voxel, floor(v / 0.064) * 0.064 + 0.032 evaluated in double and rounded
 * to float, and each distinct centre is kept once.  The reference orders
 * the centres by its hash set's iteration (an unordered_set, unspecified);
 * here they are ordered by the ring (int32, the reference reads the field
 * as int), then the index, of each centre's lowest-index point.  The order
 * changes only the float summation order of the weights. */
int tsl_filter_set_scan_cloud(tsl_filter* f, const tsl_point* raw, const int32_t* ring, size_t n,
                              const float tf_matrix[16]);

/* The filter's scan points in the order they were given (after the
 * reduction, for tsl_filter_set_scan_raw / _cloud). */
int tsl_filter_get_scan(tsl_filter* f, tsl_point* out, size_t capacity, size_t* n);

/* Sensor update on `stream`: pose matrices, evaluation, normalisation and
 * pose sums; particles' .second become the normalised weights.  Async. */
int tsl_filter_update_async(tsl_filter* f, void* stream);

/* Systematic resampling of the resident set with draw u0 on `stream`; the
 * set is replaced by the emitted particles.  Async. */
int tsl_filter_resample_async(tsl_filter* f, float u0, void* stream);

/* Wait for `stream`, then read the pose estimate of the last update and the
 * current particle count.  Returns TSL_ERR_NO_VALID_PARTICLE if the last
 * update had a zero weight sum. */
int tsl_filter_sync(tsl_filter* f, void* stream, tsl_pose* pose, size_t* count);

/* Keep a device copy of the current particle set (synchronous), and restore
 * the resident set from it on `stream` (async).  Lets a caller replay an
 * update from the same particles, e.g. to time a fixed workload.  No
 * counterpart in the reference. */
int tsl_filter_snapshot(tsl_filter* f);
int tsl_filter_restore_async(tsl_filter* f, void* stream);

/* Tile shape of the evaluation: k particles per wavefront (1 or 2; 0 =
 * automatic), points_per_lane scan points each lane handles per iteration
 * (1, 2 or 4; 0 = automatic; independent gathers in flight) and chunk_len
 * scan points per wavefront (0 = automatic).  Results depend on the tile
 * only through chunk_len (float summation order); k and points_per_lane
 * leave them unchanged.  No counterpart in the reference. */
int tsl_filter_set_tile(tsl_filter* f, int k, int points_per_lane, uint32_t chunk_len);

/* Order in which the particles are evaluated: 0 = caller order, 1 =
 * spatial order (sorted by position and heading, so particles evaluated
 * together share cached voxels), -1 = automatic (spatial from 8192
 * particles).  Results are identical in every order.  No counterpart in the
 * reference. */
int tsl_filter_set_order(tsl_filter* f, int mode);

/* How the evaluation finds a block's voxel offset: 0 = the block table in
 * device memory (one gather), 1 = an occupancy bitmap with running counts
 * staged in LDS (no gather; needs offsets in setData order and a table of
 * at most 131072 blocks), -1 = automatic.  Results are identical.  No
 * counterpart in the reference. */
int tsl_filter_set_block_lookup(tsl_filter* f, int mode);

/* Time the candidate evaluation shapes (particle order x block lookup x
 * tile) on the filter's current particles and scan (reps launches each, on
 * `stream`, synchronously) and keep the fastest; reports it through the
 * non-NULL out arguments (best_ms: evaluation kernel plus the per-update
 * cost of the particle order).  No counterpart in the reference (its launch
 * shape is fixed, src/cuda/cuda_evaluator.cu:301). */
int tsl_filter_autotune(tsl_filter* f, void* stream, int reps, int* best_order, int* best_k,
                        int* best_points_per_lane, uint32_t* best_chunk, float* best_ms);

/* Bracket every evaluation-kernel launch of this filter with HIP events on
 * the launch stream (enable resets the totals); read the accumulated kernel
 * time, number of launches and particle*point evaluations launched. */
int tsl_filter_set_timing(tsl_filter* f, int enable);
int tsl_filter_eval_timing(tsl_filter* f, double* ms, uint64_t* launches, uint64_t* pairs);

/* ---- motion update ------------------------------------------------------ */

/* The six normal distributions of a motion step (dx, dy, dz, roll, pitch,
 * yaw): means and standard deviations as the reference builds them, and the
 * seed of the noise stream.  Particle i's draw k is a pure function of
 * (seed, i, k) (SplitMix64 + Box-Muller), so a filter's particles move the
 * same way alone or in a batch; the reference draws from a sequential
 * std::mt19937 instead (resampler.h:23), which no device can reproduce. */
typedef struct tsl_motion {
    double mean[6];
    double stddev[6];
    uint64_t seed;
} tsl_motion;

/* Distributions of ParticleCloud::motion_update(dx, dy, dz, roll, pitch,
 * yaw, lin_base_noise, ang_base_noise) (src/particle_cloud.cpp:456-497) with
 * the motion parameters a_1..a_12 (particle_cloud.h:60-71). */
int tsl_motion_from_delta(float dx, float dy, float dz, float roll, float pitch, float yaw, float lin_base_noise,
                          float ang_base_noise, const float a[12], uint64_t seed, tsl_motion* out);

/* Distributions of ParticleCloud::motionUpdate(const nav_msgs::Odometry&)
 * (src/particle_cloud.cpp:153-207): forward velocity and yaw rate over
 * time_diff seconds. */
int tsl_motion_from_odom(float linear_velocity, float angular_velocity, float time_diff, const float a[12],
                         uint64_t seed, tsl_motion* out);

/* ParticleCloud::apply_model (src/particle_cloud.cpp:499-617) on the
 * filter's device-resident particles. */
int tsl_filter_motion_async(tsl_filter* f, const tsl_motion* motion, void* stream);

/* ---- batched update: many filters, one launch sequence ----------------- */

/* A batch updates and resamples n device-resident filters that share one
 * evaluator (map and sensor model) with one fixed sequence of launches: a
 * persistent evaluation kernel walks every member's work items, and every
 * per-particle kernel covers all members at once.  Each member gives the
 * same bits as when it is updated alone with the same tile (chunk length).
 * Members keep their own particles (any count) and scans; after changing a
 * member's particles, scan or snapshot outside the batch, call
 * tsl_batch_prepare.  No counterpart in the reference (one filter per
 * process, src/mcl_3d.cpp:378-425). */
typedef struct tsl_batch tsl_batch;

int tsl_batch_create(tsl_evaluator* ev, tsl_filter* const* filters, size_t n, tsl_batch** out);
int tsl_batch_destroy(tsl_batch* b);

/* Tile of the batched evaluation (k particles per wavefront, 1 or 2; points
 * per lane 1, 2 or 4; chunk_len points per chunk, 0 = automatic: 2048, or
 * the whole scan if shorter).  The workgroups of the persistent grid claim
 * work items (4k particles x one chunk) one at a time.  Re-prepares the
 * batch. */
int tsl_batch_set_tile(tsl_batch* b, int k, int points_per_lane, uint32_t chunk_len);

/* Time the batched update (every member, reps launches per candidate on
 * `stream`, synchronously) for every tile k in {1, 2} x points per lane in
 * {1, 2, 4} x chunk_len in {automatic, 1024, 4096, 8192} and keep the
 * fastest (tsl_batch_set_tile); reports it through the non-NULL out
 * arguments.  Leaves every member as one update would (weights
 * normalised, particles and counts unchanged).  No counterpart in the
 * reference (one filter per process, fixed launch shape,
 * src/cuda/cuda_evaluator.cu:301). */
int tsl_batch_autotune(tsl_batch* b, void* stream, int reps, int* best_k, int* best_points_per_lane,
                       uint32_t* best_chunk, float* best_ms);

/* Rebuild the members' device descriptors and reserve their buffers. */
int tsl_batch_prepare(tsl_batch* b);

/* Restore every member from its snapshot (tsl_filter_snapshot). */
int tsl_batch_restore_async(tsl_batch* b, void* stream);

/* Sensor update (evaluate + normalise + pose sums) of every member. */
int tsl_batch_update_async(tsl_batch* b, void* stream);

/* Systematic resampling of every member, member i with draw u0[i]; the
 * plan is built on the device from each member's device-side count. */
int tsl_batch_resample_async(tsl_batch* b, const float* u0, void* stream);

/* Motion update of every member, member i with motions[i]. */
int tsl_batch_motion_async(tsl_batch* b, const tsl_motion* motions, void* stream);

/* [restore +] update + resample of every member. */
int tsl_batch_step_async(tsl_batch* b, const float* u0, int restore, void* stream);

/* Capture one step ([restore +] update + resample, draws read from device
 * memory) into a HIP graph on `stream` (a created stream, not NULL);
 * tsl_batch_launch replays it with new draws u0[n]. */
int tsl_batch_capture(tsl_batch* b, int restore, void* stream);
int tsl_batch_launch(tsl_batch* b, const float* u0, void* stream);

/* Wait for `stream`, then report every member's pose, particle count and
 * status (0 ok, 1 no valid particle, 2 resampling capacity, 3 empty, 4 draw
 * does not advance); returns the first member error, if any. */
int tsl_batch_sync(tsl_batch* b, void* stream, tsl_pose* poses, uint64_t* counts, int* status);

/* ---- serving: resident robots, continuous batching --------------------- */

/* A server keeps many robots' particle sets resident on the device and
 * runs their requests (motion step + scan + resampling draw) with
 * continuous batching: a worker thread starts a tick with the pending
 * requests of every robot not in flight (at most max_batch, oldest first;
 * one request per robot per tick), completes earlier ticks as they finish
 * and admits requests that arrive meanwhile into the next tick.  Robots are
 * admitted against an HBM budget (TSL_ERR_ADMISSION).  A robot's result is
 * the same alone or in any mix.  No counterpart in the reference (one robot
 * per ROS node, src/mcl_3d.cpp). */
typedef struct tsl_server tsl_server;

typedef struct tsl_request {
    uint64_t id;              /* caller's tag, returned with the result */
    uint32_t robot;           /* from tsl_server_add_robot */
    uint32_t pad;
    tsl_motion motion;        /* odometry step (applied before the sensor update) */
    const tsl_point* points;  /* reduced scan, copied by tsl_server_submit */
    size_t num_points;
    float u0;                 /* systematic-resampling draw in [0, 1/M) */
    float pad2;
} tsl_request;

typedef struct tsl_result {
    uint64_t id;
    uint32_t robot;
    int32_t status;           /* 0 ok, 1 no valid particle, 2 resampling capacity, 3 empty, 4 draw does not advance */
    tsl_pose pose;            /* weighted mean pose of the sensor update */
    uint64_t count;           /* particles after resampling */
    double submit_ms, start_ms, done_ms;  /* server clock: submitted, tick started, result ready */
    uint32_t batch;           /* requests in the tick */
    uint32_t pad;
} tsl_result;

/* hbm_budget_bytes: device bytes the map and the resident robots may hold
 * (0 = no limit; the tick buffers -- three ticks' scan staging of max_batch
 * scans of the largest admitted robot -- are not charged); max_robots
 * resident at once; max_particles per robot;
 * max_batch requests per tick; chunk_len scan points per evaluation chunk
 * of robots with fewer than 65,536 particle slots (0 = each request's whole
 * scan); robots with more (the bulk class) are always evaluated a whole
 * scan per chunk.  A request's result depends on its own chunking only,
 * never on the tick it ran in. */
int tsl_server_create(tsl_evaluator* ev, uint64_t hbm_budget_bytes, uint32_t max_robots, uint64_t max_particles,
                      uint32_t max_batch, uint32_t chunk_len, tsl_server** out);
int tsl_server_destroy(tsl_server* s);

/* Admit a robot with its particle set, the largest scan it will send and its
 * scanner->base transform; every device buffer its ticks need is reserved
 * here. */
int tsl_server_add_robot(tsl_server* s, const tsl_particle* particles, size_t n, uint32_t max_points,
                         const float tf_matrix[16], uint32_t* robot);
int tsl_server_remove_robot(tsl_server* s, uint32_t robot);

/* Thread-safe; the scan is copied before returning. */
int tsl_server_submit(tsl_server* s, const tsl_request* request);

/* Up to capacity finished results (waiting up to wait_ms for the first). */
int tsl_server_poll(tsl_server* s, tsl_result* out, size_t capacity, size_t* n, int wait_ms);

/* A robot's particles (waits for its request in flight; call with none
 * pending for it). */
int tsl_server_get_particles(tsl_server* s, uint32_t robot, tsl_particle* out, size_t capacity, size_t* n);

int tsl_server_stats(tsl_server* s, uint64_t* used_bytes, uint64_t* ticks, uint64_t* requests);

/* Robots of 65,536 particle slots and more (global localisation) are served
 * on a stream of their own, apart from the others (every server of a device
 * shares the two streams), and their evaluation is cut into launches of
 * about `pairs` particle-point pairs on a grid that leaves reserved_per_cu
 * workgroup slots per compute unit free, so other requests' kernels start
 * within about one launch (default 2^27 pairs, 2 slots; pairs = 0: one
 * launch per tick on the whole grid).  Results are the same bits either way.
 * Takes effect for ticks started afterwards. */
int tsl_server_set_slice(tsl_server* s, uint64_t pairs, uint32_t reserved_per_cu);

/* Evaluation tile of the bulk class (bulk != 0) or of the others: k
 * particles per wavefront (1, 2) and points per lane (1, 2, 4); default 1
 * and 2 for the bulk class, 2 and 1 for the others.  Changes only the speed
 * (a member's chunking does not depend on it).  Takes effect for ticks
 * started afterwards. */
int tsl_server_set_tile(tsl_server* s, int bulk, int k, int points_per_lane);

/* ---- .mcl snapshot files ------------------------------------------------ */

/* MCLFile::write (src/util/mcl_file.cpp:14-64): points, their rings,
 * particles, tf[16] and pose (x, y, z, q1..q4) as text with the default
 * stream formatting (6 significant digits, so values round-trip only to
 * that precision, as in the reference).  Host only. */
int tsl_mcl_write(const char* path, const tsl_point* points, const int32_t* rings, size_t num_points,
                  const tsl_particle* particles, size_t num_particles, const float tf[16], const float pose[7]);

/* MCLFile::read (src/util/mcl_file.cpp:66-113).  Reports the sizes through
 * num_points / num_particles; TSL_ERR_CAPACITY (sizes still reported) if the
 * buffers are too small, so a first call with zero capacities sizes them.
 * tf and pose may be NULL.  Host only. */
int tsl_mcl_read(const char* path, tsl_point* points, int32_t* rings, size_t points_capacity, tsl_particle* particles,
                 size_t particles_capacity, float tf[16], float pose[7], size_t* num_points, size_t* num_particles);

/* Random-line gather probe (measurement, no counterpart in the reference):
 * wavefronts whose 64 lanes each read 2 bytes from a different pseudo-random
 * 128-byte line of a device buffer of buffer_bytes (rounded down to a power
 * of two of lines), 64 loads per lane, 8.4M lanes per launch, reps timed
 * launches.  Reports lines returned per second and the launch time: the
 * ceiling the voxel-code gathers of the evaluation work against (a buffer
 * the size of a map's code array is served by the L2 and Infinity Cache, a
 * many-GB one by HBM).  Allocates its buffer; synchronous. */
int tsl_probe_gather(uint64_t buffer_bytes, uint32_t reps, double* lines_per_s, double* ms_per_launch);

/* reps plain device-to-device copies of a bytes-sized buffer (hipMemcpyAsync):
 * the known-size reference for scaling memory-traffic counters.  No
 * counterpart in the reference.  Allocates; synchronous. */
int tsl_probe_copy(uint64_t bytes, uint32_t reps, double* ms_per_copy);

/* ---- streams and timing (plumbing for benchmarks) ---------------------- */
int tsl_stream_create(void** stream);
int tsl_stream_destroy(void* stream);
int tsl_stream_synchronize(void* stream);
int tsl_event_create(void** event);
int tsl_event_destroy(void* event);
int tsl_event_record(void* event, void* stream);
int tsl_event_elapsed_ms(void* start, void* stop, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* TSL_H */
