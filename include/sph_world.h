/*
 * sph_world.h - C-ABI of the MI355X (gfx950) SPH substep library, libsph_hip.so.
 *
 * Drop-in for the per-substep compute chain that the reference records in
 * recordCommands (src/renderer/commands.c:193-402) over the vk_context particle buffers
 * (src/renderer/vk_types.h:127-159).  The reference has no FFI; its boundary is that
 * dispatch chain plus the UBO physics fields.  Each entry point below names the reference
 * code it replaces.
 *
 * Contract:
 *   - Every int-returning call returns SPH_OK (0) or a negative sph_status; the reference
 *     prints and exit(1)s instead (src/renderer/buffer.c:30-47).  sph_last_error() gives
 *     the message of the calling thread's last failure.
 *   - The library owns device memory; callers own host arrays.
 *   - Calls on one handle are stream-ordered and must come from one thread at a time.
 *   - Particle arrays are the reference's vec4 AoS layout (x, y, z, w), 16 B per particle
 *     (predictPositions.comp:27-30).  After a step they are in the reference's permuted
 *     (cell-key-sorted) order, exactly as the reference's buffers are.
 */
#ifndef SPH_WORLD_H
#define SPH_WORLD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sph_status {
    SPH_OK = 0,
    SPH_ERR_INVALID_ARG = -1,  /* bad size, NULL pointer, NaN or out-of-box particle      */
    SPH_ERR_OUT_OF_MEMORY = -2,
    SPH_ERR_DEVICE = -3,       /* a HIP call failed                                       */
    SPH_ERR_STATE = -4,        /* call not valid in the handle's current state            */
    SPH_ERR_CAPACITY = -5      /* request does not fit the configured budget / maximum    */
} sph_status;

/* Physics parameters: the UBO fields the step shaders read
 * (UniformBufferObject, src/renderer/vk_types.h:52-65, filled by updateUniforms,
 * src/main.c:53-63), plus the table size and substep count the host code fixes
 * (renderer.c:741-754 sizes the table to numParticles; main.c:16 NUM_SUBSTEPS). */
typedef struct sph_params {
    float particle_radius;          /* wall margin of the predict pass; +0.5 in integrate */
    float smoothing_radius;         /* R: cell edge and kernel support                     */
    float g;                        /* gravity magnitude, applied along -z                 */
    float min_xyz[3];               /* box                                                 */
    float max_xyz[3];
    float dt;                       /* per-substep dt (main.c:57: frame dt / substeps)     */
    float particle_mass;
    float pressure_multiplier;
    float target_density;
    float viscosity_strength;       /* accepted and ignored, as the reference's shader
                                       never reads it (applyPressure...comp:174,313)       */
    float volume;                   /* density normaliser 2*pi*R^5/15 (main.c:349)         */
    float near_pressure_multiplier;
    uint32_t table_cells;           /* T, cell-table size; keys are hash & (T-1).
                                       0 means T = particle count, the reference's choice. */
    uint32_t substeps;              /* substeps per sph_world_step frame                   */
} sph_params;

/* Fills dt and volume the way the reference's host does: dt = (float)(frame_dt / substeps)
 * (main.c:57) and volume = (float)(2*pi*R^5/15) evaluated in double from the double
 * radius literal (main.c:349).  Also sets p->substeps. */
void sph_params_derive(sph_params *p, double frame_dt, uint32_t substeps,
                       double smoothing_radius);

typedef struct sph_world sph_world;

/* Replaces initVulkan's buffer creation (renderer.c:851-888) for the particle, cell and
 * table buffers of one scene of n particles on the current HIP device. */
int sph_world_create(const sph_params *params, uint32_t n, sph_world **out);

/* Replaces updateUniforms (main.c:37-82) for the physics fields.  n and table_cells are
 * fixed at creation; a different table_cells returns SPH_ERR_INVALID_ARG, and so does a box
 * or smoothing radius that changes the number of cells, bricks or buckets of the cell grid
 * (one that keeps them, e.g. two box extents swapped, is accepted). */
int sph_world_set_params(sph_world *w, const sph_params *params);

/* Replaces uploadDataToBufferViaStaging of the position and velocity buffers
 * (buffer.c:102, renderer.c:569-645).  Host arrays of n vec4.  Positions must be finite
 * and inside the box (the hash of an outside position is undefined in the reference). */
int sph_world_upload_state(sph_world *w, const float *pos_vec4, const float *vel_vec4);

/* Enqueues frames * substeps substeps; replaces the substep loop of recordCommands
 * (commands.c:193-402).  Asynchronous: returns once enqueued. */
int sph_world_step(sph_world *w, uint32_t frames);

/* Waits for all enqueued work of the handle. */
int sph_world_sync(sph_world *w);

/* Copies the state back (synchronising first).  Any pointer may be NULL.  color_vec4
 * receives updateVelocities.comp's speed ramp (:120-141) computed from the velocities. */
int sph_world_download_state(sph_world *w, float *pos_vec4, float *vel_vec4, float *color_vec4);

/* Overlapped readback, for a host that draws from host memory: the reference draws from the
 * device buffers it steps (commands.c:521-525) and has no counterpart.  Copies the state as
 * of this call into a device-side snapshot on the world's stream, then into pos_vec4 /
 * vel_vec4 (either may be NULL) on a second stream, and returns without waiting, so that
 * sph_world_step can run meanwhile.  The host arrays must stay valid and untouched until
 * sph_world_snapshot_wait returns; pinned arrays (hipHostMalloc) let the copy overlap the
 * step.  A snapshot call first waits for the previous one.  Holds 32 B per particle of
 * device memory from the first call on. */
int sph_world_snapshot_async(sph_world *w, float *pos_vec4, float *vel_vec4);
int sph_world_snapshot_wait(sph_world *w);

/* Device-resident state for a renderer that draws from device memory, as the reference
 * draws its particles straight from the buffers it steps (commands.c:521-525, shader.vert),
 * e.g. Vulkan buffers imported into HIP with hipImportExternalMemory.
 *   sph_world_bind_state: the world steps the caller's device buffers (n vec4 each,
 *     16-byte aligned, device memory) in place instead of its own; upload / download /
 *     snapshot then read and write them too.  The caller keeps ownership and must keep
 *     them alive while bound.  Both NULL returns to the world's own buffers (whose contents
 *     are then stale: upload a state).  Binding waits for the world's queued work.
 *   sph_world_device_state: the device pointers the world currently steps (its own or the
 *     bound ones); their contents are the state after sph_world_sync. */
int sph_world_bind_state(sph_world *w, float *pos_vec4_dev, float *vel_vec4_dev);
int sph_world_device_state(const sph_world *w, float **pos_vec4_dev, float **vel_vec4_dev);

/* Allocation-free step.  No counterpart in the reference: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Enqueues frames * params->substeps substeps of one scene on `stream` (a
 * hipStream_t; NULL = the null stream) over the caller's device state, n vec4 positions and
 * velocities (16-byte aligned device memory), updated in place into the reference's sorted
 * order, exactly as a sph_world created with these parameters, given this state and stepped
 * by sph_world_step would leave them (bit for bit).  Returns once enqueued.
 * It calls no hipMalloc / hipFree: every intermediate buffer lives in one static device
 * array of at most 72 MiB that each GPU the library runs on holds until the process ends,
 * which bounds n <= SPH_NOALLOC_MAX_PARTICLES and the table (table_cells, or n when 0) <=
 * SPH_NOALLOC_MAX_TABLE_CELLS; larger scenes return SPH_ERR_CAPACITY.  All calls on one GPU
 * share that array: calls that overlap on one GPU (other streams, other threads) race and
 * must be serialised by the caller.  Positions must be finite and inside the box, as for
 * sph_world_upload_state (not checked here: the state is never copied to the host). */
#define SPH_NOALLOC_MAX_PARTICLES 262144u
#define SPH_NOALLOC_MAX_TABLE_CELLS 524288u
int sph_step_noalloc(const sph_params *params, uint32_t n, float *pos_vec4_dev, float *vel_vec4_dev,
                     uint32_t frames, void *stream);

/* Intermediate buffers of the most recent substep, in the reference's formats, for parity
 * checks.  Stages other than CELL_KEYS, SORTED_*, START, END, PRED_POS and DENSITY need
 * debug capture (sph_world_set_debug) to have been on during that substep. */
typedef enum sph_dump_stage {
    SPH_DUMP_CELL_KEYS = 0,    /* u32[n]  CellEntry.key in the substep's input order       */
    SPH_DUMP_SORTED_KEYS = 1,  /* u32[n]  sorted CellEntry.key                             */
    SPH_DUMP_SORTED_INDEX = 2, /* u32[n]  sorted CellEntry.particleIndex (input index)     */
    SPH_DUMP_START = 3,        /* i32[T]  startIndices, -1 = empty                         */
    SPH_DUMP_END = 4,          /* i32[T]  endIndices, -1 = empty                           */
    SPH_DUMP_PRED_POS = 5,     /* f32[4n] predictedPositions, sorted order                 */
    SPH_DUMP_PRED_VEL = 6,     /* f32[4n] predictedVelocities, sorted order (debug)        */
    SPH_DUMP_DENSITY = 7,      /* f32[n]  densities                                        */
    SPH_DUMP_NEAR_DENSITY = 8, /* f32[n]  nearDensities                                    */
    SPH_DUMP_VEL_FORCE = 9     /* f32[4n] velocities after the force pass (debug)          */
} sph_dump_stage;

/* Neighbour-pass kernels.  All variants give bit-identical results.
 *   DIRECT: one lane per particle, runs read from global memory (the reference's scheme).
 *   BRICK:  workgroups own 4x4x4-cell bricks drained from a device work queue and read the
 *           6x6x6 halo's runs from LDS.
 *   AUTO:   the autotuner's choice (the default): the first substeps a world of a new
 *           size bucket (particle count, table size, brick count rounded to powers of two)
 *           steps run each configuration in turn (the brick kernels, and the per-lane
 *           kernels with 2 or 4 candidates per scan step, 64-, 128- or 256-lane workgroups,
 *           with or without force neighbour lists recorded by the density pass), two
 *           substeps each with the second timed, then the radix digit widths that need
 *           fewer passes; every configuration gives bit-identical results, so these are the
 *           world's real substeps (no state is saved, nothing is allocated).  The fastest
 *           is then kept and cached for the process; choosing it waits once for the
 *           substeps already queued.  sph_world_autotune measures on a saved copy instead.
 *           SPH_AUTOTUNE=0 in the environment disables the measurement (DIRECT is used);
 *           SPH_LANE_CONFIG=0..9 instead pins one per-lane configuration (for profiling:
 *           candidates per scan step x workgroup lanes 0 = 2x256, 1 = 4x256, 2 = 2x128,
 *           3 = 4x128, 4 = 2x64 with force neighbour lists, 5..9 the same without). */
typedef enum sph_kernel_variant {
    SPH_KERNEL_AUTO = 0,
    SPH_KERNEL_DIRECT = 1,
    SPH_KERNEL_BRICK = 2,
    SPH_KERNEL_TILE = 3
} sph_kernel_variant;
int sph_world_set_kernel(sph_world *w, int variant);
/* Runs the measurement now (`substeps` timed substeps per variant, state restored) and
 * caches the result; *chosen receives the variant, *ms_per_substep (may be NULL) its time. */
int sph_world_autotune(sph_world *w, uint32_t substeps, int *chosen, double *ms_per_substep);
/* The variant the next step will use (DIRECT or BRICK), or < 0 on error. */
int sph_world_kernel(const sph_world *w);
/* Human-readable launch configuration the next step will use (the autotuner also chooses
 * the per-lane kernels' candidates per scan step and workgroup size). */
int sph_world_kernel_config(const sph_world *w, char *buf, size_t bytes);
/* Selects DIRECT with one fixed per-lane launch configuration, 0 <= config <
 * sph_lane_configs() (the numbering of SPH_LANE_CONFIG); for reproducible timings and for
 * testing every configuration.  All configurations give bit-identical results. */
int sph_world_set_lane_config(sph_world *w, int config);
int sph_lane_configs(void);
/* Selects TILE with one fixed tile configuration, 0 <= config < sph_tile_configs() (the
 * numbering of SPH_TILE_CONFIG): 0 = 256-lane workgroups staging up to 1280 entries, 1 =
 * 128 lanes and 768 entries, both followed by the per-lane force pass over the neighbour
 * lists the tiled density pass records; 2 and 3 = 0 and 1 with the force pass tiled too;
 * 4 = configuration 0 with the approximate square root and reciprocal (needs
 * SPH_PRECISION_TOLERANCE).  The tiled kernels stage each tile's cells and their
 * neighbours in LDS and give bit-identical results to every other variant; the tile shape
 * follows the particles per cell of the last uploaded state.  AUTO measures 0 and 1 (and 4
 * in the tolerance mode); 2 and 3 run only when selected here.  Needs a box of at most
 * 64n + 65536 cells.  No counterpart in the reference (tuning). */
int sph_world_set_tile_config(sph_world *w, int config);
int sph_tile_configs(void);
/* What the last substep's tiled passes did (all 0 when it ran other kernels): out[0] tiles,
 * out[1] strays (particles walked one lane each through global memory: those no tile owns
 * and those of handed-over tiles), out[2] tiles whose region streamed through LDS in more
 * than one group of rows (it exceeded the LDS capacity), out[3] tiles handed over to the
 * strays (a single row of the region too large, more particles than the workgroup takes,
 * or, with the tiled force pass, a region larger than the LDS capacity).  Synchronises the
 * world's stream.  No counterpart in the reference (diagnostics). */
int sph_world_tile_stats(sph_world *w, uint32_t out[4]);

/* Precision of the neighbour passes.  EXACT (the default): every float operation is the
 * reference's, correctly rounded in its order, and results are bit-identical to the FP32
 * oracle.  TOLERANCE: the world may also use (and its autotuner may pick) launch
 * configurations 10 and 11, which evaluate the pair terms with the hardware's approximate
 * square root and reciprocal; cell keys, the sorted pairs and the table stay exact, and one
 * substep's densities, forces, positions and velocities agree with the reference's
 * equations within 1e-5 relative (the single-step accuracy the drop-in must keep).  Selecting a precision
 * re-tunes an AUTO world; configurations 10 and 11 need TOLERANCE. */
typedef enum sph_precision { SPH_PRECISION_EXACT = 0, SPH_PRECISION_TOLERANCE = 1 } sph_precision;
int sph_world_set_precision(sph_world *w, int mode);
/* Pins the radix sort's digit width (8, 11 or 12 bits; the autotuner otherwise chooses it
 * per size bucket among the widths that need fewer passes than 8-bit digits).  A world
 * holds look-back memory for the widths whose per-tile words stay within 16 B per particle
 * (or 8 MiB); others return SPH_ERR_CAPACITY.  All widths give the same stable order, the
 * reference's 4 x 8-bit sort (commands.c:106-161) included. */
int sph_world_set_radix_bits(sph_world *w, int bits);
/* Force neighbour lists (recorded by the density pass for the force pass; no reference
 * counterpart: the reference walks 27 cells in both passes).  A world starts with 64
 * entries per particle, an uploaded state raises that to five times the neighbours of a
 * uniform fill of its bounding box (a dense start such as the reference's own scene gets
 * up to 254 at once), and the lists grow to hold the longest that did not fit (twice the
 * capacity, or that list plus a quarter, at most 254): sph_world_step looks between frames
 * without waiting (a frame or two after the overflow it synchronises once and grows), and
 * sph_world_sync / download_state and the tuning look as well.  Eight looks in a row with
 * no list above half the capacity halve it (down to 64), at sph_world_sync /
 * download_state only, where the stream is idle.  A particle whose list does not fit walks
 * its cells in the force pass instead: results are the same bits at every capacity, only
 * speed and memory (4 B per entry per particle) change.
 *   sph_world_set_list_capacity: entries 1..254 fixes the capacity (reallocated now);
 *     0 returns to the adaptive sizing above.
 *   sph_world_list_capacity: the entries per particle allocated now.
 *   sph_world_device_bytes: device memory the world holds now.  sph_world_footprint
 *     bounds it at the largest list capacity. */
int sph_world_set_list_capacity(sph_world *w, uint32_t entries);
int sph_world_list_capacity(const sph_world *w);
uint64_t sph_world_device_bytes(const sph_world *w);

int sph_world_set_debug(sph_world *w, int enable);
int sph_world_debug_dump(sph_world *w, int stage, void *dst, size_t bytes);

/* Per-pass device time in ms averaged over `substeps` eagerly launched substeps, the
 * hipEvent counterpart of the reference's timestamp table (commands.c:3-18,
 * main.c:170-198).  The substeps advance the state like sph_world_step. */
typedef struct sph_pass_times {
    double hash_ms;      /* updateCellEntries (+ table clear, histogram)                   */
    double sort_ms;      /* multiRadixSort (all passes)                                    */
    double reorder_ms;   /* predictPositions + scatterPass + updateStartIndices            */
    double density_ms;   /* calculateDensities                                             */
    double force_ms;     /* applyPressureAndViscosityForce + updateVelocities              */
    double total_ms;
} sph_pass_times;
int sph_world_pass_times(sph_world *w, uint32_t substeps, sph_pass_times *out);

/* Density map for the volumetric renderer: replaces updateDensityMap.comp:114-148 and its
 * dispatch after the substeps (commands.c:418-424; image created by initializeDensityMap,
 * renderer.c:774-804).  For each voxel of edge voxel_size (DENSITY_CELL_SIZE, main.c:14)
 * over the box: density / volume at the voxel centre, summed over the runs of the 27 cells
 * around it in the positions and cell table the last substep left (so at least one
 * substep must have run since the last upload; SPH_ERR_STATE otherwise).  Voxels are
 * stored x fastest, dims[0] * dims[1] * dims[2] floats; dims are the image size,
 * ceil((max - min) * (1 / voxel_size)) per axis.  The cell keys use the table size
 * (table_cells - 1 as mask), which equals the shader's numParticles - 1 when T = N.  A
 * voxel the shader would not store (the corner one, when its two voxel-count formulas
 * differ) is written as 0. */
int sph_world_density_map_dims(const sph_world *w, float voxel_size, uint32_t dims[3]);
/* The same voxel counts for any scene's parameters (e.g. one scene of a batch). */
int sph_density_map_dims(const sph_params *params, float voxel_size, uint32_t dims[3]);
/* Into host memory; synchronous. */
int sph_world_density_map(sph_world *w, float voxel_size, float *out, size_t count);
/* Into device memory (e.g. memory a renderer imported); enqueued on the world's stream,
 * complete after sph_world_sync. */
int sph_world_density_map_device(sph_world *w, float voxel_size, float *dev_out, size_t count);

/* Device footprint in bytes of a world of n particles and table_cells cells (0 = n), with
 * its force lists at their largest capacity (254 entries); sph_world_device_bytes gives
 * what a world holds at its current capacity. */
uint64_t sph_world_footprint(uint32_t n, uint32_t table_cells);

void sph_world_destroy(sph_world *w);

const char *sph_last_error(void);

/* ---------------------------------------------------------------------------------------
 * Batched step: many independent scenes (own parameters, box, table size) stepped by one
 * chain of launches and one hipGraph per substep.  No reference counterpart: the reference
 * steps one scene per frame (commands.c:193-402); each scene here follows exactly those
 * semantics and its results are bit-identical to a sph_world of that scene alone.
 * Every scene of a batch must use the same substep count.  At most 65535 scenes. */
typedef struct sph_batch sph_batch;
int sph_batch_create(uint32_t scene_count, const sph_params *params, const uint32_t *counts,
                     sph_batch **out);
int sph_batch_set_kernel(sph_batch *b, int variant);  /* sph_kernel_variant; AUTO picks per
                                                         scene from its particle density */
int sph_batch_upload_scene(sph_batch *b, uint32_t scene, const float *pos_vec4,
                           const float *vel_vec4);
int sph_batch_step(sph_batch *b, uint32_t frames);
int sph_batch_sync(sph_batch *b);
int sph_batch_download_scene(sph_batch *b, uint32_t scene, float *pos_vec4, float *vel_vec4,
                             float *color_vec4);
int sph_batch_set_debug(sph_batch *b, int enable);
int sph_batch_debug_dump(sph_batch *b, uint32_t scene, int stage, void *dst, size_t bytes);
int sph_batch_scene_uses_bricks(sph_batch *b, uint32_t scene);  /* 1, 0, or < 0 on error */
/* TILE at one fixed exact tile configuration (0 to 3, see sph_world_set_tile_config); every
 * scene whose box has at most 64 cells per particle (+ 65536) gets tiles sized for its last
 * uploaded state, the others' particles take the tiled launches' stray kernels.  AUTO
 * measures tile configurations 0 and 1 beside the per-lane and brick kernels.  No
 * counterpart in the reference (tuning). */
int sph_batch_set_tile_config(sph_batch *b, int config);
/* The kernels the next step uses once tuned (DIRECT, BRICK or TILE; AUTO before its
 * measurement), or < 0 on error. */
int sph_batch_kernel(const sph_batch *b);
int sph_batch_pass_times(sph_batch *b, uint32_t substeps, sph_pass_times *out);
/* sph_world_density_map for one scene of the batch (its own box and table range). */
int sph_batch_density_map(sph_batch *b, uint32_t scene, float voxel_size, float *out, size_t count);
/* Force-list capacity of the batch (one for all its scenes), as sph_world_set_list_capacity /
 * sph_world_list_capacity: sized from each uploaded scene and grown during sph_batch_step,
 * at sph_batch_sync and at download_scene unless fixed. */
int sph_batch_set_list_capacity(sph_batch *b, uint32_t entries);
int sph_batch_list_capacity(const sph_batch *b);
uint32_t sph_batch_particles(const sph_batch *b);
void sph_batch_destroy(sph_batch *b);

/* ---------------------------------------------------------------------------------------
 * Serving: a request loop over the batched step.  No reference counterpart (the reference
 * runs one interactive scene).  Requests carry their parameters, initial state and step
 * count; a worker thread admits them in arrival order while the sum of resident footprints
 * (136 B + 4 B per force-list entry per particle: 648 B at the default 128 entries; 24 B per
 * table cell, 8 B per halo-grid cell) stays within hbm_budget_bytes and the capacities
 * hold, keeps their state resident in device memory,
 * steps every resident scene together one frame at a time, and retires finished scenes
 * with their state copied back.
 * Requests arriving while others run are admitted at the next frame boundary (continuous
 * batching); each result is bit-identical to stepping that scene alone.  Every request
 * must use the server's substep count.  Calls are thread-safe.  Resident scenes run on the
 * one-lane-per-particle kernels; SPH_SERVER_BRICKS=1 in the environment selects the brick
 * kernels instead. */
typedef struct sph_server sph_server;
typedef struct sph_server_config {
    uint64_t hbm_budget_bytes;  /* admission budget for resident scenes                   */
    uint32_t max_scenes;        /* resident scenes at once (<= 65535)                     */
    uint32_t max_particles;     /* resident particles at once (buffers sized for it)      */
    uint32_t max_table_cells;   /* resident table cells at once                           */
    uint32_t max_bricks;        /* resident bricks at once; 0 = 2^22                      */
    uint32_t substeps;          /* substeps per frame, shared by all requests             */
    uint32_t list_entries;      /* force-list entries per particle (1..254); 0 = 128       */
} sph_server_config;
typedef enum sph_request_state {
    SPH_REQ_QUEUED = 0, SPH_REQ_RUNNING = 1, SPH_REQ_DONE = 2, SPH_REQ_FAILED = 3
} sph_request_state;
typedef struct sph_server_stats {
    uint64_t submitted, completed, failed, frames;
    uint32_t resident_scenes, queued;
    uint64_t resident_particles, resident_bytes;
    double step_p50_ms, step_p99_ms;    /* wall time of one batched frame               */
    double scene_p50_ms, scene_p99_ms;  /* submit -> result ready                       */
    double wait_p50_ms, wait_p99_ms;    /* submit -> admitted                           */
    double particle_steps_per_s;        /* resident particles x frames / stepping time  */
} sph_server_stats;
int sph_server_create(const sph_server_config *cfg, sph_server **out);
/* Validates and copies the request; invalid requests (NaN or out-of-box particles, n = 0,
 * bad parameters) are refused here and never reach the batch. */
int sph_server_submit(sph_server *s, const sph_params *params, uint32_t n,
                      const float *pos_vec4, const float *vel_vec4, uint32_t steps,
                      uint64_t *request_id);
int sph_server_poll(sph_server *s, uint64_t request_id);  /* sph_request_state or < 0 */
int sph_server_wait(sph_server *s, uint64_t request_id, double timeout_s);  /* < 0: forever */
/* Copies a finished request's state (vec4 AoS, sorted order) and forgets the request. */
int sph_server_result(sph_server *s, uint64_t request_id, float *pos_vec4, float *vel_vec4);
int sph_server_get_stats(sph_server *s, sph_server_stats *out);
void sph_server_destroy(sph_server *s);

/* Seeded scene initialisers (host only).  Counterparts of initializePositionsBuffer /
 * initializeVelocitiesBuffer (renderer.c:569-645) for synthetic workloads: PCG32 streams,
 * w = 0 like the reference's initial buffers.
 *   block:   uniform in [lo, hi] per axis;  vel ~ N(0, vel_sigma) per component. */
int sph_scene_fill_block(uint32_t n, uint64_t seed, const float lo[3], const float hi[3],
                         float vel_sigma, float *pos_vec4, float *vel_vec4);

/* The reference's own start-up scene.  sph_params_reference_default fills the compile-time
 * constants of main.c:11-30 and the box of main.c:356-358: (-140,-40,-40)..(140,40,40),
 * particle radius 0.3, R 2.5, g 100, frame dt 0.01 in 1 substep, mass 1, pressure
 * multiplier 15000, target density 1, viscosity 0, near pressure 0, T = N.
 *
 * sph_scene_fill_reference replaces initializePositionsBuffer and initializeVelocitiesBuffer
 * (renderer.c:568-645): n positions uniform in the block centred in x/y/z with 25 / 45 /
 * 45 % of the box's extents (inset by the particle radius), evaluated with the reference's
 * float and double rounding, four libc rand() draws per particle (randf(),
 * renderer.c:506-508), and velocity (0, 0, -100).  `seed` is the srand() argument (0 in
 * main.c:351).  `rng` names the C library rand() of the reference build:
 * SPH_RNG_GLIBC (Linux; RAND_MAX 2^31 - 1) or SPH_RNG_MSVCRT (the Windows/MinGW build of
 * build.bat; RAND_MAX 32767).  The reference's 'R' key (main.c:85-98) refills the buffers
 * without re-seeding, so the k-th reset's scene is `resets` = k (the start-up scene is 0).
 * The caller's own rand() state is not touched. */
typedef enum sph_rng { SPH_RNG_GLIBC = 0, SPH_RNG_MSVCRT = 1 } sph_rng;
void sph_params_reference_default(sph_params *p);
int sph_scene_fill_reference(const sph_params *params, uint32_t n, uint32_t seed, int rng,
                             uint32_t resets, float *pos_vec4, float *vel_vec4);

#ifdef __cplusplus
}
#endif
#endif
